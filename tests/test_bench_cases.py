"""bench.py's barrier measurements: the cases it launches (host logic only)."""
import bench


def test_barrier_cases_leave_out_the_unbounded_and_the_unresolved():
    cases = bench.barrier_cases()
    assert (2, "xcd", 1024) in cases and (0, "flags", 256) in cases and (1, "sense", 256) in cases
    # the two-level barrier at 256-thread workgroups once did not complete: never launched
    assert all(not (var == 2 and wg == 256) for var, _, wg in cases)
    # HIP's cooperative grid sync has no spin budget: only on request
    assert all(var != 3 for var, _, _ in cases)
    coop = bench.barrier_cases(coop=True)
    assert {(3, 1024), (3, 256)} <= {(var, wg) for var, _, wg in coop}
    assert all(not (var == 2 and wg == 256) for var, _, wg in coop)
