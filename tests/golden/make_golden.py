"""Writes tests/golden/fixtures.json: the generator fixtures (n, m and FNV-1a
64 digests of row_ptr / col / weights / priorities for each configured
workload and a few small ones) and the oracle's answers on the small ones.

The reference ships no graph inputs or solutions (they are external
downloads), so these vectors come from this repository's own seeded
generators and CPU oracle; the oracle itself is pinned by
tests/test_oracle.py. Regenerate only on a deliberate change of a generator
or of the oracle:

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from gpu_workgroup_managment_amd import graphs as G  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json")


def digest(g):
    d = {"n": g.n, "m": g.m, "row_ptr": oracle.fnv1a64(g.row_ptr), "col": oracle.fnv1a64(g.col)}
    if g.wt is not None:
        d["wt"] = oracle.fnv1a64(g.wt)
    if g.prio is not None:
        d["prio"] = oracle.fnv1a64(g.prio)
    return d


def small_graphs():
    """The small workloads whose oracle answers are stored."""
    return {
        "er_small": G.add_weights(G.erdos_renyi(1024, 4096, G.SEED_ER), 1, 255, G.SEED_WEIGHTS),
        "rmat_small": G.add_weights(G.rmat(10, 16, seed=G.SEED_RMAT), 1, 255, G.SEED_WEIGHTS),
        "rmat_trunc": G.rmat(11, 16, seed=7, truncate_n=1500),
        "grid_small": G.add_priorities(G.add_weights(G.grid(32, 48), 1, 255, G.SEED_WEIGHTS), G.SEED_PRIO),
    }


def config_graphs():
    """The full configured workloads (digests only)."""
    return {
        "er_config": G.er_config(),
        "rmat_config": G.rmat_config(),
        "grid_config": G.grid_config(),
    }


def answers(name, g):
    a = {}
    lv, levels = oracle.bfs(g.row_ptr, g.col, 0)
    a["bfs"] = {"digest": oracle.fnv1a64(lv), "levels": int(levels), "reached": int((lv != oracle.INF).sum())}
    if g.wt is not None:
        d = oracle.sssp(g.row_ptr, g.col, g.wt, 0)
        a["sssp"] = {"digest": oracle.fnv1a64(d), "max": int(d[d != oracle.INF].max())}
    if g.prio is not None:
        c, rounds = oracle.color(g.row_ptr, g.col, g.prio)
        a["color"] = {"digest": oracle.fnv1a64(c), "rounds": int(rounds), "colors": int(c.max())}
        s, rounds = oracle.mis(g.row_ptr, g.col, g.prio)
        a["mis"] = {"digest": oracle.fnv1a64(s), "rounds": int(rounds), "size": int((s == 2).sum())}
        pr = oracle.pagerank_f32(g.row_ptr, g.col, np.float32(0.85), 50)
        a["pagerank_f32"] = {"digest": oracle.fnv1a64(pr)}
    return a


def matmult_answer(dim=256, seed=1234):
    """The reference's matmult self-check: A then B from one middle-square
    stream, C = A.B in int32, djb2 hash of C's diagonal."""
    a, s = oracle.middle_square_fill(dim, dim, seed)
    b, _ = oracle.middle_square_fill(dim, dim, s)
    return {"dim": dim, "seed": seed, "hash": int(oracle.hash_diag(oracle.matmul_i32(a, b)))}


def batch_digest(count=512, seed=G.SEED_BATCH):
    items = G.batch_spec(count, seed)
    raw = bytes(memoryview(items).cast("B"))
    return {"count": count, "seed": seed, "items": oracle.fnv1a64(raw), "sum_n": int(sum(it.n for it in items))}


def build():
    fx = {"splitmix64_seed_1234567": [int(x) for x in G.splitmix64(1234567, 5)]}
    fx["graphs"] = {k: digest(g) for k, g in {**small_graphs(), **config_graphs()}.items()}
    fx["answers"] = {k: answers(k, g) for k, g in small_graphs().items()}
    fx["matmult"] = matmult_answer()
    fx["batch"] = batch_digest()
    return fx


if __name__ == "__main__":
    with open(OUT, "w") as f:
        json.dump(build(), f, indent=1, sort_keys=True)
    print("wrote", OUT)
