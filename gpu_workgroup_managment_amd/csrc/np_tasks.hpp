// np_tasks.hpp -- the non-persistent tasks that borrow workgroups from a
// resident mega-kernel: int32 matmult (the reference's task), FP32 square
// matmul and the int32 minimum reduction. The task kind is chosen per task;
// every kind's buffers are resident for the life of the kernel.
#pragma once

#include <type_traits>

#include "wgm_device.hpp"
#include "wgm.h"

namespace wgm {

struct NPDev {
  int dim;  // square matrix dimension (multiple of 32)
  int len;  // reduction length
  const int* A;
  const int* B;
  int* C;
  const float* Af;
  const float* Bf;
  float* Cf;
  const int* buf;
  int* dmin;
  int* counter;                    // finished-workgroup counter (reset per task)
  double* partial;                 // FP32 matmul: one weighted-checksum partial per workgroup
  int* h_result;                   // coherent host cell
  unsigned long long* h_checksum;  // coherent host cell
};

__device__ __forceinline__ int pow2_ceil(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

// LDS of one 256-thread slice of the MFMA matmul: the accumulators of three
// of its four waves (64 lanes x 16 floats each) for the split-K reduction.
constexpr int kMfmaSliceFloats = 3 * 64 * 16;

// Bytes of dynamic LDS the tasks need at a workgroup size: the larger of the
// matmul slices (int32 tiles: 2 x 32 x 33 words; FP32 split-K accumulators)
// and the FP64 checksum reduction (bd doubles).
__host__ __device__ constexpr int np_lds_bytes(int bd) {
  return (bd >= 256 ? bd / 256 : 1) * kMfmaSliceFloats * 4 > bd * 8 ? (bd >= 256 ? bd / 256 : 1) * kMfmaSliceFloats * 4
                                                                   : bd * 8;
}

// Checksum weights of the FP32 matmul: sum_ij C[i][j] * u(i) * v(j) catches a
// transposed or permuted result, which a plain sum would not.
__host__ __device__ constexpr int np_weight_u(int i) { return i % 7 + 1; }
__host__ __device__ constexpr int np_weight_v(int j) { return j % 5 + 1; }

typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 square matmul on the matrix cores. A 256-thread slice computes one
// 32x32 tile of C: its four waves each run v_mfma_f32_32x32x2f32 over a
// quarter of K, with operands loaded straight from L2 in batches of 8 steps
// (the next batch in flight while the current one computes), and the four
// partial accumulators are summed through LDS. At dim 256 the 64 tiles keep
// 64 slices busy, one latency round per batch. Operand layout of the
// instruction (wave64): lane l supplies A[row l%32][k l/32] and
// B[k l/32][col l%32]; accumulator register r of lane l holds
// C[8(r/4) + 4(l/32) + r%4][l%32]. dim % 64 == 0. Returns this thread's share
// of the weighted checksum.
__device__ double np_matmul_f32_mfma(const float* A, const float* B, float* C, int dim, int vid, int nwg, float* lds) {
  double part = 0.0;
  const int q = blockDim.x / 256;
  const int slice = threadIdx.x / 256, lt = threadIdx.x % 256;
  const int wave = lt / 64, lane = lt % 64;
  float* red = lds + slice * kMfmaSliceFloats;
  const int tiles = dim / 32;
  const int nt = tiles * tiles;
  const int kq = dim / 4;  // K range of one wave (a multiple of 16)
  const int row = lane % 32, kl = lane / 32;
  for (int base = vid * q; base < nt; base += nwg * q) {
    const int tile = base + slice;
    const bool live = tile < nt;
    const int tr = live ? tile / tiles : 0, tc = live ? tile % tiles : 0;
    f32x16 acc;
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
    if (live) {
      const float* ap = A + (tr * 32 + row) * dim + wave * kq + kl;  // + 2 per step
      const float* bp = B + (wave * kq + kl) * dim + tc * 32 + row;  // + 2 dim per step
      float a0[8], b0[8], a1[8], b1[8];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        a0[i] = ap[2 * i];
        b0[i] = bp[2 * i * dim];
      }
      for (int k0 = 0; k0 < kq; k0 += 16) {
        const bool more = k0 + 16 < kq;
        if (more) {
#pragma unroll
          for (int i = 0; i < 8; i++) {
            a1[i] = ap[k0 + 16 + 2 * i];
            b1[i] = bp[(k0 + 16 + 2 * i) * dim];
          }
        }
#pragma unroll
        for (int i = 0; i < 8; i++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a0[i], b0[i], acc, 0, 0, 0);
        if (more) {
#pragma unroll
          for (int i = 0; i < 8; i++) {
            a0[i] = a1[i];
            b0[i] = b1[i];
          }
        }
      }
      if (wave > 0) {
#pragma unroll
        for (int r = 0; r < 16; r++) red[((wave - 1) * 16 + r) * 64 + lane] = acc[r];
      }
    }
    __syncthreads();
    if (live && wave == 0) {
      const int j = tc * 32 + row;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const float c = acc[r] + red[r * 64 + lane] + red[(16 + r) * 64 + lane] + red[(32 + r) * 64 + lane];
        const int i = tr * 32 + 8 * (r / 4) + 4 * kl + (r % 4);
        C[i * dim + j] = c;
        part += (double)c * (double)(np_weight_u(i) * np_weight_v(j));
      }
    }
    __syncthreads();
  }
  return part;
}

// Square matmul on 32x32 output tiles staged through LDS. A workgroup of
// 256*q threads works on q tiles at once (one per 256-thread slice); each
// thread owns a 2x2 block of its tile. T = unsigned: the int32 matmult with
// the reference's wrap-around arithmetic (matmult.cl:26-67; exact in any
// order). T = float: FP32 FMA accumulation in k order.
template <typename T, typename S>
__device__ void np_matmul_tiled(const S* A, const S* B, S* C, int dim, int vid, int nwg, T* lds) {
  const int tiles = dim / 32;
  const int q = blockDim.x / 256;
  const int slice = threadIdx.x / 256, lt = threadIdx.x % 256;
  T* As = lds + slice * 2 * 32 * 33;
  T* Bs = As + 32 * 33;
  const int ty = (lt / 16) * 2, tx = (lt % 16) * 2;
  const int ntiles = tiles * tiles;
  for (int base = vid * q; base < ntiles; base += nwg * q) {
    const int tile = base + slice;
    const bool live = tile < ntiles;
    const int tr = live ? tile / tiles : 0, tc = live ? tile % tiles : 0;
    T acc00 = 0, acc01 = 0, acc10 = 0, acc11 = 0;
    for (int k0 = 0; k0 < dim; k0 += 32) {
      if (live) {
        for (int e = lt; e < 32 * 32; e += 256) {
          const int r = e / 32, cc = e % 32;
          As[r * 33 + cc] = (T)A[(tr * 32 + r) * dim + k0 + cc];
          Bs[r * 33 + cc] = (T)B[(k0 + r) * dim + tc * 32 + cc];
        }
      }
      __syncthreads();
      if (live) {
#pragma unroll 8
        for (int kk = 0; kk < 32; kk++) {
          const T a0 = As[ty * 33 + kk], a1 = As[(ty + 1) * 33 + kk];
          const T b0 = Bs[kk * 33 + tx], b1 = Bs[kk * 33 + tx + 1];
          if constexpr (std::is_same<T, float>::value) {
            acc00 = __builtin_fmaf(a0, b0, acc00);
            acc01 = __builtin_fmaf(a0, b1, acc01);
            acc10 = __builtin_fmaf(a1, b0, acc10);
            acc11 = __builtin_fmaf(a1, b1, acc11);
          } else {
            acc00 += a0 * b0;
            acc01 += a0 * b1;
            acc10 += a1 * b0;
            acc11 += a1 * b1;
          }
        }
      }
      __syncthreads();
    }
    if (live) {
      S* c0 = C + (tr * 32 + ty) * dim + tc * 32 + tx;
      c0[0] = (S)acc00;
      c0[1] = (S)acc01;
      c0[dim] = (S)acc10;
      c0[dim + 1] = (S)acc11;
    }
  }
}

// Element-strided fallback for workgroups narrower than 256 threads
// (the reference's grid-stride form, matmult.cl:40-56).
template <typename T, typename S>
__device__ void np_matmul_strided(const S* A, const S* B, S* C, int dim, int vid, int nwg) {
  for (int i = vid * blockDim.x + threadIdx.x; i < dim * dim; i += nwg * blockDim.x) {
    const int r = i / dim, col = i % dim;
    T acc = 0;
    for (int j = 0; j < dim; j++) {
      if constexpr (std::is_same<T, float>::value) acc = __builtin_fmaf((T)A[r * dim + j], (T)B[j * dim + col], acc);
      else acc += (T)A[r * dim + j] * (T)B[j * dim + col];
    }
    C[i] = (S)acc;
  }
}

// MY_reduce (src/non_persistent_kernels/reduce/device/reduce.cl:4-41): the
// minimum. Each thread keeps a running minimum over 16-byte loads strided
// across the task's workgroups; one LDS tree per workgroup, one atomic min.
__device__ void np_reduce_min(const NPDev& t, int vid, int nwg, int* lds) {
  int mn = INT32_MAX;
  const int nvec = t.len / 4;
  const int4* v4 = (const int4*)t.buf;
  const int stride = nwg * blockDim.x;
  // four independent 16-byte loads in flight per iteration
  int i = vid * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < nvec; i += 4 * stride) {
    const int4 x0 = v4[i], x1 = v4[i + stride], x2 = v4[i + 2 * stride], x3 = v4[i + 3 * stride];
    mn = min(mn, min(min(min(x0.x, x0.y), min(x0.z, x0.w)), min(min(x1.x, x1.y), min(x1.z, x1.w))));
    mn = min(mn, min(min(min(x2.x, x2.y), min(x2.z, x2.w)), min(min(x3.x, x3.y), min(x3.z, x3.w))));
  }
  for (; i < nvec; i += stride) {
    const int4 x = v4[i];
    mn = min(mn, min(min(x.x, x.y), min(x.z, x.w)));
  }
  for (int i = nvec * 4 + vid * blockDim.x + threadIdx.x; i < t.len; i += stride) mn = min(mn, t.buf[i]);
  // wave minimum, then across waves
  for (int off = 32; off > 0; off >>= 1) mn = min(mn, __shfl_xor(mn, off, 64));
  const int wave = threadIdx.x / 64, nwaves = (blockDim.x + 63) / 64;
  if ((threadIdx.x & 63) == 0) lds[wave] = mn;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = lds[0];
    for (int w = 1; w < nwaves; w++) m = min(m, lds[w]);
    __hip_atomic_fetch_min(t.dmin, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
}

// Task body + completion: the workgroup that finishes last publishes the
// task's summary (djb2 diagonal hash / minimum / FP64 checksum) to the host
// cell with a system-scope release.
// `executing` (resident kernels): the task's running-workgroup count, which
// each workgroup decrements once its share is published; the last one after
// it has published the task's summary.
__device__ void np_task(const NPDev& t, int kind, int vid, int nwg, int* lds, unsigned long long* clk = nullptr,
                        int* executing = nullptr) {
  const bool tiled = blockDim.x % 256 == 0;
  if (clk && threadIdx.x == 0)
    __hip_atomic_fetch_min(&clk[0], (unsigned long long)wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (kind == WGM_NP_MATMULT_I32) {
    if (tiled) np_matmul_tiled<unsigned, int>(t.A, t.B, t.C, t.dim, vid, nwg, (unsigned*)lds);
    else np_matmul_strided<unsigned, int>(t.A, t.B, t.C, t.dim, vid, nwg);
  } else if (kind == WGM_NP_MATMUL_F32) {
    if (tiled && t.dim % 64 == 0 && t.partial) {
      // each workgroup reduces its share of the checksum; the last one adds
      // the partials in workgroup order
      double part = np_matmul_f32_mfma(t.Af, t.Bf, t.Cf, t.dim, vid, nwg, (float*)lds);
      for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
      double* red = (double*)lds;
      __syncthreads();
      if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = part;
      __syncthreads();
      if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int w = 0; w < (int)blockDim.x / 64; w++) sum += red[w];
        t.partial[vid] = sum;
      }
    } else if (tiled) {
      np_matmul_tiled<float, float>(t.Af, t.Bf, t.Cf, t.dim, vid, nwg, (float*)lds);
    } else {
      np_matmul_strided<float, float>(t.Af, t.Bf, t.Cf, t.dim, vid, nwg);
    }
  } else {
    np_reduce_min(t, vid, nwg, lds);
  }
  drain_vmem();
  __syncthreads();
  __shared__ int s_last;
  if (threadIdx.x == 0) {
    if (clk)
      __hip_atomic_fetch_max(&clk[1], (unsigned long long)wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    fence_release();
    s_last = add_acq_rel(t.counter, 1) == nwg - 1;
  }
  __syncthreads();
  if (!s_last) {
    // this share was released by the counter increment above
    if (executing && threadIdx.x == 0) add_relaxed(executing, -1);
    return;
  }
  if (kind == WGM_NP_MATMULT_I32) {
    if (threadIdx.x == 0) {
      unsigned h = 5381;
      for (int i = 0; i < t.dim; i++) h = h * 33u + (unsigned)t.C[i * t.dim + i];
      __hip_atomic_store(t.h_result, (int)h, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  } else if (kind == WGM_NP_REDUCE_MIN) {
    if (threadIdx.x == 0)
      __hip_atomic_store(t.h_result, ld_relaxed(t.dmin), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  } else if (blockDim.x % 256 == 0 && t.dim % 64 == 0 && t.partial) {
    if (threadIdx.x == 0) {
      double sum = 0.0;
      for (int w = 0; w < nwg; w++) sum += t.partial[w];
      unsigned long long bits;
      __builtin_memcpy(&bits, &sum, 8);
      __hip_atomic_store(t.h_checksum, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(t.h_result, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  } else {
    // weighted checksum in FP64: sum_ij C[i][j] u(i) v(j)
    double* red = (double*)lds;
    double acc = 0.0;
    for (int i = threadIdx.x; i < t.dim * t.dim; i += blockDim.x)
      acc += (double)t.Cf[i] * (double)(np_weight_u(i / t.dim) * np_weight_v(i % t.dim));
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = pow2_ceil(blockDim.x) / 2; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off && (int)threadIdx.x + off < (int)blockDim.x) red[threadIdx.x] += red[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      unsigned long long bits;
      __builtin_memcpy(&bits, &red[0], 8);
      __hip_atomic_store(t.h_checksum, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(t.h_result, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
  __syncthreads();
  if (executing && threadIdx.x == 0) __hip_atomic_fetch_add(executing, -1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace wgm
