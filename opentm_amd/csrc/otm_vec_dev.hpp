// otm_vec_dev.hpp -- device bodies of the vector updates and FP64
// reductions of the lock-step three-right-hand-side preconditioned
// conjugate gradient, shared by the single-cell launches (otm_vec.hip) and
// the batched multi-cell launches (otm_batch.hip). Each body works on block
// lb of the nb blocks assigned to one cell.
// The reference sums in FP32 through cub (culib/lib.cuh:595-760); here every
// inner product accumulates in FP64 with a fixed summation order.
#pragma once
#include "otm_kernels.hpp"

namespace otm {

__device__ inline void fill_body(float* a, long long n, float v, int lb, int nb) {
  const BlockRange br = block_range_v(n, lb, nb);
  for (long long i = br.b + threadIdx.x; i < br.e; i += blockDim.x) a[i] = v;
}

__device__ inline void dot3_body(long long nn, const float* __restrict__ a, const float* __restrict__ b,
                                 const Reduce& R, int lb, int nb) {
  double acc[OTM_NRHS] = {0, 0, 0};
  const BlockRange br = block_range_v(nn, lb, nb);
  for (long long i = br.b + threadIdx.x; i < br.e; i += blockDim.x) {
#pragma unroll
    for (int k = 0; k < OTM_NRHS; k++) acc[k] += (double)a[k * nn + i] * (double)b[k * nn + i];
  }
  reduce_block_at<OTM_NRHS>(acc, R, lb);
}

template <typename T>
__device__ inline void sum3_body(long long nn, const T* __restrict__ a, const Reduce& R, int lb, int nb) {
  double acc[OTM_NRHS] = {0, 0, 0};
  const BlockRange br = block_range_v(nn, lb, nb);
  for (long long i = br.b + threadIdx.x; i < br.e; i += blockDim.x) {
#pragma unroll
    for (int k = 0; k < OTM_NRHS; k++) acc[k] += (double)a[k * nn + i];
  }
  reduce_block_at<OTM_NRHS>(acc, R, lb);
}

// Iterative refinement: x64 += d (the FP32 correction), d = 0.
__device__ inline void refine_add_body(long long n, double* __restrict__ x64, float* __restrict__ d, int lb, int nb) {
  const BlockRange br = block_range_v(n, lb, nb);
  for (long long i = br.b + threadIdx.x; i < br.e; i += blockDim.x) {
    x64[i] = ld_stream(x64 + i) + (double)ld_stream(d + i);
    d[i] = 0.f;
  }
}

// x64 -= mean (per load case), x32 = x64 rounded.
__device__ inline void sub_mean_cast_body(long long nn, double* __restrict__ x64, float* __restrict__ x32,
                                          const PcgScalars* __restrict__ P, int lb, int nb) {
  const double m[OTM_NRHS] = {P->mean[0], P->mean[1], P->mean[2]};
  const BlockRange br = block_range_v(nn, lb, nb);
  for (long long i = br.b + threadIdx.x; i < br.e; i += blockDim.x) {
#pragma unroll
    for (int k = 0; k < OTM_NRHS; k++) {
      const double v = ld_stream(x64 + k * nn + i) - m[k];
      x64[k * nn + i] = v;
      x32[k * nn + i] = (float)v;
    }
  }
}

__device__ inline void sub_mean3_body(long long nn, float* a, const PcgScalars* __restrict__ P, int lb, int nb) {
  const float m0 = (float)P->mean[0], m1 = (float)P->mean[1], m2 = (float)P->mean[2];
  const BlockRange br = block_range_v(nn, lb, nb);
  for (long long i = br.b + threadIdx.x; i < br.e; i += blockDim.x) {
    a[i] -= m0;
    a[nn + i] -= m1;
    a[2 * nn + i] -= m2;
  }
}

// x += alpha p, r -= alpha q, partial r.r
// x + a*p into the unevaluated FP32 pair (hi, lo): the rounding errors of
// the product (exact through fma) and of the sum (TwoSum) go to lo, and the
// pair is renormalised, so the iterate keeps about twice FP32's precision.
// Explicitly rounded operations keep the compiler from contracting them.
__device__ inline void pair_add_prod(float& hi, float& lo, float a, float p) {
  const float b = __fmul_rn(a, p);
  const float eb = fmaf(a, p, -b);
  const float s = __fadd_rn(hi, b);
  const float bv = __fsub_rn(s, hi);
  const float es = __fadd_rn(__fsub_rn(hi, __fsub_rn(s, bv)), __fsub_rn(b, bv));
  const float l = __fadd_rn(lo, __fadd_rn(es, eb));
  const float h = __fadd_rn(s, l);
  lo = __fsub_rn(l, __fsub_rn(h, s));
  hi = h;
}

// x += alpha p (x as an FP32 pair when xlo is given), r -= alpha q, and the
// partial sums of r.r.
__device__ inline void pcg_xr_body(long long nn, float* __restrict__ x, float* __restrict__ xlo,
                                   float* __restrict__ r, const float* __restrict__ p, const float* __restrict__ q,
                                   const PcgScalars* __restrict__ P, const Reduce& R, int lb, int nb) {
  float al[OTM_NRHS];
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) al[k] = (float)P->alpha[k];
  double acc[OTM_NRHS] = {0, 0, 0};
  const BlockRange br = block_range_v(nn, lb, nb);
  for (long long i = br.b + threadIdx.x; i < br.e; i += blockDim.x) {
#pragma unroll
    for (int k = 0; k < OTM_NRHS; k++) {
      const long long j = k * nn + i;
      if (xlo) {
        float hi = ld_stream(x + j), lo = ld_stream(xlo + j);
        pair_add_prod(hi, lo, al[k], ld_stream(p + j));
        x[j] = hi;
        xlo[j] = lo;
      } else {
        x[j] = ld_stream(x + j) + al[k] * ld_stream(p + j);
      }
      const float rv = ld_stream(r + j) - al[k] * ld_stream(q + j);
      r[j] = rv;
      acc[k] += (double)rv * (double)rv;
    }
  }
  reduce_block_at<OTM_NRHS>(acc, R, lb);
}

// p = z + beta p
// p = z + beta p; first: p = z (a restart: the old p is not read).
__device__ inline void pcg_p_body(long long nn, float* __restrict__ p, const float* __restrict__ z,
                                  const PcgScalars* __restrict__ P, int lb, int nb, bool first = false) {
  float be[OTM_NRHS];
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) be[k] = (float)P->beta[k];
  const BlockRange br = block_range_v(nn, lb, nb);
  for (long long i = br.b + threadIdx.x; i < br.e; i += blockDim.x) {
#pragma unroll
    for (int k = 0; k < OTM_NRHS; k++) {
      const long long j = k * nn + i;
      p[j] = first ? ld_stream(z + j) : ld_stream(z + j) + be[k] * ld_stream(p + j);
    }
  }
}

// Stage 2 of every reduction: sums the per-block partials in block order and
// applies the finaliser (one block). With the value count a compile-time
// constant the trips over the partials unroll, so all of a thread's loads
// are in flight together (the runtime-count loop waited for each load in
// turn); each thread still adds its blocks tid, tid + 256, ... in order.
template <int NV>
__device__ inline void reduce_final_nv(const Reduce& R, int nblocks) {
  constexpr int kTrips = kMaxReduceBlocks / kBlock;
  __shared__ double sh[NV * kWaves];
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) acc[i] = 0.0;
#pragma unroll
  for (int r = 0; r < kTrips; r++) {
    const int b = threadIdx.x + kBlock * r;
    if (b < nblocks) {
#pragma unroll
      for (int i = 0; i < NV; i++) acc[i] += R.partials[(size_t)b * NV + i];
    }
  }
  block_sum<NV>(acc, sh);
  if (threadIdx.x == 0) {
    finalize(acc, NV, R);
    publish_host(acc, NV, R);
  }
}

// The same for more partial rows than kMaxReduceBlocks (the fine smoother's
// per-tile r.z): each thread sums its rows in order, four loads in flight.
template <int NV>
__device__ inline void reduce_final_many_nv(const Reduce& R, int nblocks) {
  __shared__ double sh[NV * kWaves];
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) acc[i] = 0.0;
#pragma unroll 4
  for (int b = threadIdx.x; b < nblocks; b += kBlock) {
#pragma unroll
    for (int i = 0; i < NV; i++) acc[i] += R.partials[(size_t)b * NV + i];
  }
  block_sum<NV>(acc, sh);
  if (threadIdx.x == 0) {
    finalize(acc, NV, R);
    publish_host(acc, NV, R);
  }
}

__device__ inline void reduce_final_body(const Reduce& R, int nv, int nblocks) {
  switch (nv) {
    case 2: reduce_final_nv<2>(R, nblocks); return;
    case 3: reduce_final_nv<3>(R, nblocks); return;
    case 6: reduce_final_nv<6>(R, nblocks); return;
    case 9: reduce_final_nv<9>(R, nblocks); return;
    default: break;
  }
  __shared__ double sh[kMaxReduce * kWaves];
  double acc[kMaxReduce] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblocks; b += kBlock)
    for (int i = 0; i < nv; i++) acc[i] += R.partials[(size_t)b * nv + i];
  block_sum<kMaxReduce>(acc, sh);
  if (threadIdx.x == 0) {
    finalize(acc, nv, R);
    publish_host(acc, nv, R);
  }
}

}  // namespace otm
