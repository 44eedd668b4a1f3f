// otm_vec.hip -- single-cell launches of the PCG vector updates and FP64
// reductions (device bodies in otm_vec_dev.hpp), plus copies and fills.
#include <stdexcept>

#include "otm_vec_dev.hpp"

namespace otm {

static inline unsigned nblk(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__global__ void __launch_bounds__(kBlock) k_fill(float* a, long long n, float v) {
  fill_body(a, n, v, (int)blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_dot3(long long nn, const float* __restrict__ a,
                                                 const float* __restrict__ b, Reduce R) {
  dot3_body(nn, a, b, R, (int)blockIdx.x, gridDim.x);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_sum3(long long nn, const T* __restrict__ a, Reduce R) {
  sum3_body<T>(nn, a, R, (int)blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_refine_add(long long n, double* __restrict__ x64, float* __restrict__ d) {
  refine_add_body(n, x64, d, (int)blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_sub_mean_cast(long long nn, double* __restrict__ x64,
                                                          float* __restrict__ x32, const PcgScalars* __restrict__ P) {
  sub_mean_cast_body(nn, x64, x32, P, (int)blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_sum2(long long nn, const float* __restrict__ a, const float* __restrict__ b,
                                                 Reduce R) {
  double acc[2] = {0, 0};
  const BlockRange br = stream_range(nn);
  for (long long i = br.b + threadIdx.x; i < br.e; i += blockDim.x) {
    acc[0] += (double)a[i];
    acc[1] += (double)b[i];
  }
  reduce_block<2>(acc, R);
}

__global__ void __launch_bounds__(kBlock) k_sub_mean3(long long nn, float* a, const PcgScalars* __restrict__ P) {
  sub_mean3_body(nn, a, P, (int)blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_pcg_xr(long long nn, float* __restrict__ x, float* __restrict__ xlo,
                                                   float* __restrict__ r, const float* __restrict__ p,
                                                   const float* __restrict__ q, const PcgScalars* __restrict__ P,
                                                   Reduce R) {
  pcg_xr_body(nn, x, xlo, r, p, q, P, R, (int)blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_pcg_p(long long nn, float* __restrict__ p, const float* __restrict__ z,
                                                  const PcgScalars* __restrict__ P, int first) {
  pcg_p_body(nn, p, z, P, (int)blockIdx.x, gridDim.x, first != 0);
}

// dst[v] = src[x + pitch*(y + ny*z)] + add (pitched lexicographic input)
__global__ void __launch_bounds__(kBlock) k_copy_pitched(Geom g, const float* __restrict__ src, int pitch,
                                                         float* __restrict__ dst, float add) {
  const long long t = stid();
  if (t >= g.nn) return;
  const long long x = t % g.nx, r = t / g.nx;
  dst[t] = src[x + (long long)pitch * r] + add;
}

__global__ void __launch_bounds__(kBlock) k_copy_to_pitched(Geom g, const float* __restrict__ src,
                                                            float* __restrict__ dst, int pitch) {
  const long long t = stid();
  if (t >= g.nn) return;
  const long long x = t % g.nx, r = t / g.nx;
  dst[x + (long long)pitch * r] = src[t];
}

__global__ void __launch_bounds__(kBlock) k_reduce_final(Reduce R, int nv, int nblocks) {
  reduce_final_body(R, nv, nblocks);
}

__global__ void __launch_bounds__(kBlock) k_reduce_final_many3(Reduce R, int nblocks) {
  reduce_final_many_nv<OTM_NRHS>(R, nblocks);
}

void launch_reduce_final(hipStream_t s, const Reduce& R, int nv, unsigned nblocks) {
  if (nblocks > (unsigned)kMaxReduceBlocks) {
    if (nv != OTM_NRHS) throw std::invalid_argument("more partial rows than slots for this reduction");
    hipLaunchKernelGGL(k_reduce_final_many3, dim3(1), dim3(kBlock), 0, s, R, (int)nblocks);
    return;
  }
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, s, R, nv, (int)nblocks);
}

// x += d (x an FP32 pair when xlo is given), d cleared.
__global__ void __launch_bounds__(kBlock) k_add_pair(long long n, float* __restrict__ x, float* __restrict__ xlo,
                                                     float* __restrict__ d) {
  for (long long i = stid(); i < n; i += (long long)gridDim.x * blockDim.x) {
    if (xlo) {
      float hi = x[i], lo = xlo[i];
      pair_add_prod(hi, lo, 1.f, d[i]);
      x[i] = hi;
      xlo[i] = lo;
    } else {
      x[i] += d[i];
    }
    d[i] = 0.f;
  }
}
void launch_add_pair(hipStream_t s, long long n, float* x, float* xlo, float* d) {
  unsigned nb = nblk(n);
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL(k_add_pair, dim3(nb), dim3(kBlock), 0, s, n, x, xlo, d);
}

__global__ void __launch_bounds__(kBlock) k_f2d(long long n, const float* __restrict__ in, double* __restrict__ out) {
  const long long t = stid();
  if (t < n) out[t] = (double)in[t];
}
__global__ void __launch_bounds__(kBlock) k_fill_d(double* a, long long n, double v) {
  const long long t = stid();
  if (t < n) a[t] = v;
}
void launch_f2d(hipStream_t s, long long n, const float* in, double* out) {
  hipLaunchKernelGGL(k_f2d, dim3(nblk(n)), dim3(kBlock), 0, s, n, in, out);
}
void launch_fill_d(hipStream_t s, double* a, long long n, double v) {
  hipLaunchKernelGGL(k_fill_d, dim3(nblk(n)), dim3(kBlock), 0, s, a, n, v);
}

void launch_fill(hipStream_t s, float* a, long long n, float v) {
  unsigned nb = nblk(n);
  if (nb > 4096) nb = 4096;
  if (nb == 0) nb = 1;
  hipLaunchKernelGGL(k_fill, dim3(nb), dim3(kBlock), 0, s, a, n, v);
}
void launch_dot3(hipStream_t s, long long nn, const float* a, const float* b, const Reduce& R, unsigned nblocks) {
  hipLaunchKernelGGL(k_dot3, dim3(nblocks), dim3(kBlock), 0, s, nn, a, b, R);
  launch_reduce_final(s, R, OTM_NRHS, nblocks);
}
void launch_sum3(hipStream_t s, long long nn, const float* a, const Reduce& R, unsigned nblocks) {
  hipLaunchKernelGGL(k_sum3<float>, dim3(nblocks), dim3(kBlock), 0, s, nn, a, R);
  launch_reduce_final(s, R, OTM_NRHS, nblocks);
}
void launch_sum3(hipStream_t s, long long nn, const double* a, const Reduce& R, unsigned nblocks) {
  hipLaunchKernelGGL(k_sum3<double>, dim3(nblocks), dim3(kBlock), 0, s, nn, a, R);
  launch_reduce_final(s, R, OTM_NRHS, nblocks);
}
static inline unsigned stream_blocks(long long n) {
  long long b = (n + kBlock - 1) / kBlock;
  return (unsigned)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}
void launch_refine_add(hipStream_t s, long long n, double* x64, float* d) {
  hipLaunchKernelGGL(k_refine_add, dim3(stream_blocks(n)), dim3(kBlock), 0, s, n, x64, d);
}
void launch_sub_mean_cast(hipStream_t s, long long nn, double* x64, float* x32, const PcgScalars* P) {
  hipLaunchKernelGGL(k_sub_mean_cast, dim3(stream_blocks(nn)), dim3(kBlock), 0, s, nn, x64, x32, P);
}
void launch_sum2(hipStream_t s, long long nn, const float* a, const float* b, const Reduce& R, unsigned nblocks) {
  hipLaunchKernelGGL(k_sum2, dim3(nblocks), dim3(kBlock), 0, s, nn, a, b, R);
  launch_reduce_final(s, R, 2, nblocks);
}
void launch_sub_mean3(hipStream_t s, long long nn, float* a, const PcgScalars* P) {
  unsigned nb = nblk(nn);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_sub_mean3, dim3(nb), dim3(kBlock), 0, s, nn, a, P);
}
void launch_pcg_xr(hipStream_t s, long long nn, float* x, float* xlo, float* r, const float* p, const float* q,
                   const PcgScalars* P, const Reduce& R, unsigned nblocks) {
  hipLaunchKernelGGL(k_pcg_xr, dim3(nblocks), dim3(kBlock), 0, s, nn, x, xlo, r, p, q, P, R);
  launch_reduce_final(s, R, OTM_NRHS, nblocks);
}
void launch_pcg_p(hipStream_t s, long long nn, float* p, const float* z, const PcgScalars* P, bool first) {
  unsigned nb = nblk(nn);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_pcg_p, dim3(nb), dim3(kBlock), 0, s, nn, p, z, P, first ? 1 : 0);
}
void launch_copy_pitched(hipStream_t s, const Geom& g, const float* src, int pitch, float* dst, float add) {
  hipLaunchKernelGGL(k_copy_pitched, dim3(nblk(g.nn)), dim3(kBlock), 0, s, g, src, pitch, dst, add);
}
void launch_copy_to_pitched(hipStream_t s, const Geom& g, const float* src, float* dst, int pitch) {
  hipLaunchKernelGGL(k_copy_to_pitched, dim3(nblk(g.nn)), dim3(kBlock), 0, s, g, src, dst, pitch);
}

}  // namespace otm
