// otm_coarse.hip -- single-cell launches of the coarse multigrid levels
// (device bodies in otm_coarse_dev.hpp), Galerkin assembly and the
// coarsest dense FP64 factorisation.
//
// Stencil layout: st[o*nn + v] couples node v to v + off(o) with
// off(o) = (o%3-1, o/3%3-1, o/9-1); o = 13 is the diagonal (the reference
// keeps 27 SoA arrays the same way, grid.cpp:189-220, diagonal at index 13,
// grid.cu:707-809).
//
// Galerkin coarse operators A_{l+1} = P^T A_l P with trilinear P
// (grid.cu:811-940, grid_ex.cu:16-227) are formed element by element: a
// coarse element's 8x8 matrix is the sum over its 8 children c of
// Q_c^T E_c Q_c, where Q_c interpolates the coarse element's corners to the
// child's corners. This equals the stencil-level triple product exactly.
#include "otm_coarse_dev.hpp"

namespace otm {

static inline unsigned nblk(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ---- per-level kernels ----
template <bool HALF>
__global__ void __launch_bounds__(kBlock) k_st_gs_color(Geom g, StRef S, const float* __restrict__ f, float* u, int cx,
                                                        int cy, int cz) {
  st_gs_color_body<HALF>(g, S, f, u, cx, cy, cz, gtid());
}

template <bool HALF>
__global__ void __launch_bounds__(kBlock) k_st_residual(Geom g, StRef S, const float* __restrict__ f,
                                                        const float* __restrict__ u, float* __restrict__ res) {
  st_residual_body<HALF>(g, S, f, u, res, gtid());
}

__global__ void __launch_bounds__(kBlock) k_stencil_absmax(Geom g, const float* __restrict__ st,
                                                           unsigned int* __restrict__ bits) {
  stencil_absmax_body(g, st, bits, xcd_block(), gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_stencil_pack(Geom g, const float* __restrict__ st,
                                                         const float* __restrict__ scale, _Float16* __restrict__ st16) {
  stencil_pack_body(g, st, scale, st16, gtid());
}

__global__ void __launch_bounds__(kBlock) k_restrict(Geom fg, Geom cg, const float* __restrict__ res,
                                                     float* __restrict__ fc) {
  restrict_body(fg, cg, res, fc, gtid());
}

__global__ void __launch_bounds__(kBlock) k_prolong_add(Geom fg, Geom cg, const float* __restrict__ uc,
                                                        float* __restrict__ uf) {
  prolong_body(fg, cg, uc, uf, gtid());
}

__global__ void __launch_bounds__(kBlock) k_galerkin_first(Geom fg, Geom cg, GalerkinConsts G,
                                                           const float* __restrict__ kappa, float* __restrict__ E) {
  galerkin_first_body(fg, cg, G, kappa, E, gtid());
}

__global__ void __launch_bounds__(kBlock) k_galerkin_next(Geom fg, Geom cg, const float* __restrict__ Ein,
                                                          float* __restrict__ Eout) {
  galerkin_next_body(fg, cg, Ein, Eout, gtid());
}

__global__ void __launch_bounds__(kBlock) k_stencil_from_elements(Geom g, const float* __restrict__ E,
                                                                  float* __restrict__ st) {
  stencil_from_elements_body(g, E, st, gtid());
}

__global__ void __launch_bounds__(kFactorThreads) k_coarsest_invert_lds(Geom g, const float* __restrict__ st,
                                                                        double* __restrict__ Binv) {
  extern __shared__ double A[];
  coarsest_invert_lds_body(g, st, A, Binv);
}

__global__ void __launch_bounds__(kFactorThreads) k_coarsest_invert_global(Geom g, const float* __restrict__ st,
                                                                           double* B, double* Binv) {
  coarsest_invert_global_body(g, st, B, Binv);
}

__global__ void __launch_bounds__(kBlock) k_coarsest_solve(Geom g, const double* __restrict__ Binv,
                                                           const float* __restrict__ f, float* __restrict__ u) {
  coarsest_solve_body(g, Binv, f, u, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

template <bool HALF>
__global__ void __launch_bounds__(kBrickThreads) k_st_smooth_brick(Geom g, StRef S, const float* __restrict__ f,
                                                                   const float* __restrict__ uin,
                                                                   float* __restrict__ uout, int sweeps,
                                                                   int forward, int zero_in) {
  st_brick_body<HALF>(g, S, f, uin, uout, sweeps, forward, zero_in, xcd_block());
}

// ---- launchers ----
void launch_st_gs_color(hipStream_t s, const Geom& g, const StRef& S, const float* f, float* u, int color) {
  const long long cnt = (long long)(g.nx / 2) * (g.ny / 2) * (g.nz / 2);
  if (S.f16)
    hipLaunchKernelGGL(k_st_gs_color<true>, dim3(nblk(cnt)), dim3(kBlock), 0, s, g, S, f, u, color & 1,
                       (color >> 1) & 1, color >> 2);
  else
    hipLaunchKernelGGL(k_st_gs_color<false>, dim3(nblk(cnt)), dim3(kBlock), 0, s, g, S, f, u, color & 1,
                       (color >> 1) & 1, color >> 2);
}
void launch_st_residual(hipStream_t s, const Geom& g, const StRef& S, const float* f, const float* u, float* res) {
  if (S.f16)
    hipLaunchKernelGGL(k_st_residual<true>, dim3(nblk(g.nn)), dim3(kBlock), 0, s, g, S, f, u, res);
  else
    hipLaunchKernelGGL(k_st_residual<false>, dim3(nblk(g.nn)), dim3(kBlock), 0, s, g, S, f, u, res);
}
void launch_stencil_pack(hipStream_t s, const Geom& g, const float* st, float* scale, _Float16* st16) {
  unsigned int* bits = reinterpret_cast<unsigned int*>(scale);
  launch_fill(s, scale, 1, 0.f);
  unsigned nb = nblk(g.nn);
  if (nb > 1024) nb = 1024;
  hipLaunchKernelGGL(k_stencil_absmax, dim3(nb), dim3(kBlock), 0, s, g, st, bits);
  hipLaunchKernelGGL(k_stencil_pack, dim3(nblk(g.nn)), dim3(kBlock), 0, s, g, st, scale, st16);
}
void launch_restrict(hipStream_t s, const Geom& fine, const Geom& coarse, const float* res, float* fc) {
  hipLaunchKernelGGL(k_restrict, dim3(nblk(coarse.nn)), dim3(kBlock), 0, s, fine, coarse, res, fc);
}
void launch_prolong_add(hipStream_t s, const Geom& fine, const Geom& coarse, const float* uc, float* uf) {
  hipLaunchKernelGGL(k_prolong_add, dim3(nblk(fine.nn)), dim3(kBlock), 0, s, fine, coarse, uc, uf);
}
void launch_galerkin_first(hipStream_t s, const Geom& fine, const Geom& coarse, const GalerkinConsts& G,
                           const float* kappa, float* E) {
  hipLaunchKernelGGL(k_galerkin_first, dim3(nblk(coarse.nn)), dim3(kBlock), 0, s, fine, coarse, G, kappa, E);
}
void launch_galerkin_next(hipStream_t s, const Geom& fine, const Geom& coarse, const float* Ein, float* Eout) {
  hipLaunchKernelGGL(k_galerkin_next, dim3(nblk(coarse.nn)), dim3(kBlock), 0, s, fine, coarse, Ein, Eout);
}
void launch_stencil_from_elements(hipStream_t s, const Geom& g, const float* E, float* st) {
  hipLaunchKernelGGL(k_stencil_from_elements, dim3(nblk(g.nn)), dim3(kBlock), 0, s, g, E, st);
}
void launch_coarsest_factor(hipStream_t s, const Geom& g, const float* st, double* B, double* Binv) {
  // always kFactorThreads threads: the assembly's trace sum, and so the
  // inverse, is then bitwise the same as in the batched setup
  const int n = (int)g.nn;
  const int nt = kFactorThreads;
  if (n <= kInvertLdsMax) {
    // per device and cheap; set on every call (once per design update)
    (void)hipFuncSetAttribute((const void*)k_coarsest_invert_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(kInvertLdsMax * kInvertLdsMax * sizeof(double)));
    hipLaunchKernelGGL(k_coarsest_invert_lds, dim3(1), dim3(nt), (size_t)n * n * sizeof(double), s, g, st, Binv);
  } else {
    hipLaunchKernelGGL(k_coarsest_invert_global, dim3(1), dim3(nt), 0, s, g, st, B, Binv);
  }
}
bool brick_ok(const Geom& g) { return g.nx >= 2 && g.ny >= 2 && g.nz >= 2; }
void launch_st_smooth_brick(hipStream_t s, const Geom& g, const StRef& S, const float* f, const float* uin,
                            float* uout, int sweeps, bool forward, bool zero_in) {
  const unsigned nb = (unsigned)brick_count(g);
  if (S.f16)
    hipLaunchKernelGGL(k_st_smooth_brick<true>, dim3(nb), dim3(kBrickThreads), 0, s, g, S, f, uin, uout, sweeps,
                       forward ? 1 : 0, zero_in ? 1 : 0);
  else
    hipLaunchKernelGGL(k_st_smooth_brick<false>, dim3(nb), dim3(kBrickThreads), 0, s, g, S, f, uin, uout, sweeps,
                       forward ? 1 : 0, zero_in ? 1 : 0);
}
void launch_coarsest_solve(hipStream_t s, const Geom& g, const double* Binv, const float* f, float* u) {
  hipLaunchKernelGGL(k_coarsest_solve, dim3(nblk(g.nn), OTM_NRHS), dim3(kBlock), 0, s, g, Binv, f, u);
}

}  // namespace otm
