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
#include <algorithm>

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

__global__ void __launch_bounds__(kBlock) k_st_residual8(Geom g, const float* __restrict__ st,
                                                         const float* __restrict__ f, const float* __restrict__ u,
                                                         float* __restrict__ res) {
  st_residual8_body(g, st, f, u, res, (long long)blockIdx.x * blockDim.x + threadIdx.x);
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

__global__ void __launch_bounds__(kBlock) k_restrict8(Geom fg, Geom cg, const float* __restrict__ res,
                                                      float* __restrict__ fc) {
  restrict8_body(fg, cg, res, fc, (long long)blockIdx.x * blockDim.x + threadIdx.x);
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


// ---- blocked Gauss-Jordan inverse for coarsest levels beyond LDS size -----
// One workgroup per elimination step cannot move an n x n FP64 matrix fast
// enough (n = 1331 at 176^3 cells: 28 MB through one CU per pivot). The
// blocked form eliminates kGjB pivots per step with every workgroup:
//   A_KK <- A_KK^-1 = D,  A_KO <- D A_KO,  A_OO <- A_OO - A_OK D A_KO,
//   A_OK <- -A_OK D,
// the block version of the scalar update used in LDS (K: the step's pivot
// rows, O: all others). Work layout in B: the n x n matrix, then the old
// column panel C (n x kGjB), then D (kGjB x kGjB), then the diagonal (n).
constexpr int kGjB = 32;     // pivots per step
constexpr int kGjTile = 64;  // output tile edge of the update kernels

__global__ void __launch_bounds__(kBlock) k_gj_assemble(Geom g, const float* __restrict__ st, double* __restrict__ B,
                                                        double* __restrict__ diag) {
  // row v of the symmetrised coarse matrix: coupling (v, w) = mean of the
  // two stencil entries that describe it; zero row sums
  const int n = (int)g.nn;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int x = v % g.nx, y = (v / g.nx) % g.ny, z = v / (g.nx * g.ny);
  double off = 0.0;
  for (int o = 0; o < 27; o++) {
    if (o == 13) continue;
    const int w = wrapm(x + o % 3 - 1, g.nx) + g.nx * (wrapm(y + (o / 3) % 3 - 1, g.ny) +
                                                       g.ny * wrapm(z + o / 9 - 1, g.nz));
    const double a = 0.5 * ((double)st[(long long)o * n + v] + (double)st[(long long)(26 - o) * n + w]);
    if (w != v) {
      B[(long long)v * n + w] += a;
      off += a;
    }
  }
  B[(long long)v * n + v] = -off;
  diag[v] = -off;
}

__global__ void __launch_bounds__(kBlock) k_gj_fill(double* __restrict__ a, long long n, double v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    a[i] = v;
}

// B += (tr / n^2) on every entry: removes the constant null space.
__global__ void __launch_bounds__(kBlock) k_gj_shift(double* __restrict__ B, const double* __restrict__ diag, int n) {
  __shared__ double sh[kWaves];
  double acc[1] = {0.0};
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc[0] += diag[i];
  block_sum<1>(acc, sh);
  __shared__ double alpha;
  if (threadIdx.x == 0) alpha = acc[0] / ((double)n * (double)n);
  __syncthreads();
  const long long nn = (long long)n * n;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (long long)gridDim.x * blockDim.x)
    B[i] += alpha;
}

// Step k, part 1 (one workgroup): D = A_KK^-1 in LDS, written to A_KK and D.
__global__ void __launch_bounds__(kBlock) k_gj_pivot(double* __restrict__ B, int n, int k0, int kb,
                                                     double* __restrict__ D) {
  __shared__ double A[kGjB * kGjB], colk[kGjB], rowk[kGjB];
  for (int q = threadIdx.x; q < kb * kb; q += blockDim.x) A[q] = B[(long long)(k0 + q / kb) * n + k0 + q % kb];
  __syncthreads();
  for (int k = 0; k < kb; k++) {
    const double ip = 1.0 / A[k * kb + k];
    if (threadIdx.x < kb) {
      const int i = threadIdx.x;
      colk[i] = A[i * kb + k];
      rowk[i] = (i == k ? 1.0 : A[k * kb + i]) * ip;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kb * kb; q += blockDim.x) {
      const int i = q / kb, j = q % kb;
      A[q] = (i == k) ? rowk[j] : ((j == k ? 0.0 : A[q]) - colk[i] * rowk[j]);
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < kb * kb; q += blockDim.x) {
    D[q] = A[q];
    B[(long long)(k0 + q / kb) * n + k0 + q % kb] = A[q];
  }
}

// Step k, part 2: row panel A_KO <- D A_KO in place (column tiles), and the
// old column panel C = A_OK copied out (row tiles). Grid: column tiles, then
// row tiles.
__global__ void __launch_bounds__(kBlock) k_gj_panels(double* __restrict__ B, int n, int k0, int kb,
                                                      const double* __restrict__ D, double* __restrict__ C) {
  __shared__ double sD[kGjB * kGjB], sA[kGjB * kGjTile];
  const int ntile = (n + kGjTile - 1) / kGjTile;
  if ((int)blockIdx.x < ntile) {
    const int j0 = blockIdx.x * kGjTile;
    for (int q = threadIdx.x; q < kb * kb; q += blockDim.x) sD[q] = D[q];
    for (int q = threadIdx.x; q < kb * kGjTile; q += blockDim.x) {
      const int r = q / kGjTile, j = j0 + q % kGjTile;
      sA[q] = j < n ? B[(long long)(k0 + r) * n + j] : 0.0;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kb * kGjTile; q += blockDim.x) {
      const int r = q / kGjTile, jj = q % kGjTile, j = j0 + jj;
      if (j >= n || (j >= k0 && j < k0 + kb)) continue;
      double s = 0.0;
      for (int t = 0; t < kb; t++) s += sD[r * kb + t] * sA[t * kGjTile + jj];
      B[(long long)(k0 + r) * n + j] = s;
    }
  } else {
    const int i0 = ((int)blockIdx.x - ntile) * kGjTile;
    for (int q = threadIdx.x; q < kGjTile * kb; q += blockDim.x) {
      const int i = i0 + q / kb, c = q % kb;
      if (i < n) C[(long long)i * kGjB + c] = B[(long long)i * n + k0 + c];
    }
  }
}

// Step k, part 3: for rows i outside K, A_ij -= C_i R_j (j outside K) and
// A_iK = -C_i D. Tiles of kGjTile x kGjTile, 4 x 4 outputs per thread.
__global__ void __launch_bounds__(kBlock) k_gj_update(double* __restrict__ B, int n, int k0, int kb,
                                                      const double* __restrict__ D, const double* __restrict__ C) {
  __shared__ double sC[kGjTile * kGjB], sR[kGjB * kGjTile];
  const int ntile = (n + kGjTile - 1) / kGjTile;
  const int i0 = (blockIdx.x / ntile) * kGjTile, j0 = (blockIdx.x % ntile) * kGjTile;
  for (int q = threadIdx.x; q < kGjTile * kb; q += blockDim.x) {
    const int i = i0 + q / kb, c = q % kb;
    sC[(q / kb) * kGjB + c] = i < n ? C[(long long)i * kGjB + c] : 0.0;
  }
  for (int q = threadIdx.x; q < kb * kGjTile; q += blockDim.x) {
    const int r = q / kGjTile, j = j0 + q % kGjTile;
    double v = 0.0;
    if (j < n) v = (j >= k0 && j < k0 + kb) ? -D[r * kb + (j - k0)] : B[(long long)(k0 + r) * n + j];
    sR[r * kGjTile + q % kGjTile] = v;
  }
  __syncthreads();
  const int ti = threadIdx.x / 16, tj = threadIdx.x % 16;  // 16 x 16 threads, 4 x 4 outputs each
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b2 = 0; b2 < 4; b2++) acc[a][b2] = 0.0;
  for (int t = 0; t < kb; t++) {
    double cv[4], rv[4];
#pragma unroll
    for (int a = 0; a < 4; a++) cv[a] = sC[(ti + 16 * a) * kGjB + t];
#pragma unroll
    for (int b2 = 0; b2 < 4; b2++) rv[b2] = sR[t * kGjTile + tj + 16 * b2];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b2 = 0; b2 < 4; b2++) acc[a][b2] += cv[a] * rv[b2];
  }
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int i = i0 + ti + 16 * a;
    if (i >= n || (i >= k0 && i < k0 + kb)) continue;
#pragma unroll
    for (int b2 = 0; b2 < 4; b2++) {
      const int j = j0 + tj + 16 * b2;
      if (j >= n) continue;
      const long long q = (long long)i * n + j;
      if (j >= k0 && j < k0 + kb)
        B[q] = acc[a][b2];  // -C_i D
      else
        B[q] -= acc[a][b2];
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_gj_symmetrise(const double* __restrict__ A, double* __restrict__ Binv,
                                                          int n) {
  const long long nn = (long long)n * n;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nn; q += (long long)gridDim.x * blockDim.x) {
    const long long a = q / n, b = q - a * n;
    Binv[q] = 0.5 * (A[a * n + b] + A[b * n + a]);
  }
}

size_t coarsest_work_doubles(long long n) { return (size_t)(n * n + n * kGjB + kGjB * kGjB + n + 64); }

void launch_coarsest_blocked(hipStream_t s, const Geom& g, const float* st, double* B, double* Binv) {
  const int n = (int)g.nn;
  const long long nn = (long long)n * n;
  double* C = B + nn;
  double* D = C + (long long)n * kGjB;
  double* diag = D + kGjB * kGjB;
  const unsigned fill_blocks = (unsigned)std::min<long long>(4096, (nn + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_gj_fill, dim3(fill_blocks), dim3(kBlock), 0, s, B, nn, 0.0);
  hipLaunchKernelGGL(k_gj_assemble, dim3(nblk(n)), dim3(kBlock), 0, s, g, st, B, diag);
  hipLaunchKernelGGL(k_gj_shift, dim3(fill_blocks), dim3(kBlock), 0, s, B, diag, n);
  const int ntile = (n + kGjTile - 1) / kGjTile;
  for (int k0 = 0; k0 < n; k0 += kGjB) {
    const int kb = std::min(kGjB, n - k0);
    hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(kBlock), 0, s, B, n, k0, kb, D);
    hipLaunchKernelGGL(k_gj_panels, dim3(2 * ntile), dim3(kBlock), 0, s, B, n, k0, kb, D, C);
    hipLaunchKernelGGL(k_gj_update, dim3(ntile * ntile), dim3(kBlock), 0, s, B, n, k0, kb, D, C);
  }
  hipLaunchKernelGGL(k_gj_symmetrise, dim3(fill_blocks), dim3(kBlock), 0, s, B, Binv, n);
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

__global__ void __launch_bounds__(kTailThreads) k_tail_vcycle(TailLevels T) { tail_vcycle_body(T); }

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
  else if (st_resid8(g, false))
    hipLaunchKernelGGL(k_st_residual8, dim3(nblk(8 * g.nn)), dim3(kBlock), 0, s, g, S.f32, f, u, res);
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
  if (restrict8(coarse))
    hipLaunchKernelGGL(k_restrict8, dim3(nblk(8 * coarse.nn)), dim3(kBlock), 0, s, fine, coarse, res, fc);
  else
    hipLaunchKernelGGL(k_restrict, dim3(nblk(coarse.nn)), dim3(kBlock), 0, s, fine, coarse, res, fc);
}
void launch_prolong_add(hipStream_t s, const Geom& fine, const Geom& coarse, const float* uc, float* uf) {
  hipLaunchKernelGGL(k_prolong_add, dim3(nblk(coarse.nn)), dim3(kBlock), 0, s, fine, coarse, uc, uf);
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
    launch_coarsest_blocked(s, g, st, B, Binv);
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

void launch_tail_vcycle(hipStream_t s, const TailLevels& T) {
  hipLaunchKernelGGL(k_tail_vcycle, dim3(1), dim3(kTailThreads), 0, s, T);
}

}  // namespace otm
