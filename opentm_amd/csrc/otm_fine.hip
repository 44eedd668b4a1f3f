// otm_fine.hip -- single-cell launches of the matrix-free fine-level
// kernels; the device bodies are in otm_fine_dev.hpp.
#include <stdexcept>

#include "otm_fine_dev.hpp"

namespace otm {

__global__ void __launch_bounds__(kBlock) k_fine_gs_color(Geom g, FineConsts fc, const float* __restrict__ kappa,
                                                          const float* __restrict__ f, float* u, int cy, int cz) {
  fine_gs_color_body(g, fc, kappa, f, u, cy, cz, gtid());
}

__global__ void __launch_bounds__(kBlock) k_fine_residual(Geom g, FineConsts fc, const float* __restrict__ kappa,
                                                          const float* __restrict__ f, const float* __restrict__ u,
                                                          float* __restrict__ res) {
  fine_residual_body(g, fc, kappa, f, u, res, gtid());
}

__global__ void __launch_bounds__(kBlock) k_fine_apply_dot(Geom g, FineConsts fc, const float* __restrict__ kappa,
                                                           const float* __restrict__ p, float* __restrict__ q,
                                                           Reduce R) {
  fine_apply_dot_body(g, fc, kappa, p, q, R, xcd_block(), gridDim.x);
}

// LDS-tiled forms of the two (kBlock threads: BX x kBlock/BX per tile)
template <int BX>
__global__ void __launch_bounds__(kBlock) k_fine_residual_tile(Geom g, FineConsts fc, const float* __restrict__ kappa,
                                                               const float* __restrict__ f,
                                                               const float* __restrict__ u, float* __restrict__ res) {
  fine_residual_tile_body<BX, kBlock / BX>(g, fc, kappa, f, u, res, xcd_block());
}

template <int BX>
__global__ void __launch_bounds__(kBlock) k_fine_apply_tile(Geom g, FineConsts fc, const float* __restrict__ kappa,
                                                            const float* __restrict__ p, float* __restrict__ q,
                                                            Reduce R) {
  fine_apply_tile_body<BX, kBlock / BX>(g, fc, kappa, p, q, R, xcd_block(), gridDim.x);
}

// grid (nblocks, 3): one load case per block row
template <typename XT>
__global__ void __launch_bounds__(kBlock) k_fine_init_residual_case(Geom g, FineConsts fc,
                                                                    const float* __restrict__ kappa,
                                                                    const XT* __restrict__ x_,
                                                                    const float* __restrict__ xlo,
                                                                    float* __restrict__ rvec, Reduce R) {
  fine_init_residual_case_body<XT>(g, fc, kappa, x_, xlo, rvec, R, xcd_block(), gridDim.x, (int)blockIdx.y);
}

// the same by whole x-rows (fine_init_residual_rows_body), nx <= kBlock
template <typename XT>
__global__ void __launch_bounds__(kBlock) k_fine_init_residual_rows(Geom g, FineConsts fc,
                                                                    const float* __restrict__ kappa,
                                                                    const XT* __restrict__ x_,
                                                                    const float* __restrict__ xlo,
                                                                    float* __restrict__ rvec, Reduce R) {
  fine_init_residual_rows_body<XT>(g, fc, kappa, x_, xlo, rvec, R, xcd_block(), gridDim.x, (int)blockIdx.y);
}

__global__ void __launch_bounds__(kBlock) k_fine_kappa_h(Geom g, FineConsts fc, const float* __restrict__ kappa,
                                                         ChiRef chi, Reduce R, float* __restrict__ S6) {
  fine_kappa_h_body(g, fc, kappa, chi, R, xcd_block(), gridDim.x, S6);
}

__global__ void __launch_bounds__(kBlock) k_sens_from_energy(Geom g, const float* __restrict__ S6, float w0, float w1,
                                                             float w2, float w3, float w4, float w5,
                                                             float* __restrict__ sens, int pitch,
                                                             const float* __restrict__ scale) {
  const float w[6] = {w0, w1, w2, w3, w4, w5};
  sens_from_energy_body(g, S6, w, sens, pitch, stid(), scale);
}

__global__ void __launch_bounds__(kBlock) k_fine_sensitivity(Geom g, FineConsts fc, ChiRef chi,
                                                             float w0, float w1, float w2, float w3, float w4, float w5,
                                                             float* __restrict__ sens, int pitch,
                                                             const float* __restrict__ scale) {
  const float w[6] = {w0, w1, w2, w3, w4, w5};
  fine_sensitivity_body(g, fc, chi, w, sens, pitch, gtid(), scale);
}

__global__ void __launch_bounds__(kBlock) k_fine_energy6(Geom g, FineConsts fc, ChiRef chi,
                                                         float* __restrict__ S6) {
  fine_energy6_body(g, fc, chi, S6, gtid());
}

// The occupancy hint keeps the fully unrolled colour loop from holding
// every node's loads in flight at once (which costs occupancy); the 30 KB
// LDS footprint of a 32x4x8 brick allows five workgroups per CU.
template <int BX, int BY, int BZ, int TR>
__global__ void __launch_bounds__(BX * TR) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_fine_smooth_brick(Geom g, FineConsts fc, const float* __restrict__ kappa, const float* __restrict__ f,
                    const float* __restrict__ uin, float* __restrict__ uout, int sweeps, int forward, int zero_in,
                    int dot, Reduce RD) {
  fine_brick_body<BX, BY, BZ, TR>(g, fc, kappa, f, uin, uout, sweeps, forward, zero_in, xcd_block(), dot, RD);
}

bool fine_tile_ok(const Geom& g, int v) {
  if (v < 0 || v >= kFineTileVariants) return false;
  const FineTileShape& t = kFineTileShapes[v];
  return g.nx % t.bx == 0 && g.ny % t.by == 0 && g.nz % t.bz == 0;
}
int fine_tile_default(const Geom& g) { return fine_tile_ok(g, 0) ? 0 : (fine_tile_ok(g, 1) ? 1 : -1); }
int fine_brick_x(const Geom& g) {
  const int v = fine_tile_default(g);
  return v < 0 ? 0 : kFineTileShapes[v].bx;
}
static long long fine_variant_tiles(const Geom& g, int v) {
  const FineTileShape& t = kFineTileShapes[v];
  return (long long)(g.nx / t.bx) * (g.ny / t.by) * (g.nz / t.bz);
}

// r.z in the smoother: 256-thread tiles, one partial-sum row per tile in
// the context's tile partials (one row per 512 nodes, at most kTileDotSlots)
bool fine_brick_dot_ok(const Geom& g, int v) {
  if (!fine_tile_ok(g, v)) return false;
  const FineTileShape& t = kFineTileShapes[v];
  return t.bx == 32 && t.bx * t.tr == kBlock && fine_variant_tiles(g, v) <= kTileDotSlots;
}

void launch_fine_smooth_brick(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa, const float* f,
                              const float* uin, float* uout, int sweeps, bool forward, bool zero_in, int v,
                              const Reduce* dot) {
  if (!fine_tile_ok(g, v)) throw std::invalid_argument("the cell does not tile with this fine tile shape");
  if (dot && !fine_brick_dot_ok(g, v)) throw std::invalid_argument("r.z in the smoother needs 256-thread 32-wide tiles");
  const unsigned nb = (unsigned)fine_variant_tiles(g, v);
  const Reduce RD = dot ? *dot : Reduce{};
  const int fw = forward ? 1 : 0, zi = zero_in ? 1 : 0, d = dot ? 1 : 0;
  switch (v) {
    case 0:
      hipLaunchKernelGGL((k_fine_smooth_brick<32, 4, 8, 8>), dim3(nb), dim3(256), 0, s, g, fc, kappa, f, uin, uout,
                         sweeps, fw, zi, d, RD);
      break;
    case 1:
      hipLaunchKernelGGL((k_fine_smooth_brick<16, 4, 8, 8>), dim3(nb), dim3(128), 0, s, g, fc, kappa, f, uin, uout,
                         sweeps, fw, zi, 0, RD);
      break;
    case 2:
      hipLaunchKernelGGL((k_fine_smooth_brick<32, 8, 4, 8>), dim3(nb), dim3(256), 0, s, g, fc, kappa, f, uin, uout,
                         sweeps, fw, zi, d, RD);
      break;
    case 3:
      hipLaunchKernelGGL((k_fine_smooth_brick<32, 4, 4, 4>), dim3(nb), dim3(128), 0, s, g, fc, kappa, f, uin, uout,
                         sweeps, fw, zi, 0, RD);
      break;
    default:
      hipLaunchKernelGGL((k_fine_smooth_brick<32, 8, 8, 8>), dim3(nb), dim3(256), 0, s, g, fc, kappa, f, uin, uout,
                         sweeps, fw, zi, d, RD);
      break;
  }
}

void launch_fine_brick_dot_final(hipStream_t s, const Geom& g, const Reduce& R, int v) {
  launch_reduce_final(s, R, OTM_NRHS, (unsigned)fine_variant_tiles(g, v));
}

// ---- launchers ---------------------------------------------------------
static inline unsigned nblk(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

void launch_fine_gs_color(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa, const float* f,
                          float* u, int color) {
  const long long cnt = (long long)g.nx * (g.ny / 2) * (g.nz / 2);
  hipLaunchKernelGGL(k_fine_gs_color, dim3(nblk(cnt)), dim3(kBlock), 0, s, g, fc, kappa, f, u, color & 1, color >> 1);
}

void launch_fine_residual(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa, const float* f,
                          const float* u, float* res) {
  const int bx = fine_tile_x(g);
  if (bx == 32)
    hipLaunchKernelGGL(k_fine_residual_tile<32>, dim3(fine_tiles(g, 32)), dim3(kBlock), 0, s, g, fc, kappa, f, u, res);
  else if (bx == 16)
    hipLaunchKernelGGL(k_fine_residual_tile<16>, dim3(fine_tiles(g, 16)), dim3(kBlock), 0, s, g, fc, kappa, f, u, res);
  else
    hipLaunchKernelGGL(k_fine_residual, dim3(nblk(g.nn)), dim3(kBlock), 0, s, g, fc, kappa, f, u, res);
}

// Levels that tile use the LDS-tiled kernel with its own workgroup count
// (fine_apply_tile_blocks); the reduction's partial count follows it.
void launch_fine_apply_dot(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa, const float* p,
                           float* q, const Reduce& R, unsigned nblocks) {
  const int bx = fine_tile_x(g);
  if (bx) {
    const unsigned nb = fine_apply_tile_blocks(g, bx);
    if (bx == 32)
      hipLaunchKernelGGL(k_fine_apply_tile<32>, dim3(nb), dim3(kBlock), 0, s, g, fc, kappa, p, q, R);
    else
      hipLaunchKernelGGL(k_fine_apply_tile<16>, dim3(nb), dim3(kBlock), 0, s, g, fc, kappa, p, q, R);
    launch_reduce_final(s, R, OTM_NRHS, nb);
    return;
  }
  hipLaunchKernelGGL(k_fine_apply_dot, dim3(nblocks), dim3(kBlock), 0, s, g, fc, kappa, p, q, R);
  launch_reduce_final(s, R, OTM_NRHS, nblocks);
}

// The FP64 residual by whole x-rows where rows have at most kBlock nodes
// (3 new iterate values per node and case); longer rows take the pointwise
// form. An LDS-tiled form (180 VGPRs, two waves per SIMD) measured 25%
// slower than the pointwise one.
template <typename XT>
static void launch_init_residual_xt(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa,
                                    const XT* x, const float* xlo, float* r, const Reduce& R, unsigned nblocks) {
  const dim3 grid(nblocks, OTM_NRHS), blk(kBlock);
  if (fine_rows_ok(g.nx))
    hipLaunchKernelGGL(k_fine_init_residual_rows<XT>, grid, blk, 0, s, g, fc, kappa, x, xlo, r, R);
  else
    hipLaunchKernelGGL(k_fine_init_residual_case<XT>, grid, blk, 0, s, g, fc, kappa, x, xlo, r, R);
  launch_reduce_final(s, R, 3 * OTM_NRHS, nblocks);
}

void launch_fine_init_residual(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa,
                               const float* x, const float* xlo, float* r, const Reduce& R, unsigned nblocks) {
  launch_init_residual_xt<float>(s, g, fc, kappa, x, xlo, r, R, nblocks);
}

void launch_fine_init_residual(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa,
                               const double* x, float* r, const Reduce& R, unsigned nblocks) {
  launch_init_residual_xt<double>(s, g, fc, kappa, x, nullptr, r, R, nblocks);
}

void launch_fine_kappa_h(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa, const ChiRef& chi,
                         const Reduce& R, unsigned nblocks, float* S6) {
  hipLaunchKernelGGL(k_fine_kappa_h, dim3(nblocks), dim3(kBlock), 0, s, g, fc, kappa, chi, R, S6);
  launch_reduce_final(s, R, 6, nblocks);
}

void launch_sens_from_energy(hipStream_t s, const Geom& g, const float* S6, const float w[6], float* sens, int pitch,
                             const float* scale) {
  hipLaunchKernelGGL(k_sens_from_energy, dim3(nblk(g.nn)), dim3(kBlock), 0, s, g, S6, w[0], w[1], w[2], w[3], w[4],
                     w[5], sens, pitch, scale);
}

void launch_fine_sensitivity(hipStream_t s, const Geom& g, const FineConsts& fc, const ChiRef& chi, const float w[6],
                             float* sens, int pitch, const float* scale) {
  hipLaunchKernelGGL(k_fine_sensitivity, dim3(nblk(g.nn)), dim3(kBlock), 0, s, g, fc, chi, w[0], w[1], w[2], w[3],
                     w[4], w[5], sens, pitch, scale);
}

void launch_fine_energy6(hipStream_t s, const Geom& g, const FineConsts& fc, const ChiRef& chi, float* S6) {
  hipLaunchKernelGGL(k_fine_energy6, dim3(nblk(g.nn)), dim3(kBlock), 0, s, g, fc, chi, S6);
}

}  // namespace otm
