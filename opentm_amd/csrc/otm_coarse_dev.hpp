// otm_coarse_dev.hpp -- device bodies of the coarse multigrid levels
// (stored 27-point stencils, transfers, coarsest solve), shared by the
// single-cell launches (otm_coarse.hip) and the batched multi-cell launches
// (otm_batch.hip).
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
#pragma once
#include "otm_kernels.hpp"

namespace otm {

__device__ inline long long lin(const Geom& g, int x, int y, int z) {
  return (long long)x + (long long)g.nx * ((long long)y + (long long)g.ny * z);
}

// ---- per-node operations of the stored-stencil levels ----
// Stencil entries of node v. FP16 storage (HALF) keeps the 26 couplings
// divided by a per-level scale; the diagonal is rebuilt as minus their sum,
// so rows keep a zero sum and the constants stay in the null space.
template <bool HALF>
__device__ inline void load_stencil(const StRef& S, long long nn, long long v, float a[27]) {
  if constexpr (HALF) {
    const float sc = *S.scale;
    float off = 0.f;
#pragma unroll
    for (int o = 0; o < 27; o++) {
      if (o == 13) continue;
      a[o] = (float)S.f16[o * nn + v] * sc;
      off += a[o];
    }
    a[13] = -off;
  } else {
#pragma unroll
    for (int o = 0; o < 27; o++) a[o] = S.f32[o * nn + v];
  }
}

// GS update at node (x,y,z): u = (f - sum_{o != 13} a_o u_o) / a_13.
template <bool HALF>
__device__ inline void st_gs_node(const Geom& g, const StRef& S, const float* __restrict__ f, float* u, int x, int y,
                                  int z) {
  const long long v = lin(g, x, y, z);
  const int xs[3] = {x == 0 ? g.nx - 1 : x - 1, x, x == g.nx - 1 ? 0 : x + 1};
  const int ys[3] = {y == 0 ? g.ny - 1 : y - 1, y, y == g.ny - 1 ? 0 : y + 1};
  const int zs[3] = {z == 0 ? g.nz - 1 : z - 1, z, z == g.nz - 1 ? 0 : z + 1};
  float a[27];
  load_stencil<HALF>(S, g.nn, v, a);
  const float inv = 1.f / a[13];
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) {
    const float* uk = u + k * g.nn;
    float s = 0.f;
#pragma unroll
    for (int o = 0; o < 27; o++) {
      if (o == 13) continue;
      s += a[o] * uk[lin(g, xs[o % 3], ys[(o / 3) % 3], zs[o / 9])];
    }
    u[k * g.nn + v] = (f[k * g.nn + v] - s) * inv;
  }
}

// res = f - A u at node t.
template <bool HALF>
__device__ inline void st_resid_node(const Geom& g, const StRef& S, const float* __restrict__ f, const float* u,
                                     float* res, long long t) {
  int x, y, z;
  split_node(g.nx, g.ny, t, x, y, z);
  const int xs[3] = {x == 0 ? g.nx - 1 : x - 1, x, x == g.nx - 1 ? 0 : x + 1};
  const int ys[3] = {y == 0 ? g.ny - 1 : y - 1, y, y == g.ny - 1 ? 0 : y + 1};
  const int zs[3] = {z == 0 ? g.nz - 1 : z - 1, z, z == g.nz - 1 ? 0 : z + 1};
  float a[27];
  if constexpr (HALF) {
    load_stencil<HALF>(S, g.nn, t, a);
  } else {
#pragma unroll
    for (int o = 0; o < 27; o++) a[o] = ld_off(S.f32 + (size_t)o * (size_t)g.nn, 4u * (unsigned)t);
  }
  // neighbour byte offsets from wave-uniform bases in 32 bits (64-bit
  // per-neighbour addresses were half of this kernel's instructions)
  const unsigned nx = (unsigned)g.nx, nxy = (unsigned)g.nx * (unsigned)g.ny;
  unsigned off[27];
#pragma unroll
  for (int o = 0; o < 27; o++)
    off[o] = 4u * ((unsigned)xs[o % 3] + nx * (unsigned)ys[(o / 3) % 3] + nxy * (unsigned)zs[o / 9]);
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) {
    const float* uk = u + (size_t)k * (size_t)g.nn;
    float s = 0.f;
#pragma unroll
    for (int o = 0; o < 27; o++) s += a[o] * ld_off(uk, off[o]);
    res[k * g.nn + t] = f[k * g.nn + t] - s;
  }
}

// f_c(I) = sum_{d in {-1,0,1}^3} w(d) res(2I + d), w = prod(d_i == 0 ? 1 : 1/2)
// (restrict_residual_kernel_H, grid.cu:811-878, weights (2-|d|)/2 per axis).
__device__ inline void restrict_node(const Geom& fg, const Geom& cg, const float* res, float* fc, long long t) {
  int X, Y, Z;
  split_node(cg.nx, cg.ny, t, X, Y, Z);
  const int x = 2 * X, y = 2 * Y, z = 2 * Z;
  const int xs[3] = {x == 0 ? fg.nx - 1 : x - 1, x, x + 1};
  const int ys[3] = {y == 0 ? fg.ny - 1 : y - 1, y, y + 1};
  const int zs[3] = {z == 0 ? fg.nz - 1 : z - 1, z, z + 1};
  const float w1[3] = {0.5f, 1.f, 0.5f};
  // byte offsets in 32 bits from wave-uniform bases
  unsigned rb[9];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int b = 0; b < 3; b++) rb[b + 3 * c] = (unsigned)fg.nx * ((unsigned)ys[b] + (unsigned)fg.ny * (unsigned)zs[c]);
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) {
    const float* rk = res + (size_t)k * (size_t)fg.nn;
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        float row = 0.f;
#pragma unroll
        for (int a = 0; a < 3; a++) row += w1[a] * ld_off(rk, 4u * (rb[b + 3 * c] + (unsigned)xs[a]));
        s += w1[b] * w1[c] * row;
      }
    fc[k * cg.nn + t] = s;
  }
}

// Restriction with 8 lanes per coarse node for small coarse levels (the
// same weights; lane j takes the fine offsets o = j + 8q, the group adds
// with group8_sum, lanes 0-2 write). Thread t serves coarse node t / 8.
constexpr long long kRestrict8MaxNodes = 4096;
__host__ __device__ inline bool restrict8(const Geom& cg) { return cg.nn <= kRestrict8MaxNodes; }
__device__ inline void restrict8_body(const Geom& fg, const Geom& cg, const float* __restrict__ res,
                                      float* __restrict__ fc, long long t) {
  const long long node = t >> 3;
  const int j = (int)(t & 7);
  if (node >= cg.nn) return;
  int X, Y, Z;
  split_node(cg.nx, cg.ny, node, X, Y, Z);
  const int x = 2 * X, y = 2 * Y, z = 2 * Z;
  float a[OTM_NRHS] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int o = j + 8 * q;
    if (o < 27) {
      const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
      const float w = (dx ? 0.5f : 1.f) * (dy ? 0.5f : 1.f) * (dz ? 0.5f : 1.f);
      const unsigned off = 4u * ((unsigned)wrapm(x + dx, fg.nx) +
                                 (unsigned)fg.nx * ((unsigned)wrapm(y + dy, fg.ny) + (unsigned)fg.ny * (unsigned)wrapm(z + dz, fg.nz)));
#pragma unroll
      for (int k = 0; k < OTM_NRHS; k++) a[k] += w * ld_off(res + (size_t)k * (size_t)fg.nn, off);
    }
  }
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) a[k] = group8_sum(a[k]);
  if (j < OTM_NRHS) fc[(size_t)j * (size_t)cg.nn + node] = j == 0 ? a[0] : (j == 1 ? a[1] : a[2]);
}

// u_f(i) += sum_I P(i, I) u_c(I), trilinear (prolongate_correction_kernel_H,
// grid.cu:881-940).
__device__ inline void prolong_node(const Geom& fg, const Geom& cg, const float* uc, float* uf, long long t) {
  int x, y, z;
  split_node(fg.nx, fg.ny, t, x, y, z);
  const ProlongW P = prolong_weights(cg, x, y, z);
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) uf[k * fg.nn + t] += prolong_sum(cg, uc + k * cg.nn, P);
}

// ---- per-level bodies ----
// GS update of one of 8 colours (parity of x, y, z), colour = cx + 2cy + 4cz.
template <bool HALF>
__device__ inline void st_gs_color_body(const Geom& g, const StRef& S, const float* __restrict__ f, float* u, int cx,
                                        int cy, int cz, long long t) {
  const int hx = g.nx >> 1, hy = g.ny >> 1, hz = g.nz >> 1;
  const long long cnt = (long long)hx * hy * hz;
  if (t >= cnt) return;
  int x, y, z;
  split_node(hx, hy, t, x, y, z);
  x = 2 * x + cx;
  y = 2 * y + cy;
  z = 2 * z + cz;
  st_gs_node<HALF>(g, S, f, u, x, y, z);
}

// Levels up to this many nodes take the 8-lanes-per-node residual (latency
// bound: 10 -> 3.4-5.4 us at 16^3 and 32^3); larger ones one node per thread
// (at 64^3 the 8-lane form's stencil reads coalesce worse: 12.7 -> 26 us).
constexpr long long kResid8MaxNodes = 65536;
__host__ __device__ inline bool st_resid8(const Geom& g, bool half) { return !half && g.nn <= kResid8MaxNodes; }

// res = f - A u with 8 lanes per node (FP32 stencils): lane j of a group
// takes the entries o = j, j + 8, j + 16, j + 24 (o < 27) with their three
// load cases' neighbours, the group adds its partial sums (group8_sum) and
// lanes 0-2 write. Thread t serves node t / 8 (a whole group is past the
// level when its node is). One node per thread left small levels waiting
// on its 111 loads one after another (10 us at 16^3 and 32^3).
__device__ inline void st_residual8_body(const Geom& g, const float* __restrict__ st, const float* __restrict__ f,
                                         const float* __restrict__ u, float* __restrict__ res, long long t) {
  const long long node = t >> 3;
  const int j = (int)(t & 7);
  if (node >= g.nn) return;
  int x, y, z;
  split_node(g.nx, g.ny, node, x, y, z);
  const unsigned nx = (unsigned)g.nx, nxy = (unsigned)g.nx * (unsigned)g.ny;
  float a[OTM_NRHS] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int o = j + 8 * q;
    if (o < 27) {
      const float c = ld_off(st + (size_t)o * (size_t)g.nn, 4u * (unsigned)node);
      const unsigned w = 4u * ((unsigned)wrapm(x + o % 3 - 1, g.nx) + nx * (unsigned)wrapm(y + (o / 3) % 3 - 1, g.ny) +
                               nxy * (unsigned)wrapm(z + o / 9 - 1, g.nz));
#pragma unroll
      for (int k = 0; k < OTM_NRHS; k++) a[k] += c * ld_off(u + (size_t)k * (size_t)g.nn, w);
    }
  }
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) a[k] = group8_sum(a[k]);
  if (j < OTM_NRHS) {
    const float s = j == 0 ? a[0] : (j == 1 ? a[1] : a[2]);
    res[(size_t)j * (size_t)g.nn + node] = f[(size_t)j * (size_t)g.nn + node] - s;
  }
}

template <bool HALF>
__device__ inline void st_residual_body(const Geom& g, const StRef& S, const float* __restrict__ f,
                                        const float* __restrict__ u, float* __restrict__ res, long long t) {
  if (t >= g.nn) return;
  st_resid_node<HALF>(g, S, f, u, res, t);
}

// restrict_node when coarse rows are whole waves (cg.nx a multiple of 64,
// t = this wave's first coarse node + lane): a lane loads the fine pair
// (2X, 2X+1) of each of its nine rows as one 8-byte load (a wave reads 512
// contiguous bytes) and takes 2X-1 from the lane before it by DPP; lane 0
// loads it. The same terms in the same order: the same bits.
__device__ inline void restrict_node_pairs(const Geom& fg, const Geom& cg, const float* res, float* fc, long long t) {
  int X, Y, Z;
  split_node(cg.nx, cg.ny, t, X, Y, Z);
  const int x = 2 * X, y = 2 * Y, z = 2 * Z;
  const int xm = x == 0 ? fg.nx - 1 : x - 1;
  const int ys[3] = {y == 0 ? fg.ny - 1 : y - 1, y, y + 1};
  const int zs[3] = {z == 0 ? fg.nz - 1 : z - 1, z, z + 1};
  const float w1[3] = {0.5f, 1.f, 0.5f};
  const bool lane0 = (threadIdx.x & 63) == 0;
  unsigned rb[9];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int b = 0; b < 3; b++) rb[b + 3 * c] = (unsigned)fg.nx * ((unsigned)ys[b] + (unsigned)fg.ny * (unsigned)zs[c]);
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) {
    const float* rk = res + (size_t)k * (size_t)fg.nn;
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        const unsigned r = rb[b + 3 * c];
        const float2 v = *(const float2*)((const char*)rk + 4u * (r + (unsigned)x));
        float left = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v.y), 0x138, 0xF, 0xF, false));
        if (lane0) left = ld_off(rk, 4u * (r + (unsigned)xm));
        float row = 0.f;
        row += w1[0] * left;
        row += w1[1] * v.x;
        row += w1[2] * v.y;
        s += w1[b] * w1[c] * row;
      }
    fc[k * cg.nn + t] = s;
  }
}

// One thread per coarse node (t); the grid covers whole waves, so the
// paired form may be taken whenever coarse rows are whole waves.
__device__ inline void restrict_body(const Geom& fg, const Geom& cg, const float* __restrict__ res,
                                     float* __restrict__ fc, long long t) {
  if (cg.nx % 64 == 0 && blockDim.x % 64 == 0) {
    if (t < cg.nn) restrict_node_pairs(fg, cg, res, fc, t);
    return;
  }
  if (t >= cg.nn) return;
  restrict_node(fg, cg, res, fc, t);
}

// Prolongation by coarse cell: thread t owns coarse node (X, Y, Z) and the
// eight fine nodes (2X+dx, 2Y+dy, 2Z+dz). It loads the eight coarse values
// around the cell once per right-hand side and adds to every fine node the
// terms of prolong_sum in its order (zero weights skipped), so the result
// is prolong_node's bit for bit with an eighth of the coarse loads and
// paired, coalesced fine accesses. The fine level is twice the coarse one
// along every axis.
__device__ inline void prolong_body(const Geom& fg, const Geom& cg, const float* __restrict__ uc,
                                    float* __restrict__ uf, long long t) {
  if (t >= cg.nn) return;
  int X, Y, Z;
  split_node(cg.nx, cg.ny, t, X, Y, Z);
  const int Xs[2] = {X, X + 1 == cg.nx ? 0 : X + 1};
  const int Ys[2] = {Y, Y + 1 == cg.ny ? 0 : Y + 1};
  const int Zs[2] = {Z, Z + 1 == cg.nz ? 0 : Z + 1};
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) {
    const float* ck = uc + k * cg.nn;
    float v[2][2][2];
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int a = 0; a < 2; a++)
          v[c][b][a] = ck[(long long)Xs[a] + (long long)cg.nx * ((long long)Ys[b] + (long long)cg.ny * Zs[c])];
    float* fk = uf + k * fg.nn;
#pragma unroll
    for (int dz = 0; dz < 2; dz++)
#pragma unroll
      for (int dy = 0; dy < 2; dy++)
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
          const ProlongW P = prolong_weights(cg, 2 * X + dx, 2 * Y + dy, 2 * Z + dz);
          float s = 0.f;
#pragma unroll
          for (int c = 0; c < 2; c++)
#pragma unroll
            for (int b = 0; b < 2; b++)
#pragma unroll
              for (int a = 0; a < 2; a++) {
                const float w = P.wx[a] * P.wy[b] * P.wz[c];
                if (w != 0.f) s += w * v[c & dz][b & dy][a & dx];
              }
          const long long fi = (long long)(2 * X + dx) +
                               (long long)fg.nx * ((long long)(2 * Y + dy) + (long long)fg.ny * (2 * Z + dz));
          fk[fi] += s;
        }
  }
}

// u = Binv f for right-hand side k, row i, FP64 accumulation.
__device__ inline void coarsest_solve_body(const Geom& g, const double* __restrict__ Binv,
                                           const float* __restrict__ f, float* __restrict__ u, int i, int k) {
  const int n = (int)g.nn;
  if (i >= n) return;
  const float* fk = f + (long long)k * n;
  double s = 0.0;
  for (int j = 0; j < n; j++) s += Binv[(long long)i * n + j] * (double)fk[j];
  u[(long long)k * n + i] = (float)s;
}

// Brick smoother for stored-stencil levels. One 512-thread workgroup owns an
// 8^3 brick (cut at the level's edge when an edge is not a multiple of 8):
// its nodes' stencils sit in registers, u of the brick plus a one-node halo
// sits in LDS. All colours of `sweeps` Gauss-Seidel sweeps run inside the
// brick with workgroup barriers; halo values stay frozen at their values on
// entry (block Gauss-Seidel across bricks, exact multicolour Gauss-Seidel
// inside; a level smaller than a brick is smoothed exactly except across
// its periodic seam). The result goes to a second buffer so bricks never
// read a neighbour's new values. The post-smoother runs the colours in
// reverse order, which is the adjoint of the pre-smoother: the V-cycle
// stays symmetric.
constexpr int kBrick = 8;
constexpr int kBrickThreads = kBrick * kBrick * kBrick;
constexpr int kHalo = kBrick + 2;
__host__ __device__ inline int brick_count(const Geom& g) {
  return ((g.nx + kBrick - 1) / kBrick) * ((g.ny + kBrick - 1) / kBrick) * ((g.nz + kBrick - 1) / kBrick);
}
__device__ inline int wrap_any(int a, int n) {
  a %= n;
  return a < 0 ? a + n : a;
}
// Coupling o of the stored stencil at node v (FP16: scaled back to FP32).
template <bool HALF>
__device__ inline float stencil_coef(const StRef& S, long long nn, long long v, int o) {
  if constexpr (HALF)
    return (float)S.f16[o * nn + v] * *S.scale;
  else
    return S.f32[o * nn + v];
}

// Work split of a colour step: the 64 nodes of one colour in a brick are
// updated by 64 groups of 8 lanes. Lane j of a group holds the couplings
// o = j, j + 8, j + 16, j + 24 (o != 13, o < 27: 2 to 4 of the 26) of its
// node for every colour, forms their products with the three load cases'
// values, and the group adds its eight partial sums (group8_sum); lanes
// 0-2 then write load cases 0-2. With one node per thread
// only 16 lanes of half the waves worked in a colour step, each through 78
// dependent multiply-adds.
// LDS of the brick smoother, one allocation for both stencil precisions
// (a batched launch instantiates both bodies in one kernel).
struct StBrickLds {
  float su[OTM_NRHS][kHalo * kHalo * kHalo];  // u of the brick and its halo
  float scf[kBrickThreads * 27];              // stencils, colour-major node order
};
__device__ inline StBrickLds& st_brick_lds() {
  __shared__ StBrickLds lds;
  return lds;
}

template <bool HALF>
__device__ inline void st_brick_body(const Geom& g, const StRef& S, const float* __restrict__ f,
                                     const float* __restrict__ uin, float* __restrict__ uout, int sweeps, int forward,
                                     int zero_in, int b) {
  auto& su = st_brick_lds().su;
  float* scf = st_brick_lds().scf;
  const int bx = (g.nx + kBrick - 1) / kBrick, by = (g.ny + kBrick - 1) / kBrick;
  const int ox = (b % bx) * kBrick, oy = ((b / bx) % by) * kBrick, oz = (b / (bx * by)) * kBrick;
  const int t = threadIdx.x;
  const long long nn = g.nn;
  // halo load (or zero initial guess): loads first, then the LDS stores
  {
    constexpr int kH3 = kHalo * kHalo * kHalo;
    constexpr int kTrips = (kH3 + kBrickThreads - 1) / kBrickThreads;
    float hv[kTrips][OTM_NRHS];
#pragma unroll
    for (int r = 0; r < kTrips; r++) {
      const int i = t + kBrickThreads * r;
#pragma unroll
      for (int k = 0; k < OTM_NRHS; k++) hv[r][k] = 0.f;
      if (i < kH3 && !zero_in) {
        const int hx = i % kHalo, hy = (i / kHalo) % kHalo, hz = i / (kHalo * kHalo);
        const int X = wrap_any(ox + hx - 1, g.nx), Y = wrap_any(oy + hy - 1, g.ny), Z = wrap_any(oz + hz - 1, g.nz);
        const long long w = (long long)X + (long long)g.nx * ((long long)Y + (long long)g.ny * Z);
#pragma unroll
        for (int k = 0; k < OTM_NRHS; k++) hv[r][k] = uin[k * nn + w];
      }
    }
#pragma unroll
    for (int r = 0; r < kTrips; r++) {
      const int i = t + kBrickThreads * r;
      if (i < kH3) {
#pragma unroll
        for (int k = 0; k < OTM_NRHS; k++) su[k][i] = hv[r][k];
      }
    }
  }
  const int m = t >> 3, j = t & 7;
  int doff[4];
  bool slot[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int o = j + 8 * q;
    slot[q] = o < 27 && o != 13;
    doff[q] = slot[q] ? (o % 3 - 1) + kHalo * (((o / 3) % 3 - 1) + kHalo * (o / 9 - 1)) : 0;
  }
  // Stencils: each thread loads its own node's 27 entries (neighbouring
  // threads read neighbouring entries; load_stencil's values, FP16 diagonal
  // included) into LDS in colour-major node order, from where each lane
  // takes its slots for every colour. (Loaded straight into the slots, a
  // wave's 64 loads touched 8 stencil arrays at stride-2 nodes.)
  float cf[8][4], inv[8], fv[8];
  {
    const int lx = t & 7, ly = (t >> 3) & 7, lz = t >> 6;
    const int x = ox + lx, y = oy + ly, z = oz + lz;
    const bool act = x < g.nx && y < g.ny && z < g.nz;
    float a[27];
    if (act) {
      load_stencil<HALF>(S, nn, (long long)x + (long long)g.nx * ((long long)y + (long long)g.ny * z), a);
    } else {
#pragma unroll
      for (int o = 0; o < 27; o++) a[o] = 0.f;
    }
    const int c = (lx & 1) + 2 * (ly & 1) + 4 * (lz & 1), mm = (lx >> 1) + 4 * (ly >> 1) + 16 * (lz >> 1);
    float* d = scf + (c * 64 + mm) * 27;
#pragma unroll
    for (int o = 0; o < 27; o++) d[o] = a[o];
  }
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int x = ox + 2 * (m & 3) + (c & 1), y = oy + 2 * ((m >> 2) & 3) + ((c >> 1) & 1),
              z = oz + 2 * (m >> 4) + (c >> 2);
    const bool act = x < g.nx && y < g.ny && z < g.nz;
    const long long v = act ? (long long)x + (long long)g.nx * ((long long)y + (long long)g.ny * z) : 0;
    fv[c] = act && j < OTM_NRHS ? f[j * nn + v] : 0.f;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const float* d = scf + (c * 64 + m) * 27;
#pragma unroll
    for (int q = 0; q < 4; q++) cf[c][q] = slot[q] ? d[j + 8 * q] : 0.f;
    const float dg = d[13];
    inv[c] = dg != 0.f ? 1.f / dg : 0.f;  // zero for nodes past the level's edge
  }
  auto step = [&](const int c) {
    const int c0 = (2 * (m & 3) + (c & 1) + 1) + kHalo * ((2 * ((m >> 2) & 3) + ((c >> 1) & 1) + 1) +
                                                          kHalo * (2 * (m >> 4) + (c >> 2) + 1));
    float a[OTM_NRHS] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int p = c0 + doff[q];
#pragma unroll
      for (int k = 0; k < OTM_NRHS; k++) a[k] += cf[c][q] * su[k][p];
    }
#pragma unroll
    for (int k = 0; k < OTM_NRHS; k++) a[k] = group8_sum(a[k]);
    if (j < OTM_NRHS && inv[c] != 0.f) {
      const float acc = j == 0 ? a[0] : (j == 1 ? a[1] : a[2]);
      su[j][c0] = (fv[c] - acc) * inv[c];
    }
    __syncthreads();
  };
  for (int s = 0; s < sweeps; s++) {
    if (forward) {
#pragma unroll
      for (int c = 0; c < 8; c++) step(c);
    } else {
#pragma unroll
      for (int c = 7; c >= 0; c--) step(c);
    }
  }
  const int lx = t & 7, ly = (t >> 3) & 7, lz = t >> 6;
  const int x = ox + lx, y = oy + ly, z = oz + lz;
  if (x < g.nx && y < g.ny && z < g.nz) {
    const long long v = (long long)x + (long long)g.nx * ((long long)y + (long long)g.ny * z);
    const int c0 = (lx + 1) + kHalo * ((ly + 1) + kHalo * (lz + 1));
#pragma unroll
    for (int k = 0; k < OTM_NRHS; k++) uout[k * nn + v] = su[k][c0];
  }
}

// ---- setup bodies (Galerkin hierarchy, FP16 packing, coarsest inverse) ----
// Level-1 element matrices from kappa: E(E) = sum_c kappa_c G[c].
__device__ inline void galerkin_first_body(const Geom& fg, const Geom& cg, const GalerkinConsts& G,
                                           const float* __restrict__ kappa, float* __restrict__ E, long long t) {
  if (t >= cg.nn) return;
  int X, Y, Z;
  split_node(cg.nx, cg.ny, t, X, Y, Z);
  float kc[8];
#pragma unroll
  for (int c = 0; c < 8; c++)
    kc[c] = kappa[lin(fg, 2 * X + (c & 1), 2 * Y + ((c >> 1) & 1), 2 * Z + (c >> 2))];
#pragma unroll
  for (int q = 0; q < 36; q++) {
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < 8; c++) s += kc[c] * G.G[c][q];
    E[q * cg.nn + t] = s;
  }
}

// 1-D interpolation weight of a fine point p (0..2) from coarse corner M (0/1).
__device__ constexpr float w1d(int p, int M) {
  return (p - 2 * M == 0) ? 1.f : ((p - 2 * M == 1 || p - 2 * M == -1) ? 0.5f : 0.f);
}

// Deeper levels: Eout(E) = sum_c Q_c^T Ein(child c) Q_c.
__device__ inline void galerkin_next_body(const Geom& fg, const Geom& cg, const float* __restrict__ Ein,
                                          float* __restrict__ Eout, long long t) {
  if (t >= cg.nn) return;
  int X, Y, Z;
  split_node(cg.nx, cg.ny, t, X, Y, Z);
  float acc[36];
#pragma unroll
  for (int q = 0; q < 36; q++) acc[q] = 0.f;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
    const long long child = lin(fg, 2 * X + cx, 2 * Y + cy, 2 * Z + cz);
    float e[36];
#pragma unroll
    for (int q = 0; q < 36; q++) e[q] = Ein[q * fg.nn + child];
    // T = E Q  (8 x 8)
    float T[8][8];
#pragma unroll
    for (int m = 0; m < 8; m++)
#pragma unroll
      for (int b = 0; b < 8; b++) {
        float s = 0.f;
#pragma unroll
        for (int n = 0; n < 8; n++) {
          const float qnb = w1d(cx + (n & 1), b & 1) * w1d(cy + ((n >> 1) & 1), (b >> 1) & 1) *
                            w1d(cz + (n >> 2), b >> 2);
          if (qnb != 0.f) s += e[tri8(m, n)] * qnb;
        }
        T[m][b] = s;
      }
    // acc(a,b) += sum_m Q(m,a) T(m,b)
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
      for (int b = a; b < 8; b++) {
        float s = 0.f;
#pragma unroll
        for (int m = 0; m < 8; m++) {
          const float qma = w1d(cx + (m & 1), a & 1) * w1d(cy + ((m >> 1) & 1), (a >> 1) & 1) *
                            w1d(cz + (m >> 2), a >> 2);
          if (qma != 0.f) s += qma * T[m][b];
        }
        acc[tri8(a, b)] += s;
      }
  }
#pragma unroll
  for (int q = 0; q < 36; q++) Eout[q * cg.nn + t] = acc[q];
}

// Node stencil from element matrices; the diagonal is reset to minus the sum
// of the off-diagonal couplings so that constants stay in the null space.
__device__ inline void stencil_from_elements_body(const Geom& g, const float* __restrict__ E, float* __restrict__ st,
                                                  long long t) {
  if (t >= g.nn) return;
  int x, y, z;
  split_node(g.nx, g.ny, t, x, y, z);
  float a[27];
#pragma unroll
  for (int o = 0; o < 27; o++) a[o] = 0.f;
#pragma unroll
  for (int e = 0; e < 8; e++) {
    const int ax = e & 1, ay = (e >> 1) & 1, az = e >> 2;
    const int ex = wrapm(x - 1 + ax, g.nx), ey = wrapm(y - 1 + ay, g.ny), ez = wrapm(z - 1 + az, g.nz);
    const long long eid = lin(g, ex, ey, ez);
    const int lv = (1 - ax) + 2 * (1 - ay) + 4 * (1 - az);
#pragma unroll
    for (int m = 0; m < 8; m++) {
      const int o = (ax + (m & 1)) + 3 * (ay + ((m >> 1) & 1)) + 9 * (az + (m >> 2));
      a[o] += E[tri8(lv, m) * g.nn + eid];
    }
  }
  double off = 0.0;
#pragma unroll
  for (int o = 0; o < 27; o++)
    if (o != 13) off += (double)a[o];
  a[13] = (float)(-off);
#pragma unroll
  for (int o = 0; o < 27; o++) st[o * g.nn + t] = a[o];
}

// Largest |coupling| of a level (as float bits: non-negative floats order
// like their bit patterns), for the FP16 scale; block lb of nb.
__device__ inline void stencil_absmax_body(const Geom& g, const float* __restrict__ st, unsigned int* __restrict__ bits,
                                           int lb, int nb) {
  float m = 0.f;
  for (long long v = (long long)lb * blockDim.x + threadIdx.x; v < g.nn; v += (long long)nb * blockDim.x)
#pragma unroll
    for (int o = 0; o < 27; o++)
      if (o != 13) m = fmaxf(m, fabsf(st[o * g.nn + v]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_down(m, o, 64));
  __shared__ float wm[kWaves];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kWaves; w++) m = fmaxf(m, wm[w]);
    atomicMax(bits, __float_as_uint(m));
  }
}

// FP16 copy of a level's couplings divided by the level scale. Coupling
// (v, v+d) for d before the centre in offset order is taken from the
// neighbour's entry (v+d, v), so the stored operator is exactly symmetric.
__device__ inline void stencil_pack_body(const Geom& g, const float* __restrict__ st, const float* __restrict__ scale,
                                         _Float16* __restrict__ st16, long long t) {
  if (t >= g.nn) return;
  int x, y, z;
  split_node(g.nx, g.ny, t, x, y, z);
  const float sc = *scale;
  const float inv = sc > 0.f ? 1.f / sc : 0.f;
#pragma unroll
  for (int o = 0; o < 27; o++) {
    float val = 0.f;
    if (o > 13) {
      val = st[o * g.nn + t];
    } else if (o < 13) {
      const long long w = lin(g, wrapm(x + o % 3 - 1, g.nx), wrapm(y + (o / 3) % 3 - 1, g.ny),
                              wrapm(z + o / 9 - 1, g.nz));
      val = st[(26 - o) * g.nn + w];
    }
    st16[o * g.nn + t] = (_Float16)(val * inv);
  }
}

// Coarsest level: dense FP64 matrix from the stencil, symmetrised, zero row
// sums, plus (tr/n^2) * 1 1^T to remove the constant null space (the
// reference solves its coarsest system on the host with Eigen and a
// null-space projection, grid.cpp:272-310, :432-510). The result is its
// explicit inverse Binv, so every V-cycle's coarsest solve is one small
// dense matrix-vector product on the device.
constexpr int kFactorThreads = 1024;
constexpr int kInvertLdsMax = 128;  // n*n doubles in LDS up to n = 128 (128 KiB)

__device__ inline void coarsest_assemble(const Geom& g, const float* __restrict__ st, double* B, double* red,
                                         double* alpha_sh) {
  const int n = (int)g.nn;
  const int tid = threadIdx.x, bs = blockDim.x;
  for (int i = tid; i < n * n; i += bs) B[i] = 0.0;
  __syncthreads();
  for (int v = tid; v < n; v += bs) {
    const int x = v % g.nx, y = (v / g.nx) % g.ny, z = v / (g.nx * g.ny);
    for (int o = 0; o < 27; o++) {
      const int w = wrapm(x + o % 3 - 1, g.nx) + g.nx * (wrapm(y + (o / 3) % 3 - 1, g.ny) +
                                                         g.ny * wrapm(z + o / 9 - 1, g.nz));
      B[v * n + w] += (double)st[o * n + v];
    }
  }
  __syncthreads();
  for (int i = tid; i < n * n; i += bs) {
    const int a = i / n, b = i - a * n;
    if (a < b) {
      const double m = 0.5 * (B[a * n + b] + B[b * n + a]);
      B[a * n + b] = m;
      B[b * n + a] = m;
    }
  }
  __syncthreads();
  double tr = 0.0;
  for (int a = tid; a < n; a += bs) {
    double s = 0.0;
    for (int b = 0; b < n; b++)
      if (b != a) s += B[a * n + b];
    B[a * n + a] = -s;
    tr += -s;
  }
  tr = wave_sum(tr);
  if ((tid & 63) == 0) red[tid >> 6] = tr;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int w = 0; w < (bs + 63) / 64; w++) s += red[w];
    *alpha_sh = s / ((double)n * (double)n);
  }
  __syncthreads();
  const double alpha = *alpha_sh;
  for (int i = tid; i < n * n; i += bs) B[i] += alpha;
  __syncthreads();
}

// In-place Gauss-Jordan inversion of the SPD matrix A (n x n, LDS or global
// memory): no pivoting is needed; every elimination step updates all n^2
// entries in parallel between two barriers (2n barriers in total instead of
// the long dependent chains of a triangular inverse). colk/rowk hold n
// doubles each. Binv gets the symmetrised inverse.
__device__ inline void gauss_jordan_invert(int n, double* A, double* colk, double* rowk, double* __restrict__ Binv) {
  const int tid = threadIdx.x, bs = blockDim.x;
  for (int k = 0; k < n; k++) {
    const double ip = 1.0 / A[k * n + k];
    for (int i = tid; i < n; i += bs) {
      colk[i] = A[i * n + k];
      rowk[i] = (i == k ? 1.0 : A[k * n + i]) * ip;
    }
    __syncthreads();
    if (n <= bs) {
      const int rstep = bs / n;  // rows updated per pass
      const int j = tid % n, i0 = tid / n;
      if (i0 < rstep)
        for (int i = i0; i < n; i += rstep) {
          const int q = i * n + j;
          A[q] = (i == k) ? rowk[j] : ((j == k ? 0.0 : A[q]) - colk[i] * rowk[j]);
        }
    } else {
      for (long long q = tid; q < (long long)n * n; q += bs) {
        const int i = (int)(q / n), j = (int)(q - (long long)i * n);
        A[q] = (i == k) ? rowk[j] : ((j == k ? 0.0 : A[q]) - colk[i] * rowk[j]);
      }
    }
    __syncthreads();
  }
  for (long long q = tid; q < (long long)n * n; q += bs) {
    const int a = (int)(q / n), b = (int)(q - (long long)a * n);
    Binv[q] = 0.5 * (A[(long long)a * n + b] + A[(long long)b * n + a]);
  }
}

// n <= kInvertLdsMax: the work matrix lives in (dynamic) LDS.
__device__ inline void coarsest_invert_lds_body(const Geom& g, const float* __restrict__ st, double* A,
                                                double* __restrict__ Binv) {
  __shared__ double colk[kInvertLdsMax], rowk[kInvertLdsMax];
  __shared__ double red[kFactorThreads / 64];
  __shared__ double alpha_sh;
  coarsest_assemble(g, st, A, red, &alpha_sh);
  gauss_jordan_invert((int)g.nn, A, colk, rowk, Binv);
}



// ---- one-workgroup tail V-cycle ---------------------------------------------
// The last levels of a hierarchy (from the first coarse level with at most
// kTailStartNodes nodes down to the coarsest) hold a few hundred nodes: as
// separate launches their smoothers, residuals and transfers are pure
// launch and memory latency. One workgroup runs the whole tail V-cycle
// instead. The stencils of its non-coarsest levels are staged into LDS once
// (kTailStNodes nodes at most) and the corrections of every tail level stay
// in LDS; smoothing is exact 8-colour Gauss-Seidel (forward before
// restriction, reverse after prolongation, so the V-cycle stays symmetric),
// residuals and restricted right-hand sides go through the level buffers in
// global memory (visible to the workgroup after each barrier), and the
// coarsest level is solved with its explicit inverse. A hierarchy whose
// coarse levels are all larger keeps only the coarsest level in the tail
// (up to kTailLdsNodes nodes).
constexpr int kTailMaxLevels = 6;
constexpr int kTailStartNodes = 512;                 // largest first tail level (8^3)
constexpr int kTailStNodes = 512 + 64 + 8 + 1;       // stencil nodes held in LDS
constexpr int kTailThreads = 512;  // (1024 would cap registers at 128 and spill)
constexpr int kTailLdsNodes = 4096;                  // correction nodes held in LDS
constexpr int kTailBinvNodes = 64;                   // coarsest inverse held in LDS up to this size
static_assert(kTailStartNodes / 8 <= kTailThreads / 8, "a tail colour step has one node per group of 8 lanes");

struct TailLevels {
  int n;  // levels in the tail; the last is the coarsest
  int sweeps;
  Geom g[kTailMaxLevels];
  const float* f0;                // right-hand side of the first tail level
  float* f[kTailMaxLevels];       // right-hand sides of the deeper levels (written here)
  float* res[kTailMaxLevels];     // residual scratch of the non-coarsest levels
  StRef st[kTailMaxLevels];
  const double* Binv;             // inverse of the coarsest level
  float* uout;                    // correction of the first tail level (output)
};

void launch_tail_vcycle(hipStream_t s, const TailLevels& T);

__device__ inline void tail_stencil(const StRef& S, long long nn, long long v, float a[27]) {
  if (S.f16)
    load_stencil<true>(S, nn, v, a);
  else
    load_stencil<false>(S, nn, v, a);
}

// `sweeps` exact 8-colour Gauss-Seidel sweeps on u of one tail level with
// its stencil sa[o*nn + v] and right-hand side f (all in LDS). A node of
// the active colour is updated by a group of 8 lanes as in the coarse brick
// smoother (lane j takes the couplings o = j + 8q, the group adds with
// group8_sum, lanes 0-2 write the three load cases). A tail level has at
// most kTailStartNodes nodes, so each group owns one node position per
// colour: its coordinates are decoded once per call, not per colour step
// (the runtime divisions and the FP32 division per step made a step 2700
// cycles long).
template <bool FORWARD>
__device__ inline void tail_smooth(const Geom& g, const float* sa, const float* f, float* u, int sweeps) {
  const int hx = g.nx >> 1, hy = g.ny >> 1, hz = g.nz >> 1;
  const int cnt = hx * hy * hz;  // nodes per colour, at most blockDim.x / 8
  const int nn = (int)g.nn;
  const int grp = (int)threadIdx.x >> 3, j = (int)threadIdx.x & 7;
  const bool act = grp < cnt;
  const int tx = act ? grp % hx : 0, tr = act ? grp / hx : 0;
  const int ty = tr % hy, tz = tr / hy;
  for (int s = 0; s < sweeps; s++)
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int col = FORWARD ? i : 7 - i;
      if (act) {
        const int x = 2 * tx + (col & 1), y = 2 * ty + ((col >> 1) & 1), z = 2 * tz + (col >> 2);
        const int v = x + g.nx * (y + g.ny * z);
        float a[OTM_NRHS] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int o = j + 8 * q;
          if (o < 27 && o != 13) {
            const int X = wrapm(x + o % 3 - 1, g.nx), Y = wrapm(y + (o / 3) % 3 - 1, g.ny),
                      Z = wrapm(z + o / 9 - 1, g.nz);
            const int w = X + g.nx * (Y + g.ny * Z);
            const float c = sa[o * nn + v];
#pragma unroll
            for (int k = 0; k < OTM_NRHS; k++) a[k] += c * u[k * nn + w];
          }
        }
#pragma unroll
        for (int k = 0; k < OTM_NRHS; k++) a[k] = group8_sum(a[k]);
        if (j < OTM_NRHS) {
          const float acc = j == 0 ? a[0] : (j == 1 ? a[1] : a[2]);
          u[j * nn + v] = (f[j * nn + v] - acc) * __builtin_amdgcn_rcpf(sa[13 * nn + v]);
        }
      }
      __syncthreads();
    }
}

// res = f - A u at node t of a tail level (st_resid_node's arithmetic with
// the stencil from LDS).
__device__ inline void tail_resid_node(const Geom& g, const float* sa, const float* __restrict__ f, const float* u,
                                       float* res, int t) {
  const int nn = (int)g.nn;
  int x, y, z;
  split_node(g.nx, g.ny, t, x, y, z);
  const int xs[3] = {x == 0 ? g.nx - 1 : x - 1, x, x == g.nx - 1 ? 0 : x + 1};
  const int ys[3] = {y == 0 ? g.ny - 1 : y - 1, y, y == g.ny - 1 ? 0 : y + 1};
  const int zs[3] = {z == 0 ? g.nz - 1 : z - 1, z, z == g.nz - 1 ? 0 : z + 1};
  float a[27];
#pragma unroll
  for (int o = 0; o < 27; o++) a[o] = sa[o * nn + t];
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) {
    const float* uk = u + k * nn;
    float s = 0.f;
#pragma unroll
    for (int o = 0; o < 27; o++) s += a[o] * uk[xs[o % 3] + g.nx * (ys[(o / 3) % 3] + g.ny * zs[o / 9])];
    res[k * nn + t] = f[k * nn + t] - s;
  }
}

// Offsets of tail level l in the LDS corrections and in the LDS stencils
// (every level before it stores OTM_NRHS, resp. 27, values per node).
__device__ inline int tail_nodes_before(const TailLevels& T, int l) {
  int o = 0;
  for (int m = 0; m < l; m++) o += (int)T.g[m].nn;
  return o;
}

__device__ inline void tail_vcycle_body(const TailLevels& T) {
  __shared__ float su[OTM_NRHS * kTailLdsNodes];
  __shared__ float sst[27 * kTailStNodes];
  __shared__ float sf[OTM_NRHS * kTailStNodes];        // right-hand sides of the tail levels
  __shared__ float sres[OTM_NRHS * kTailStartNodes];   // residual of the level being restricted
  __shared__ double sbinv[kTailBinvNodes * kTailBinvNodes];
  const int tid = threadIdx.x, bs = blockDim.x;
  const int L = T.n - 1;
  const int nc = (int)T.g[L].nn;
  auto off = [&](int l) { return OTM_NRHS * tail_nodes_before(T, l); };
  auto soff = [&](int l) { return 27 * tail_nodes_before(T, l); };
  // Stencils of the non-coarsest levels (the values load_stencil gives), the
  // first level's right-hand side and a small coarsest inverse: everything
  // the cycle reads more than once comes from LDS. (With the right-hand
  // sides, residuals and inverse in global memory, every colour step and
  // the coarsest product waited for L2: 33 us per cycle at 128^3.)
  for (int l = 0; l < L; l++) {
    const int nn = (int)T.g[l].nn;
    for (int v = tid; v < nn; v += bs) {
      float a[27];
      tail_stencil(T.st[l], nn, v, a);
#pragma unroll
      for (int o = 0; o < 27; o++) sst[soff(l) + o * nn + v] = a[o];
    }
  }
  if (L > 0)
    for (int i = tid; i < OTM_NRHS * (int)T.g[0].nn; i += bs) sf[i] = T.f0[i];
  const bool binv_lds = nc <= kTailBinvNodes;
  if (binv_lds) {
    // all loads in flight before the stores (a fixed trip count)
    constexpr int kTrips = (kTailBinvNodes * kTailBinvNodes + kTailThreads - 1) / kTailThreads;
    double bv[kTrips];
#pragma unroll
    for (int r = 0; r < kTrips; r++) {
      const int i = tid + kTailThreads * r;
      bv[r] = i < nc * nc ? T.Binv[i] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < kTrips; r++) {
      const int i = tid + kTailThreads * r;
      if (i < nc * nc) sbinv[i] = bv[r];
    }
  }
  // down: smooth from zero, residual, restriction (into the next level's
  // right-hand side in LDS)
  for (int l = 0; l < L; l++) {
    const Geom& g = T.g[l];
    float* u = su + off(l);
    const float* f = sf + off(l);
    for (int i = tid; i < OTM_NRHS * (int)g.nn; i += bs) u[i] = 0.f;
    __syncthreads();
    tail_smooth<true>(g, sst + soff(l), f, u, T.sweeps);
    for (int t = tid; t < (int)g.nn; t += bs) tail_resid_node(g, sst + soff(l), f, u, sres, t);
    __syncthreads();
    for (int t = tid; t < (int)T.g[l + 1].nn; t += bs) restrict_node(g, T.g[l + 1], sres, sf + off(l + 1), t);
    __syncthreads();
  }
  // coarsest: u = Binv f (FP64 accumulation). From LDS, row i is read as
  // column i (Binv is symmetric): neighbouring lanes read neighbouring
  // entries, no bank conflicts.
  {
    const float* f = L == 0 ? T.f0 : sf + off(L);
    for (int q = tid; q < OTM_NRHS * nc; q += bs) {
      const int k = q / nc, i = q - k * nc;
      double s = 0.0;
      if (binv_lds) {
        for (int j = 0; j < nc; j++) s += sbinv[j * nc + i] * (double)f[k * nc + j];
      } else {
        for (int j = 0; j < nc; j++) s += T.Binv[(long long)i * nc + j] * (double)f[k * nc + j];
      }
      su[off(L) + q] = (float)s;
    }
    __syncthreads();
  }
  // up: prolongation, reverse smoothing
  for (int l = L - 1; l >= 0; l--) {
    const Geom& g = T.g[l];
    float* u = su + off(l);
    for (int t = tid; t < (int)g.nn; t += bs) prolong_node(g, T.g[l + 1], su + off(l + 1), u, t);
    __syncthreads();
    tail_smooth<false>(g, sst + soff(l), sf + off(l), u, T.sweeps);
  }
  for (int i = tid; i < OTM_NRHS * (int)T.g[0].nn; i += bs) T.uout[i] = su[i];
}

}  // namespace otm
