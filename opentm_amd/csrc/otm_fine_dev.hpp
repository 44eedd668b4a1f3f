// otm_fine_dev.hpp -- device bodies of the matrix-free fine-level kernels
// (gfx950, wave64), shared by the single-cell launches (otm_fine.hip) and
// the batched multi-cell launches (otm_batch.hip). Bodies take the thread's
// item index t, or for reductions the block's index lb among the nb blocks
// that work on the cell.
//
// The fine operator is the SIMP-weighted trilinear conductivity stencil
// assembled on the fly from per-element kappa (reference: grid.cu:318-471
// smoother, grid.cu:1027-1164 residual, grid.cu:1282-1342 load vectors).
// With the element matrix Ke = c*[4 diag, 0 one-axis neighbours, -1 else],
// node v couples only to its 12 edge and 8 corner neighbours:
//   (K u)_v = c * sum_{e ni v} kappa_e * (4 u_v - T_e(u)),
// where T_e sums u over the three face-diagonal and the body-diagonal nodes
// of v inside element e. Because one-axis neighbours do not couple, the
// reference's 8-colour Gauss-Seidel (colour = parity of x,y,z, in order
// 0..7) is exactly a 4-colour sweep over whole x-rows with colour
// (y%2) + 2*(z%2): colours 2k and 2k+1 of the 8-colour order never couple.
#pragma once
#include "otm_kernels.hpp"

namespace otm {


// kappa of the 8 elements around node (x,y,z) and their sum.
struct NodeElems {
  float k[8];  // index ax + 2 ay + 4 az: element at (x-1+ax, y-1+ay, z-1+az)
  float ksum;
};

struct NodeIdx {
  int x, y, z, xm, xp, ym, yp, zm, zp;
  long long v;
};

__device__ inline NodeIdx node_idx(const Geom& g, int x, int y, int z) {
  NodeIdx n;
  n.x = x; n.y = y; n.z = z;
  n.xm = x == 0 ? g.nx - 1 : x - 1;
  n.xp = x == g.nx - 1 ? 0 : x + 1;
  n.ym = y == 0 ? g.ny - 1 : y - 1;
  n.yp = y == g.ny - 1 ? 0 : y + 1;
  n.zm = z == 0 ? g.nz - 1 : z - 1;
  n.zp = z == g.nz - 1 ? 0 : z + 1;
  n.v = (long long)x + (long long)g.nx * ((long long)y + (long long)g.ny * z);
  return n;
}

__device__ inline long long rowb(const Geom& g, int y, int z) {
  return (long long)g.nx * ((long long)y + (long long)g.ny * z);
}

__device__ inline NodeElems load_elems(const Geom& g, const NodeIdx& n, const float* __restrict__ kappa) {
  NodeElems E;
  const long long r00 = rowb(g, n.ym, n.zm), r10 = rowb(g, n.y, n.zm);
  const long long r01 = rowb(g, n.ym, n.z), r11 = rowb(g, n.y, n.z);
  E.k[0] = kappa[r00 + n.xm]; E.k[1] = kappa[r00 + n.x];
  E.k[2] = kappa[r10 + n.xm]; E.k[3] = kappa[r10 + n.x];
  E.k[4] = kappa[r01 + n.xm]; E.k[5] = kappa[r01 + n.x];
  E.k[6] = kappa[r11 + n.xm]; E.k[7] = kappa[r11 + n.x];
  E.ksum = ((E.k[0] + E.k[1]) + (E.k[2] + E.k[3])) + ((E.k[4] + E.k[5]) + (E.k[6] + E.k[7]));
  return E;
}

// The same from an LDS tile of element conductivities: element
// (x-1+ax, y-1+ay, z-1+az) at sk[kb + ax + KX*(ay + KY*az)].
__device__ inline NodeElems tile_elems(const float* sk, int kb, int KX, int KY) {
  NodeElems E;
#pragma unroll
  for (int e = 0; e < 8; e++) E.k[e] = sk[kb + (e & 1) + KX * (((e >> 1) & 1) + KY * (e >> 2))];
  E.ksum = ((E.k[0] + E.k[1]) + (E.k[2] + E.k[3])) + ((E.k[4] + E.k[5]) + (E.k[6] + E.k[7]));
  return E;
}

// The 20 neighbours a node couples to (one-axis neighbours have a zero
// coefficient), named by offset: first letter x (m/0/p), then y, then z.
struct Far20 {
  float mm0, pm0, mp0, pp0, m0m, p0m, m0p, p0p;
  float mmm, _0mm, pmm, mmp, _0mp, pmp, mpm, _0pm, ppm, mpp, _0pp, ppp;
};

// sum_e kappa_e T_e(u): element (ax,ay,az) lies on side s = 2a-1 and
// couples to far nodes (sx,sy,0) (sx,0,sz) (0,sy,sz) (sx,sy,sz). One
// expression for every kernel, so each evaluates it in the same order.
__device__ inline float far_combine(const NodeElems& E, const Far20& u) {
  float s = 0.f;
  s += E.k[0] * (u.mm0 + u.m0m + u._0mm + u.mmm);
  s += E.k[1] * (u.pm0 + u.p0m + u._0mm + u.pmm);
  s += E.k[2] * (u.mp0 + u.m0m + u._0pm + u.mpm);
  s += E.k[3] * (u.pp0 + u.p0m + u._0pm + u.ppm);
  s += E.k[4] * (u.mm0 + u.m0p + u._0mp + u.mmp);
  s += E.k[5] * (u.pm0 + u.p0p + u._0mp + u.pmp);
  s += E.k[6] * (u.mp0 + u.m0p + u._0pp + u.mpp);
  s += E.k[7] * (u.pp0 + u.p0p + u._0pp + u.ppp);
  return s;
}

// The 20 neighbour values of node n from global memory.
__device__ inline Far20 load_far20(const Geom& g, const NodeIdx& n, const float* __restrict__ u) {
  // rows touched: (y+-1, z), (y, z+-1), (y+-1, z+-1)
  const long long rMm = rowb(g, n.ym, n.zm), rMp = rowb(g, n.ym, n.zp);
  const long long rPm = rowb(g, n.yp, n.zm), rPp = rowb(g, n.yp, n.zp);
  const long long rM0 = rowb(g, n.ym, n.z), rP0 = rowb(g, n.yp, n.z);
  const long long r0m = rowb(g, n.y, n.zm), r0p = rowb(g, n.y, n.zp);
  Far20 v;
  v.mm0 = u[rM0 + n.xm]; v.pm0 = u[rM0 + n.xp];
  v.mp0 = u[rP0 + n.xm]; v.pp0 = u[rP0 + n.xp];
  v.m0m = u[r0m + n.xm]; v.p0m = u[r0m + n.xp];
  v.m0p = u[r0p + n.xm]; v.p0p = u[r0p + n.xp];
  v.mmm = u[rMm + n.xm]; v._0mm = u[rMm + n.x]; v.pmm = u[rMm + n.xp];
  v.mmp = u[rMp + n.xm]; v._0mp = u[rMp + n.x]; v.pmp = u[rMp + n.xp];
  v.mpm = u[rPm + n.xm]; v._0pm = u[rPm + n.x]; v.ppm = u[rPm + n.xp];
  v.mpp = u[rPp + n.xm]; v._0pp = u[rPp + n.x]; v.ppp = u[rPp + n.xp];
  return v;
}

// sum_e kappa_e T_e(u) for one right-hand side, neighbours from global memory.
__device__ inline float far_sum(const Geom& g, const NodeIdx& n, const NodeElems& E,
                                const float* __restrict__ u) {
  return far_combine(E, load_far20(g, n, u));
}

// The same from an LDS tile: node at U[c0], row strides DY (y) and DZ (z).
template <typename F20>
__device__ inline F20 tile_far20(const float* U, int c0, int DY, int DZ) {
  F20 v;
  v.mm0 = U[c0 - 1 - DY]; v.pm0 = U[c0 + 1 - DY];
  v.mp0 = U[c0 - 1 + DY]; v.pp0 = U[c0 + 1 + DY];
  v.m0m = U[c0 - 1 - DZ]; v.p0m = U[c0 + 1 - DZ];
  v.m0p = U[c0 - 1 + DZ]; v.p0p = U[c0 + 1 + DZ];
  v.mmm = U[c0 - 1 - DY - DZ]; v._0mm = U[c0 - DY - DZ]; v.pmm = U[c0 + 1 - DY - DZ];
  v.mmp = U[c0 - 1 - DY + DZ]; v._0mp = U[c0 - DY + DZ]; v.pmp = U[c0 + 1 - DY + DZ];
  v.mpm = U[c0 - 1 + DY - DZ]; v._0pm = U[c0 + DY - DZ]; v.ppm = U[c0 + 1 + DY - DZ];
  v.mpp = U[c0 - 1 + DY + DZ]; v._0pp = U[c0 + DY + DZ]; v.ppp = U[c0 + 1 + DY + DZ];
  return v;
}
__device__ inline float far_sum_tile(const NodeElems& E, const float* U, int c0, int DY, int DZ) {
  return far_combine(E, tile_far20<Far20>(U, c0, DY, DZ));
}

// (K u)_v in difference form: the rows of K sum to zero, so
// (K u)_v = c sum_e kappa_e sum_{j far in e} (u_v - u_j). For smooth u the
// differences are small and exact (Sterbenz), so the FP32 result keeps FP32
// precision relative to itself; the direct form c(4 ksum u_v - far) loses
// it to cancellation.
__device__ inline float apply_diff(const NodeElems& E, float c, float uv, Far20 v) {
  v.mm0 = uv - v.mm0; v.pm0 = uv - v.pm0; v.mp0 = uv - v.mp0; v.pp0 = uv - v.pp0;
  v.m0m = uv - v.m0m; v.p0m = uv - v.p0m; v.m0p = uv - v.m0p; v.p0p = uv - v.p0p;
  v.mmm = uv - v.mmm; v._0mm = uv - v._0mm; v.pmm = uv - v.pmm; v.mmp = uv - v.mmp;
  v._0mp = uv - v._0mp; v.pmp = uv - v.pmp; v.mpm = uv - v.mpm; v._0pm = uv - v._0pm;
  v.ppm = uv - v.ppm; v.mpp = uv - v.mpp; v._0pp = uv - v._0pp; v.ppp = uv - v.ppp;
  return c * far_combine(E, v);
}

// FP64 variant of far_sum for residuals that must resolve below FP32 rounding.
struct Far20d {
  double mm0, pm0, mp0, pp0, m0m, p0m, m0p, p0p;
  double mmm, _0mm, pmm, mmp, _0mp, pmp, mpm, _0pm, ppm, mpp, _0pp, ppp;
};
__device__ inline double far_combine_d(const NodeElems& E, const Far20d& u) {
  double s = 0.0;
  s += (double)E.k[0] * (u.mm0 + u.m0m + u._0mm + u.mmm);
  s += (double)E.k[1] * (u.pm0 + u.p0m + u._0mm + u.pmm);
  s += (double)E.k[2] * (u.mp0 + u.m0m + u._0pm + u.mpm);
  s += (double)E.k[3] * (u.pp0 + u.p0m + u._0pm + u.ppm);
  s += (double)E.k[4] * (u.mm0 + u.m0p + u._0mp + u.mmp);
  s += (double)E.k[5] * (u.pm0 + u.p0p + u._0mp + u.pmp);
  s += (double)E.k[6] * (u.mp0 + u.m0p + u._0pp + u.mpp);
  s += (double)E.k[7] * (u.pp0 + u.p0p + u._0pp + u.ppp);
  return s;
}
template <typename XT>
__device__ inline double far_sum_d(const Geom& g, const NodeIdx& n, const NodeElems& E,
                                   const XT* __restrict__ u) {
  const long long rMm = rowb(g, n.ym, n.zm), rMp = rowb(g, n.ym, n.zp);
  const long long rPm = rowb(g, n.yp, n.zm), rPp = rowb(g, n.yp, n.zp);
  const long long rM0 = rowb(g, n.ym, n.z), rP0 = rowb(g, n.yp, n.z);
  const long long r0m = rowb(g, n.y, n.zm), r0p = rowb(g, n.y, n.zp);
  Far20d v;
  v.mm0 = u[rM0 + n.xm]; v.pm0 = u[rM0 + n.xp];
  v.mp0 = u[rP0 + n.xm]; v.pp0 = u[rP0 + n.xp];
  v.m0m = u[r0m + n.xm]; v.p0m = u[r0m + n.xp];
  v.m0p = u[r0p + n.xm]; v.p0p = u[r0p + n.xp];
  v.mmm = u[rMm + n.xm]; v._0mm = u[rMm + n.x]; v.pmm = u[rMm + n.xp];
  v.mmp = u[rMp + n.xm]; v._0mp = u[rMp + n.x]; v.pmp = u[rMp + n.xp];
  v.mpm = u[rPm + n.xm]; v._0pm = u[rPm + n.x]; v.ppm = u[rPm + n.xp];
  v.mpp = u[rPp + n.xm]; v._0pp = u[rPp + n.x]; v.ppp = u[rPp + n.xp];
  return far_combine_d(E, v);
}
// The same for an iterate held as an FP32 pair (hi + lo); both arrays are
// read at the same 32-bit byte offsets from wave-uniform bases.
__device__ inline double far_sum_pair(const Geom& g, const NodeIdx& n, const NodeElems& E,
                                      const float* __restrict__ u, const float* __restrict__ ul) {
  const unsigned nx = (unsigned)g.nx, nxy = (unsigned)g.nx * (unsigned)g.ny;
  auto row = [&](int y, int z) { return 4u * (nx * (unsigned)y + nxy * (unsigned)z); };
  const unsigned rMm = row(n.ym, n.zm), rMp = row(n.ym, n.zp);
  const unsigned rPm = row(n.yp, n.zm), rPp = row(n.yp, n.zp);
  const unsigned rM0 = row(n.ym, n.z), rP0 = row(n.yp, n.z);
  const unsigned r0m = row(n.y, n.zm), r0p = row(n.y, n.zp);
  const unsigned xm = 4u * (unsigned)n.xm, xp = 4u * (unsigned)n.xp, x0 = 4u * (unsigned)n.x;
  auto at = [&](unsigned i) { return (double)ld_off(u, i) + (double)ld_off(ul, i); };
  Far20d v;
  v.mm0 = at(rM0 + xm); v.pm0 = at(rM0 + xp);
  v.mp0 = at(rP0 + xm); v.pp0 = at(rP0 + xp);
  v.m0m = at(r0m + xm); v.p0m = at(r0m + xp);
  v.m0p = at(r0p + xm); v.p0p = at(r0p + xp);
  v.mmm = at(rMm + xm); v._0mm = at(rMm + x0); v.pmm = at(rMm + xp);
  v.mmp = at(rMp + xm); v._0mp = at(rMp + x0); v.pmp = at(rMp + xp);
  v.mpm = at(rPm + xm); v._0pm = at(rPm + x0); v.ppm = at(rPm + xp);
  v.mpp = at(rPp + xm); v._0pp = at(rPp + x0); v.ppp = at(rPp + xp);
  return far_combine_d(E, v);
}

// Load vector of case k at the node (grid.cu:1282-1342): sum_e kappa_e feMu[7-e][k].
__device__ inline float load_vec(const NodeElems& E, const FineConsts& fc, int k) {
  float s = 0.f;
#pragma unroll
  for (int e = 0; e < 8; e++) s += E.k[e] * fc.femu[7 - e][k];
  return s;
}

// One colour of the x-row Gauss-Seidel sweep: u_v = (f_v + c*far)/(4c*ksum).
__device__ inline void fine_gs_color_body(const Geom& g, const FineConsts& fc, const float* __restrict__ kappa,
                                          const float* __restrict__ f, float* u, int cy, int cz, long long t) {
  const int hy = g.ny >> 1, hz = g.nz >> 1;
  const long long cnt = (long long)g.nx * hy * hz;
  if (t >= cnt) return;
  int x, y, z;
  split_node(g.nx, hy, t, x, y, z);
  y = 2 * y + cy;
  z = 2 * z + cz;
  const NodeIdx n = node_idx(g, x, y, z);
  const NodeElems E = load_elems(g, n, kappa);
  const float inv = 1.f / (4.f * fc.c * E.ksum);
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) {
    const float* uk = u + k * g.nn;
    const float fs = far_sum(g, n, E, uk);
    u[k * g.nn + n.v] = (f[k * g.nn + n.v] + fc.c * fs) * inv;
  }
}

// res = f - K u on every node.
__device__ inline void fine_residual_body(const Geom& g, const FineConsts& fc, const float* __restrict__ kappa,
                                          const float* __restrict__ f, const float* __restrict__ u,
                                          float* __restrict__ res, long long t) {
  if (t >= g.nn) return;
  int x, y, z;
  split_node(g.nx, g.ny, t, x, y, z);
  const NodeIdx n = node_idx(g, x, y, z);
  const NodeElems E = load_elems(g, n, kappa);
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) {
    const float* uk = u + k * g.nn;
    const float ku = fc.c * (4.f * E.ksum * uk[n.v] - far_sum(g, n, E, uk));
    res[k * g.nn + n.v] = f[k * g.nn + n.v] - ku;
  }
}

// q = K p and partial sums of p.q (PCG step).
__device__ inline void fine_apply_dot_body(const Geom& g, const FineConsts& fc, const float* __restrict__ kappa,
                                           const float* __restrict__ p, float* __restrict__ q, const Reduce& R,
                                           int lb, int nb) {
  double acc[OTM_NRHS] = {0.0, 0.0, 0.0};
  const BlockRange br = block_range_v(g.nn, lb, nb);
  for (long long t = br.b + threadIdx.x; t < br.e; t += blockDim.x) {
    int x, y, z;
    split_node(g.nx, g.ny, t, x, y, z);
    const NodeIdx n = node_idx(g, x, y, z);
    const NodeElems E = load_elems(g, n, kappa);
#pragma unroll
    for (int k = 0; k < OTM_NRHS; k++) {
      const float* pk = p + k * g.nn;
      const float pv = pk[n.v];
      const float kp = apply_diff(E, fc.c, pv, load_far20(g, n, pk));
      q[k * g.nn + n.v] = kp;
      acc[k] += (double)pv * (double)kp;
    }
  }
  reduce_block_at<OTM_NRHS>(acc, R, lb);
}

// r = b - K x in FP64 with b the load vectors built on the fly (the FP64
// sum of the eight element contributions cancels exactly wherever kappa
// does not vary along the load direction); stored rounded to FP32. Partial
// sums of b.b, r.r and sum(r) (PCG start and restarts). XT is the iterate's
// type: FP32 (with xlo: an FP32 pair), or FP64 for the iterative refinement.
template <typename XT>
__device__ inline void fine_init_residual_body(const Geom& g, const FineConsts& fc, const float* __restrict__ kappa,
                                               const XT* __restrict__ x_, const float* __restrict__ xlo,
                                               float* __restrict__ rvec, const Reduce& R, int lb, int nb) {
  double acc[3 * OTM_NRHS] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const double c = (double)fc.c;
  const BlockRange br = block_range_v(g.nn, lb, nb);
  for (long long t = br.b + threadIdx.x; t < br.e; t += blockDim.x) {
    // 64-bit decode on purpose: this FP64-heavy loop measured twice as fast
    // with it on gfx950 as with split_node (47 against 95 us per launch at
    // 128^3), although the 32-bit decode is shorter code
    const int x = (int)(t % g.nx);
    const long long r = t / g.nx;
    const int y = (int)(r % g.ny), z = (int)(r / g.ny);
    const NodeIdx n = node_idx(g, x, y, z);
    const NodeElems E = load_elems(g, n, kappa);
    double ksum = 0.0;
#pragma unroll
    for (int e = 0; e < 8; e++) ksum += (double)E.k[e];
#pragma unroll
    for (int k = 0; k < OTM_NRHS; k++) {
      const XT* xk = x_ + k * g.nn;
      double b = 0.0;
#pragma unroll
      for (int e = 0; e < 8; e++) b += (double)E.k[e] * (double)fc.femu[7 - e][k];
      double kx;
      if constexpr (sizeof(XT) == 4) {
        if (xlo) {
          const float* xl = xlo + k * g.nn;
          kx = c * (4.0 * ksum * ((double)xk[n.v] + (double)xl[n.v]) - far_sum_pair(g, n, E, xk, xl));
        } else {
          kx = c * (4.0 * ksum * (double)xk[n.v] - far_sum_d(g, n, E, xk));
        }
      } else {
        kx = c * (4.0 * ksum * (double)xk[n.v] - far_sum_d(g, n, E, xk));
      }
      const double rv = b - kx;
      rvec[k * g.nn + n.v] = (float)rv;
      acc[k] += b * b;
      acc[OTM_NRHS + k] += rv * rv;
      acc[2 * OTM_NRHS + k] += rv;
    }
  }
  reduce_block_at<3 * OTM_NRHS>(acc, R, lb);
}

// fine_init_residual_body for load case k alone (a grid of nb x 3 blocks):
// block (lb, k) covers the nodes of block lb of the one-dimensional form
// in the same order, so r and the partial sums carry the same bits; it
// writes values k, 3 + k and 6 + k of partial slot lb. With one load case
// per thread, the gathered neighbours of three cases are not live together
// (222 registers and two waves per SIMD for the three-case form).
template <typename XT>
__device__ inline void fine_init_residual_case_body(const Geom& g, const FineConsts& fc,
                                                    const float* __restrict__ kappa, const XT* __restrict__ x_,
                                                    const float* __restrict__ xlo, float* __restrict__ rvec,
                                                    const Reduce& R, int lb, int nb, int k) {
  double acc[3] = {0.0, 0.0, 0.0};
  const double c = (double)fc.c;
  float fe[8];
#pragma unroll
  for (int e = 0; e < 8; e++) fe[e] = k == 0 ? fc.femu[7 - e][0] : (k == 1 ? fc.femu[7 - e][1] : fc.femu[7 - e][2]);
  const XT* xk = x_ + (size_t)k * (size_t)g.nn;
  const float* xl = xlo ? xlo + (size_t)k * (size_t)g.nn : nullptr;
  const BlockRange br = block_range_v(g.nn, lb, nb);
  for (long long t = br.b + threadIdx.x; t < br.e; t += blockDim.x) {
    // 64-bit decode as in fine_init_residual_body
    const int x = (int)(t % g.nx);
    const long long r = t / g.nx;
    const int y = (int)(r % g.ny), z = (int)(r / g.ny);
    const NodeIdx n = node_idx(g, x, y, z);
    const NodeElems E = load_elems(g, n, kappa);
    double ksum = 0.0;
#pragma unroll
    for (int e = 0; e < 8; e++) ksum += (double)E.k[e];
    double b = 0.0;
#pragma unroll
    for (int e = 0; e < 8; e++) b += (double)E.k[e] * (double)fe[e];
    double kx;
    if constexpr (sizeof(XT) == 4) {
      if (xl)
        kx = c * (4.0 * ksum * ((double)xk[n.v] + (double)xl[n.v]) - far_sum_pair(g, n, E, xk, xl));
      else
        kx = c * (4.0 * ksum * (double)xk[n.v] - far_sum_d(g, n, E, xk));
    } else {
      kx = c * (4.0 * ksum * (double)xk[n.v] - far_sum_d(g, n, E, xk));
    }
    const double rv = b - kx;
    rvec[(size_t)k * (size_t)g.nn + n.v] = (float)rv;
    acc[0] += b * b;
    acc[1] += rv * rv;
    acc[2] += rv;
  }
  __shared__ double sh[3 * kWaves];
  block_sum<3>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 3; i++) R.partials[(size_t)lb * (3 * OTM_NRHS) + k + OTM_NRHS * i] = acc[i];
  }
}

// fine_init_residual_case_body by whole x-rows, for rows of at most kBlock
// nodes: the neighbour sum of node (x, y, z) regrouped by element quadrant
// q = (ay, az) (sides sy = 2ay - 1, sz = 2az - 1), with G_q(x') =
// u(x', y+sy, z) + u(x', y, z+sz) + u(x', y+sy, z+sz) from column x' alone:
//   far(x) = A(x-1) + B(x+1) + sum_q (k(x-1,q) + k(x,q)) u(x, y+sy, z+sz),
//   A(x') = sum_q k(x',q) G_q(x'),  B(x') = sum_q k(x'-1,q) G_q(x'),
// k(x', q) the element at (x', y-1+ay, z-1+az). A thread holds the 3 x 3
// iterate values of its column around its row and the conductivities of
// the two element rows y-1 and y, puts A and B in LDS and after a barrier
// takes A(x-1) and B(x+1) of its row (periodic in x). Block lb of nb takes
// a contiguous range of rows (y fastest) and splits it into one segment per
// row group of nx threads; a group walks its segment in +y, so a step loads
// only the values of the row entering the window (3 of the iterate at
// y+2, 4 conductivities of element row y+1), issued before the barrier of
// the step before. The window is a ring of three rows (the loop is unrolled
// by three so no values move) and LDS is double-buffered: one barrier per
// step. Writes values k, 3 + k and 6 + k of partial slot lb.
__host__ __device__ inline bool fine_rows_ok(int nx) { return nx <= kBlock; }

template <typename XT>
struct RowsWindow {
  double V[3][3];       // V[slot][bz]: iterate at row slot, plane z - 1 + bz
  double KM[3][2], KP[3][2];  // element row slot: conductivity at x - 1 / x, plane z - 1 + az
  double SQ[3][2], H[3];      // KM + KP, and its sum over az
  double pu[3];         // the entering row's values
  float pk[4];
  unsigned y, z, pz[3];  // current row, node index of (0, 0, z - 1 + bz)
};

template <typename XT>
__device__ inline double rows_val(const XT* __restrict__ xk, const float* __restrict__ xl, unsigned v) {
  if constexpr (sizeof(XT) == 4) return xl ? (double)xk[v] + (double)xl[v] : (double)xk[v];
  else return xk[v];
}

template <typename XT>
__device__ inline void rows_elem_slot(RowsWindow<XT>& W, int e, float m0, float p0, float m1, float p1) {
  W.KM[e][0] = (double)m0;
  W.KP[e][0] = (double)p0;
  W.KM[e][1] = (double)m1;
  W.KP[e][1] = (double)p1;
  W.SQ[e][0] = W.KM[e][0] + W.KP[e][0];
  W.SQ[e][1] = W.KM[e][1] + W.KP[e][1];
  W.H[e] = W.SQ[e][0] + W.SQ[e][1];
}

// One step of a row group at ring phase P: rows y-1, y, y+1 in slots P,
// P+1, P+2 (mod 3); element rows y-1, y in slots P, P+1.
template <typename XT, int P>
__device__ inline void rows_step(RowsWindow<XT>& W, bool& have, unsigned row, bool act, unsigned ge, int x, int xm,
                                 int j, int nx, unsigned ny, unsigned nz, double (*sA)[kBlock],
                                 double (*sB)[kBlock], int buf, const float* __restrict__ kappa,
                                 const XT* __restrict__ xk, const float* __restrict__ xl, float* __restrict__ rk,
                                 const double fe[8], double c, double acc[3]) {
  constexpr int s0 = P % 3, s1 = (P + 1) % 3, s2 = (P + 2) % 3;
  const unsigned plane = (unsigned)nx * ny;
  if (act && !have) {
    W.y = row % ny;
    W.z = row / ny;
    const unsigned zm = W.z == 0 ? nz - 1 : W.z - 1, zp = W.z == nz - 1 ? 0 : W.z + 1;
    W.pz[0] = zm * plane;
    W.pz[1] = W.z * plane;
    W.pz[2] = zp * plane;
    const unsigned ym = W.y == 0 ? ny - 1 : W.y - 1, yp = W.y == ny - 1 ? 0 : W.y + 1;
    const unsigned yo[3] = {ym * (unsigned)nx, W.y * (unsigned)nx, yp * (unsigned)nx};
    constexpr int sl[3] = {s0, s1, s2};
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int bz = 0; bz < 3; bz++) W.V[sl[a]][bz] = rows_val(xk, xl, W.pz[bz] + yo[a] + (unsigned)x);
#pragma unroll
    for (int ay = 0; ay < 2; ay++) {
      const unsigned r0 = W.pz[0] + yo[ay], r1 = W.pz[1] + yo[ay];
      rows_elem_slot(W, sl[ay], kappa[r0 + xm], kappa[r0 + x], kappa[r1 + xm], kappa[r1 + x]);
    }
  }
  double Lv = 0.0, bl = 0.0;
  if (act) {
    double A = 0.0, B = 0.0, L = 0.0;
#pragma unroll
    for (int az = 0; az < 2; az++)
#pragma unroll
      for (int ay = 0; ay < 2; ay++) {
        const int a = ay ? s2 : s0, e = ay ? s1 : s0, bz = 2 * az;
        const double G = W.V[a][1] + W.V[s1][bz] + W.V[a][bz];
        A += W.KP[e][az] * G;
        B += W.KM[e][az] * G;
        L += W.SQ[e][az] * W.V[a][bz];
      }
    // load vector: element (x-1+ax, y-1+ay, z-1+az) weighs fe[ax + 2 ay + 4 az]
#pragma unroll
    for (int az = 0; az < 2; az++)
#pragma unroll
      for (int ay = 0; ay < 2; ay++) {
        const int e = ay ? s1 : s0;
        bl += W.KM[e][az] * fe[2 * ay + 4 * az];
        bl += W.KP[e][az] * fe[1 + 2 * ay + 4 * az];
      }
    const double ksum = W.H[s0] + W.H[s1];
    sA[buf][j] = A;
    sB[buf][j] = B;
    Lv = 4.0 * ksum * W.V[s1][1] - L;
  }
  // the next row of the segment is (y + 1, z): fetch its entering values now
  const bool slide = act && row + 1 < ge && W.y + 1 < ny;
  if (slide) {
    const unsigned y1 = (W.y + 1) * (unsigned)nx, y2 = W.y + 2 == ny ? 0u : y1 + (unsigned)nx;
#pragma unroll
    for (int bz = 0; bz < 3; bz++) W.pu[bz] = rows_val(xk, xl, W.pz[bz] + y2 + (unsigned)x);
    const unsigned r0 = W.pz[0] + y1, r1 = W.pz[1] + y1;
    W.pk[0] = kappa[r0 + xm];
    W.pk[1] = kappa[r0 + x];
    W.pk[2] = kappa[r1 + xm];
    W.pk[3] = kappa[r1 + x];
  }
  __syncthreads();
  if (act) {
    const int jl = x == 0 ? j + nx - 1 : j - 1, jr = x == nx - 1 ? j - nx + 1 : j + 1;
    const double kx = c * (Lv - (sA[buf][jl] + sB[buf][jr]));
    const double rv = bl - kx;
    rk[W.pz[1] + W.y * (unsigned)nx + (unsigned)x] = (float)rv;
    acc[0] += bl * bl;
    acc[1] += rv * rv;
    acc[2] += rv;
  }
  if (slide) {
    // row y-1 leaves slot s0 for row y+2; element row y+1 takes slot s2
#pragma unroll
    for (int bz = 0; bz < 3; bz++) W.V[s0][bz] = W.pu[bz];
    rows_elem_slot(W, s2, W.pk[0], W.pk[1], W.pk[2], W.pk[3]);
    W.y += 1;
  }
  have = slide;
}

template <typename XT>
__device__ inline void fine_init_residual_rows_body(const Geom& g, const FineConsts& fc,
                                                    const float* __restrict__ kappa, const XT* __restrict__ x_,
                                                    const float* __restrict__ xlo, float* __restrict__ rvec,
                                                    const Reduce& R, int lb, int nb, int k) {
  __shared__ double sA[2][kBlock], sB[2][kBlock];
  double acc[3] = {0.0, 0.0, 0.0};
  const double c = (double)fc.c;
  double fe[8];
#pragma unroll
  for (int e = 0; e < 8; e++)
    fe[e] = (double)(k == 0 ? fc.femu[7 - e][0] : (k == 1 ? fc.femu[7 - e][1] : fc.femu[7 - e][2]));
  const XT* xk = x_ + (size_t)k * (size_t)g.nn;
  const float* xl = xlo ? xlo + (size_t)k * (size_t)g.nn : nullptr;
  float* rk = rvec + (size_t)k * (size_t)g.nn;
  const int nx = g.nx;
  const unsigned ny = (unsigned)g.ny, nz = (unsigned)g.nz, rows = ny * nz;
  const unsigned per = (rows + (unsigned)nb - 1) / (unsigned)nb;
  const unsigned rb = min((unsigned)lb * per, rows), re = min(rb + per, rows);
  const int rpc = kBlock / nx;  // row groups
  const int grp = (int)threadIdx.x / nx;
  const int x = (int)threadIdx.x - grp * nx, xm = x == 0 ? nx - 1 : x - 1;
  const unsigned seg = (re - rb + (unsigned)rpc - 1) / (unsigned)rpc;
  const unsigned gs = rb + (unsigned)grp * seg, ge = min(gs + seg, re);
  const bool grp_ok = grp < rpc;
  RowsWindow<XT> W;
  W.y = W.z = 0;
  W.pz[0] = W.pz[1] = W.pz[2] = 0;
  bool have = false;
  for (unsigned t = 0; t < seg;) {
    rows_step<XT, 0>(W, have, gs + t, grp_ok && gs + t < ge, ge, x, xm, (int)threadIdx.x, nx, ny, nz, sA, sB,
                     (int)(t & 1), kappa, xk, xl, rk, fe, c, acc);
    if (++t >= seg) break;
    rows_step<XT, 1>(W, have, gs + t, grp_ok && gs + t < ge, ge, x, xm, (int)threadIdx.x, nx, ny, nz, sA, sB,
                     (int)(t & 1), kappa, xk, xl, rk, fe, c, acc);
    if (++t >= seg) break;
    rows_step<XT, 2>(W, have, gs + t, grp_ok && gs + t < ge, ge, x, xm, (int)threadIdx.x, nx, ny, nz, sA, sB,
                     (int)(t & 1), kappa, xk, xl, rk, fe, c, acc);
    ++t;
  }
  __shared__ double sh[3 * kWaves];
  block_sum<3>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 3; i++) R.partials[(size_t)lb * (3 * OTM_NRHS) + k + OTM_NRHS * i] = acc[i];
  }
  __syncthreads();  // sh is free again for a following call in the same block
}

// Element energies (homogenization.cu:13-153 and :255-412). For element e
// with nodal fields d_k = disp[:,k] - chi_k(nodes): S_ij = d_i^T Ke d_j.
// Ke has zero row sums and couples a node only to its 3 face-diagonal and
// its body-diagonal partners (-c each), so S_ij = c sum over those 16 node
// pairs (a, b) of D_i(a,b) D_j(a,b) with D_k(a,b) = d_k,a - d_k,b. The
// differences are formed from the fluctuation field first: disp differences
// are 0 or +-h exactly, and chi differences of neighbouring nodes are exact
// or nearly so, so S keeps the relative precision of its own value (the
// direct form d^T (Ke d) cancels to it from O(1) products and loses a factor
// of the cell edge: 1e-5 at 256^3 in FP32). The fluctuation field is read
// as FP64 (iterative refinement), or as the FP32 pair hi + lo of the plain
// solve, or as FP32 alone.
struct ChiRef {
  const float* hi;
  const float* lo;    // may be null
  const double* d64;  // when set, used instead of hi/lo
};

// the 16 coupled node pairs of an element (local index ax + 2 ay + 4 az)
__device__ constexpr int kPairA[16] = {0, 0, 0, 1, 1, 1, 2, 2, 3, 3, 4, 5, 0, 1, 2, 3};
__device__ constexpr int kPairB[16] = {3, 5, 6, 2, 4, 7, 4, 7, 5, 6, 7, 6, 7, 6, 5, 4};

__device__ inline void element_S(const Geom& g, const FineConsts& fc, const ChiRef& chi, int x, int y, int z,
                                 float S[6]) {
  const int xp = x == g.nx - 1 ? 0 : x + 1;
  const int yp = y == g.ny - 1 ? 0 : y + 1;
  const int zp = z == g.nz - 1 ? 0 : z + 1;
  const long long r0 = rowb(g, y, z), r1 = rowb(g, yp, z), r2 = rowb(g, y, zp), r3 = rowb(g, yp, zp);
  const long long nid[8] = {r0 + x, r0 + xp, r1 + x, r1 + xp, r2 + x, r2 + xp, r3 + x, r3 + xp};
  if (chi.d64) {
    double v[3][8];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int m = 0; m < 8; m++) v[k][m] = chi.d64[k * g.nn + nid[m]];
    double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int a = kPairA[q], b = kPairB[q];
      double D[3];
#pragma unroll
      for (int k = 0; k < 3; k++) D[k] = (double)(fc.disp[a][k] - fc.disp[b][k]) - (v[k][a] - v[k][b]);
      acc[0] += D[0] * D[0]; acc[1] += D[0] * D[1]; acc[2] += D[0] * D[2];
      acc[3] += D[1] * D[1]; acc[4] += D[1] * D[2]; acc[5] += D[2] * D[2];
    }
#pragma unroll
    for (int c = 0; c < 6; c++) S[c] = (float)((double)fc.c * acc[c]);
    return;
  }
  float v[3][8], w[3][8];
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int m = 0; m < 8; m++) {
      v[k][m] = chi.hi[k * g.nn + nid[m]];
      w[k][m] = chi.lo ? chi.lo[k * g.nn + nid[m]] : 0.f;
    }
  float acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int a = kPairA[q], b = kPairB[q];
    float D[3];
#pragma unroll
    for (int k = 0; k < 3; k++)
      D[k] = (fc.disp[a][k] - fc.disp[b][k]) - ((v[k][a] - v[k][b]) + (w[k][a] - w[k][b]));
    acc[0] += D[0] * D[0]; acc[1] += D[0] * D[1]; acc[2] += D[0] * D[2];
    acc[3] += D[1] * D[1]; acc[4] += D[1] * D[2]; acc[5] += D[2] * D[2];
  }
#pragma unroll
  for (int c = 0; c < 6; c++) S[c] = fc.c * acc[c];
}

// kappa^H_ij = sum_e (kappa_e) S_e(i,j) with FP64 accumulation; kappa
// already holds the 1e-9 floor.
// S6 (may be null) receives the element energies as fine_energy6_body
// writes them, for the sensitivities that follow the same solve.
__device__ inline void fine_kappa_h_body(const Geom& g, const FineConsts& fc, const float* __restrict__ kappa,
                                         const ChiRef& chi, const Reduce& R, int lb, int nb,
                                         float* __restrict__ S6 = nullptr) {
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const BlockRange br = block_range_v(g.nn, lb, nb);
  for (long long t = br.b + threadIdx.x; t < br.e; t += blockDim.x) {
    int x, y, z;
    split_node(g.nx, g.ny, t, x, y, z);
    float S[6];
    element_S(g, fc, chi, x, y, z, S);
    if (S6) {
#pragma unroll
      for (int c = 0; c < 6; c++) S6[c * g.nn + t] = S[c];
    }
    const double ke = (double)kappa[t];
#pragma unroll
    for (int c = 0; c < 6; c++) acc[c] += ke * (double)S[c];
  }
  reduce_block_at<6>(acc, R, lb);
}

// sens[x + pitch*(y + ny*z)] = sum_c w_c S_c (contracted sensitivity with
// respect to kappa; the penalty exponent is 1, homogenization.cu:344-346).
// scale (may be null): the sensitivity times scale[e] (d kappa / d rho of
// the element, for the design gradient).
__device__ inline void fine_sensitivity_body(const Geom& g, const FineConsts& fc, const ChiRef& chi, const float w[6],
                                             float* __restrict__ sens, int pitch, long long t,
                                             const float* __restrict__ scale = nullptr) {
  if (t >= g.nn) return;
  int x, y, z;
  split_node(g.nx, g.ny, t, x, y, z);
  float S[6];
  element_S(g, fc, chi, x, y, z, S);
  float s = S[0] * w[0] + S[1] * w[1] + S[2] * w[2] + S[3] * w[3] + S[4] * w[4] + S[5] * w[5];
  if (scale) s = s * scale[t];
  sens[(long long)x + (long long)pitch * ((long long)y + (long long)g.ny * z)] = s;
}

// The contraction of fine_sensitivity_body from stored element energies
// (the same float expression, so the same bits).
__device__ inline void sens_from_energy_body(const Geom& g, const float* __restrict__ S6, const float w[6],
                                             float* __restrict__ sens, int pitch, long long t,
                                             const float* __restrict__ scale = nullptr) {
  if (t >= g.nn) return;
  const long long nn = g.nn;
  float s = S6[t] * w[0] + S6[nn + t] * w[1] + S6[2 * nn + t] * w[2] + S6[3 * nn + t] * w[3] +
            S6[4 * nn + t] * w[4] + S6[5 * nn + t] * w[5];
  if (scale) s = s * scale[t];
  if (pitch == g.nx) {
    sens[t] = s;
  } else {
    const long long x = t % g.nx, r = t / g.nx;
    sens[x + (long long)pitch * r] = s;
  }
}

// S6[c*nn + e]: the six element energies, for callers that keep them.
__device__ inline void fine_energy6_body(const Geom& g, const FineConsts& fc, const ChiRef& chi,
                                         float* __restrict__ S6, long long t) {
  if (t >= g.nn) return;
  int x, y, z;
  split_node(g.nx, g.ny, t, x, y, z);
  float S[6];
  element_S(g, fc, chi, x, y, z, S);
#pragma unroll
  for (int c = 0; c < 6; c++) S6[c * g.nn + t] = S[c];
}

// ---- LDS tiles of the fine level ----------------------------------------------
// A tile is BX x BY x BZ nodes; a workgroup of TR thread rows of BX threads
// stages the conductivity of every element touching the tile and up to
// three fields with a one-node halo. Staging reads whole x-rows of the tile
// (coalesced) and the two halo columns separately, with no per-element
// index division. The default shape, 32 x 4 x 8 nodes for 256 threads,
// measured fastest for the smoother on gfx950 among the shapes tried
// (32x8x8, 32x8x4, 32x4x8, 32x4x4, 32x2x8, 32x4x16, 64x4x4, 64x4x8,
// 64x8x4, 64x8x8, 16x8x8, 16x4x8 at 128 to 512 threads): more resident
// workgroups hide the staging latency better than the smaller halo share of
// larger tiles saves.
constexpr int kFineBY = 4, kFineBZ = 8, kFineTR = 8;

template <int BX, int BY, int BZ, int TR>
struct FineTile {
  static constexpr int HX = BX + 2, HY = BY + 2, HZ = BZ + 2;  // nodes with the halo
  static constexpr int KX = BX + 1, KY = BY + 1, KZ = BZ + 1;  // elements touching the tile
  static constexpr int NT = BX * TR;                           // threads
  static constexpr int kU = HX * HY * HZ, kK = KX * KY * KZ;   // LDS floats per field / of kappa
};

// Tile width along x for a level (0: the level does not tile); tiles per level.
__host__ __device__ inline int fine_tile_x(const Geom& g) {
  if (g.ny % kFineBY || g.nz % kFineBZ) return 0;
  if (g.nx % 32 == 0) return 32;
  if (g.nx % 16 == 0) return 16;
  return 0;
}
__host__ __device__ inline int fine_tiles(const Geom& g, int bx) {
  return (g.nx / bx) * (g.ny / kFineBY) * (g.nz / kFineBZ);
}
// Workgroups of the tiled operator apply: one tile each up to the partial-sum
// capacity of a reduction, several tiles each beyond.
constexpr int kTileRedBlocks = kMaxReduceBlocks;
// Rows of the fine smoother's per-tile r.z partials (a context's tile
// partials hold kTileDotSlots x 3 doubles): 65536 tiles of 512 nodes and
// more cover a 1024^3 cell.
constexpr int kTileDotSlots = 65536;
__host__ __device__ inline unsigned fine_apply_tile_blocks(const Geom& g, int bx) {
  const int t = fine_tiles(g, bx);
  return (unsigned)(t > kTileRedBlocks ? kTileRedBlocks : t);
}

struct TileOrigin {
  int ox, oy, oz;
};
template <int BX, int BY, int BZ>
__device__ inline TileOrigin tile_origin(const Geom& g, int b) {
  const int bx = g.nx / BX, by = g.ny / BY;
  return {(b % bx) * BX, ((b / bx) % by) * BY, (b / (bx * by)) * BZ};
}

// Staging runs in two phases: every global load of a tile goes into
// registers before the first LDS store, so a workgroup waits for HBM once
// per tile instead of once per loop trip (the loops' trip counts are fixed,
// so they unroll with a predicate on the last trip).
//
// Address arithmetic: a wave holds 64 / BX thread rows, so the x-runs it
// loads in one trip are a few rows known to the whole wave. Their offsets
// (periodic wrap, row and plane strides) are computed once per wave on the
// scalar unit and each lane picks its row's; loads then take a wave-uniform
// base and a 32-bit byte offset (a lattice has at most 2^30 nodes). Done
// per lane, this arithmetic was about 40% of the smoother's vector
// instructions.
template <int BX, int BY, int BZ, int TR>
struct TileStage {
  using T = FineTile<BX, BY, BZ, TR>;
  static constexpr int WR = 64 / BX;                   // thread rows per wave
  static constexpr int KR = T::KY * T::KZ;             // kappa rows (x-runs of KX elements)
  static constexpr int KJ = (KR + TR - 1) / TR;        // interior runs per thread
  static constexpr int KH = (KR + T::NT - 1) / T::NT;  // left-column elements per thread
  static constexpr int UR = T::HY * T::HZ;             // field rows (x-runs of HX nodes)
  static constexpr int UJ = (UR + TR - 1) / TR;
  static constexpr int UH = (2 * UR + T::NT - 1) / T::NT;  // halo-column nodes per thread
  float k[KJ], kh[KH];
  float u[OTM_NRHS][UJ], uh[OTM_NRHS][UH];
};

// first thread row of this wave (wave-uniform) and the lane's row within it
template <int BX>
__device__ inline int wave_first_row() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * (64 / BX);
}
template <int BX>
__device__ inline int lane_row() {
  return (int)(threadIdx.x & 63) / BX;
}
// node offset of row (y, z) of a cell, wrapped periodically
__device__ inline unsigned wrapped_row(const Geom& g, int y, int z) {
  return (unsigned)g.nx * (unsigned)wrapm(y, g.ny) + (unsigned)g.nx * (unsigned)g.ny * (unsigned)wrapm(z, g.nz);
}
// Byte offset of the x-run (row r of an H-wide y range starting at y0,
// z0) of this lane's thread row: the wave's rows are evaluated as uniform
// values and the lane selects its own. Rows past `rows` are clamped.
template <int BX, int H>
__device__ inline unsigned lane_run_offset(const Geom& g, int rb, int rows, int y0, int z0) {
  unsigned off = 0u;
  const int sub = lane_row<BX>();
#pragma unroll
  for (int q = 0; q < 64 / BX; q++) {
    int r = rb + q;
    r = r < rows ? r : rows - 1;
    const unsigned o = wrapped_row(g, y0 + r % H, z0 + r / H);
    off = sub == q ? o : off;
  }
  return 4u * off;
}

// kappa of element (ox-1+kx, oy-1+ky, oz-1+kz), for sk[kx + KX*(ky + KY*kz)]
template <int BX, int BY, int BZ, int TR>
__device__ inline void load_tile_kappa(const Geom& g, const float* __restrict__ kappa, const TileOrigin& o,
                                       TileStage<BX, BY, BZ, TR>& st) {
  using S = TileStage<BX, BY, BZ, TR>;
  using T = FineTile<BX, BY, BZ, TR>;
  const int t = threadIdx.x, tx = t % BX;
  const int ty = wave_first_row<BX>() + lane_row<BX>();
  const int wr0 = wave_first_row<BX>();
#pragma unroll
  for (int j = 0; j < S::KJ; j++) {
    const int rb = wr0 + TR * j;  // wave-uniform
    const unsigned off = lane_run_offset<BX, T::KY>(g, rb, S::KR, o.oy - 1, o.oz - 1);
    if (ty + TR * j < S::KR) st.k[j] = ld_off(kappa, off + 4u * (unsigned)(o.ox + tx));
  }
  const int xm = wrapm(o.ox - 1, g.nx);
#pragma unroll
  for (int j = 0; j < S::KH; j++) {
    const int r = t + T::NT * j;
    if (r < S::KR) {
      const int ky = r % T::KY, kz = r / T::KY;
      st.kh[j] = ld_off(kappa, 4u * (wrapped_row(g, o.oy - 1 + ky, o.oz - 1 + kz) + (unsigned)xm));
    }
  }
}
template <int BX, int BY, int BZ, int TR>
__device__ inline void store_tile_kappa(float* sk, const TileStage<BX, BY, BZ, TR>& st) {
  using S = TileStage<BX, BY, BZ, TR>;
  using T = FineTile<BX, BY, BZ, TR>;
  const int t = threadIdx.x, tx = t % BX, ty = t / BX;
#pragma unroll
  for (int j = 0; j < S::KJ; j++) {
    const int r = ty + TR * j;
    if (r < S::KR) sk[r * T::KX + 1 + tx] = st.k[j];
  }
#pragma unroll
  for (int j = 0; j < S::KH; j++) {
    const int r = t + T::NT * j;
    if (r < S::KR) sk[r * T::KX] = st.kh[j];
  }
}

// field k (k < OTM_NRHS) of node (ox-1+hx, oy-1+hy, oz-1+hz), for
// su[k*kU + hx + HX*(hy + HY*hz)]
template <int BX, int BY, int BZ, int TR>
__device__ inline void load_tile_fields(const Geom& g, const float* __restrict__ src, const TileOrigin& o,
                                        TileStage<BX, BY, BZ, TR>& st) {
  using S = TileStage<BX, BY, BZ, TR>;
  using T = FineTile<BX, BY, BZ, TR>;
  const int t = threadIdx.x, tx = t % BX;
  const int wr0 = wave_first_row<BX>();
  const int ty = wr0 + lane_row<BX>();
  const float* srck[OTM_NRHS];
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) srck[k] = src + (size_t)k * (size_t)g.nn;
#pragma unroll
  for (int j = 0; j < S::UJ; j++) {
    const int rb = wr0 + TR * j;
    const unsigned off = lane_run_offset<BX, T::HY>(g, rb, S::UR, o.oy - 1, o.oz - 1) + 4u * (unsigned)(o.ox + tx);
    if (ty + TR * j < S::UR) {
#pragma unroll
      for (int k = 0; k < OTM_NRHS; k++) st.u[k][j] = ld_off(srck[k], off);
    }
  }
  const int xm = wrapm(o.ox - 1, g.nx), xp = wrapm(o.ox + BX, g.nx);
#pragma unroll
  for (int j = 0; j < S::UH; j++) {
    const int i = t + T::NT * j;
    if (i < 2 * S::UR) {
      const int r = i >> 1, side = i & 1;
      const int hy = r % T::HY, hz = r / T::HY;
      const unsigned w = 4u * (wrapped_row(g, o.oy - 1 + hy, o.oz - 1 + hz) + (unsigned)(side ? xp : xm));
#pragma unroll
      for (int k = 0; k < OTM_NRHS; k++) st.uh[k][j] = ld_off(srck[k], w);
    }
  }
}
template <int BX, int BY, int BZ, int TR>
__device__ inline void store_tile_fields(float* su, const TileStage<BX, BY, BZ, TR>& st) {
  using S = TileStage<BX, BY, BZ, TR>;
  using T = FineTile<BX, BY, BZ, TR>;
  const int t = threadIdx.x, tx = t % BX, ty = t / BX;
#pragma unroll
  for (int j = 0; j < S::UJ; j++) {
    const int r = ty + TR * j;
    if (r < S::UR) {
#pragma unroll
      for (int k = 0; k < OTM_NRHS; k++) su[k * T::kU + r * T::HX + 1 + tx] = st.u[k][j];
    }
  }
#pragma unroll
  for (int j = 0; j < S::UH; j++) {
    const int i = t + T::NT * j;
    if (i < 2 * S::UR) {
      const int r = i >> 1, hx = (i & 1) ? T::HX - 1 : 0;
#pragma unroll
      for (int k = 0; k < OTM_NRHS; k++) su[k * T::kU + r * T::HX + hx] = st.uh[k][j];
    }
  }
}

// Both at once (the tiled apply and residual).
template <int BX, int BY, int BZ, int TR>
__device__ inline void stage_tile(const Geom& g, const float* __restrict__ kappa, float* sk,
                                  const float* __restrict__ src, float* su, const TileOrigin& o) {
  TileStage<BX, BY, BZ, TR> st;
  load_tile_kappa<BX, BY, BZ, TR>(g, kappa, o, st);
  load_tile_fields<BX, BY, BZ, TR>(g, src, o, st);
  store_tile_kappa<BX, BY, BZ, TR>(sk, st);
  store_tile_fields<BX, BY, BZ, TR>(su, st);
}

// q = K p on the tiles lb, lb + nb, ... and the block's partial sums of p.q
// (slot lb). Each node's arithmetic is fine_apply_dot_body's: the
// difference form (apply_diff), because with the direct form the PCG
// residual recurrence drifted from the true residual.
template <int BX, int TR>
__device__ inline void fine_apply_tile_body(const Geom& g, const FineConsts& fc, const float* __restrict__ kappa,
                                            const float* __restrict__ p, float* __restrict__ q, const Reduce& R,
                                            int lb, int nb) {
  using T = FineTile<BX, kFineBY, kFineBZ, TR>;
  __shared__ float su[OTM_NRHS * T::kU];
  __shared__ float sk[T::kK];
  const int tx = threadIdx.x % BX, ty = threadIdx.x / BX;
  const int tiles = fine_tiles(g, BX);
  double acc[OTM_NRHS] = {0.0, 0.0, 0.0};
  for (int b = lb; b < tiles; b += nb) {
    const TileOrigin o = tile_origin<BX, kFineBY, kFineBZ>(g, b);
    stage_tile<BX, kFineBY, kFineBZ, TR>(g, kappa, sk, p, su, o);
    __syncthreads();
    for (int r = ty; r < kFineBY * kFineBZ; r += TR) {
      const int ly = r % kFineBY, lz = r / kFineBY;
      const NodeElems E = tile_elems(sk, tx + T::KX * (ly + T::KY * lz), T::KX, T::KY);
      const int c0 = (tx + 1) + T::HX * ((ly + 1) + T::HY * (lz + 1));
      const long long v = (long long)(o.ox + tx) + (long long)g.nx * ((long long)(o.oy + ly) + (long long)g.ny * (o.oz + lz));
#pragma unroll
      for (int k = 0; k < OTM_NRHS; k++) {
        const float* U = su + k * T::kU;
        const float pv = U[c0];
        const float kp = apply_diff(E, fc.c, pv, tile_far20<Far20>(U, c0, T::HX, T::HX * T::HY));
        q[k * g.nn + v] = kp;
        acc[k] += (double)pv * (double)kp;
      }
    }
    __syncthreads();
  }
  reduce_block_at<OTM_NRHS>(acc, R, lb);
}

// res = f - K u on tile b; each node's arithmetic is fine_residual_body's.
template <int BX, int TR>
__device__ inline void fine_residual_tile_body(const Geom& g, const FineConsts& fc, const float* __restrict__ kappa,
                                               const float* __restrict__ f, const float* __restrict__ u,
                                               float* __restrict__ res, int b) {
  using T = FineTile<BX, kFineBY, kFineBZ, TR>;
  __shared__ float su[OTM_NRHS * T::kU];
  __shared__ float sk[T::kK];
  const int tx = threadIdx.x % BX, ty = threadIdx.x / BX;
  const TileOrigin o = tile_origin<BX, kFineBY, kFineBZ>(g, b);
  stage_tile<BX, kFineBY, kFineBZ, TR>(g, kappa, sk, u, su, o);
  __syncthreads();
  for (int r = ty; r < kFineBY * kFineBZ; r += TR) {
    const int ly = r % kFineBY, lz = r / kFineBY;
    const NodeElems E = tile_elems(sk, tx + T::KX * (ly + T::KY * lz), T::KX, T::KY);
    const int c0 = (tx + 1) + T::HX * ((ly + 1) + T::HY * (lz + 1));
    const long long v = (long long)(o.ox + tx) + (long long)g.nx * ((long long)(o.oy + ly) + (long long)g.ny * (o.oz + lz));
#pragma unroll
    for (int k = 0; k < OTM_NRHS; k++) {
      const float* U = su + k * T::kU;
      const float ku = fc.c * (4.f * E.ksum * U[c0] - far_sum_tile(E, U, c0, T::HX, T::HX * T::HY));
      res[k * g.nn + v] = f[k * g.nn + v] - ku;
    }
  }
}

// Fine-level brick smoother on a tile: `sweeps` x-row Gauss-Seidel sweeps
// (4 colours) in LDS with workgroup barriers. Halo values stay frozen at
// their entry values and the result goes to a separate buffer (block
// Gauss-Seidel across tiles, exact colour order inside); the post-smoother
// runs the colours in reverse. Each thread updates (BY/2)(BZ/2)/TR nodes of
// the active colour per step (one with the default tile and 256 threads).
//
// A node couples to its 12 face-diagonal neighbours with the summed
// conductivity of the two elements they share and to its 8 body-diagonal
// neighbours with one element's conductivity (one-axis neighbours have a
// zero coefficient), so an update is 20 LDS reads and 20 FMAs per load case
// plus 12 adds shared by the three.
//
// With dot set (post-smoother of a PCG preconditioner application, 256
// threads), the block also forms its partial sums of f.u over its nodes
// with the final u (r.z: f is the PCG residual, u the preconditioned
// residual), from the values it already holds, into slot b of RD.
template <int BX, int BY = kFineBY, int BZ = kFineBZ, int TR = kFineTR>
__device__ inline void fine_brick_body(const Geom& g, const FineConsts& fc, const float* __restrict__ kappa,
                                       const float* __restrict__ f, const float* __restrict__ uin,
                                       float* __restrict__ uout, int sweeps, int forward, int zero_in, int b,
                                       int dot = 0, const Reduce& RD = Reduce{}) {
  using T = FineTile<BX, BY, BZ, TR>;
  constexpr int HX = T::HX, HY = T::HY, KX = T::KX, KY = T::KY;
  constexpr int R = (BY / 2) * (BZ / 2);  // x-rows of one colour
  constexpr int H = R / TR;               // of them per thread row
  static_assert(R % TR == 0, "thread rows must divide the rows of a colour");
  __shared__ float su[OTM_NRHS][T::kU];
  __shared__ float sk[T::kK];
  const TileOrigin o = tile_origin<BX, BY, BZ>(g, b);
  const int ox = o.ox, oy = o.oy, oz = o.oz;
  const int t = threadIdx.x;
  const int tx = t % BX, ty = t / BX;  // ty in [0, TR)
  const long long nx = g.nx, nxy = (long long)g.nx * g.ny;
  // the tile's global loads (kappa, the entry field) are all issued before
  // the first LDS store; the right-hand side's loads then fly during the
  // stores and the barrier
  {
    TileStage<BX, BY, BZ, TR> st;
    load_tile_kappa<BX, BY, BZ, TR>(g, kappa, o, st);
    if (!zero_in) load_tile_fields<BX, BY, BZ, TR>(g, uin, o, st);
    store_tile_kappa<BX, BY, BZ, TR>(sk, st);
    if (zero_in) {
      for (int i = t; i < T::kU; i += T::NT) {
#pragma unroll
        for (int k = 0; k < OTM_NRHS; k++) su[k][i] = 0.f;
      }
    } else {
      store_tile_fields<BX, BY, BZ, TR>(&su[0][0], st);
    }
  }
  // node ownership: x = tx; of colour (cy, cz) the x-rows rr = ty + TR*h,
  // y = cy + 2*(rr % (BY/2)), z = cz + 2*(rr / (BY/2)); row offsets are
  // evaluated per wave as in the staging
  const int lx = tx;
  const int wr0 = wave_first_row<BX>(), sub = lane_row<BX>();
  const unsigned nxu = (unsigned)g.nx, nxyu = (unsigned)g.nx * (unsigned)g.ny;
  float fv[4][H][OTM_NRHS];
#pragma unroll
  for (int col = 0; col < 4; col++)
#pragma unroll
    for (int h = 0; h < H; h++) {
      unsigned off = 0u;
#pragma unroll
      for (int q = 0; q < 64 / BX; q++) {
        const int rr = wr0 + q + TR * h;
        const int ly = (col & 1) + 2 * (rr % (BY / 2)), lz = (col >> 1) + 2 * (rr / (BY / 2));
        const unsigned oq = nxu * (unsigned)(oy + ly) + nxyu * (unsigned)(oz + lz);
        off = sub == q ? oq : off;
      }
      off = 4u * (off + (unsigned)(ox + lx));
#pragma unroll
      for (int k = 0; k < OTM_NRHS; k++) fv[col][h][k] = ld_off(f + (size_t)k * (size_t)g.nn, off);
    }
  __syncthreads();
  const float c = fc.c;
  constexpr int DY = HX, DZ = HX * HY;
  constexpr int LO = DY + DZ;  // reads start at the node's (y-1, z-1) row
  // The neighbours at x-1 and x+1 of a node are summed in the lanes that
  // hold them (the same tile row, the next lower/higher lane), with the
  // conductivities of the elements on their side, and moved over by DPP:
  // with G_e(x') = u(x', y+sy, z) + u(x', y, z+sz) + u(x', y+sy, z+sz) for
  // the element quadrant e = (ay, az) (sy = 2ay - 1, sz = 2az - 1),
  //   sum_j w_j u_j = sum_e k(x-1,e) G_e(x-1) + k(x,e) G_e(x+1)
  //                 + (k(x-1,e) + k(x,e)) u(x, y+sy, z+sz),
  // so lane x' forms T1 = sum_e k(x',e) G_e(x') (needed by x'+1) and
  // T0 = sum_e k(x'-1,e) G_e(x') (needed by x'-1): a node reads only the
  // 8 rows of its own column from LDS, not 20 neighbours. The lanes at the
  // tile's x edges would take T from the next tile row; their sends are
  // zeroed and the frozen halo columns' terms are added to f once.
  auto from_left = [](float v) {  // value of lane x - 1
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), BX == 16 ? 0x111 : 0x138, 0xF, 0xF, true));
  };
  auto from_right = [](float v) {  // value of lane x + 1
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), BX == 16 ? 0x101 : 0x130, 0xF, 0xF, true));
  };
  static_assert(BX == 16 || BX == 32, "tile rows are one DPP row (16) or half a wave (32)");
  const bool send_left = lx != BX - 1;  // T1 goes to lane x + 1 in the same tile row
  const bool send_right = lx != 0;      // T0 goes to lane x - 1
  float fe[4][H][OTM_NRHS];
#pragma unroll
  for (int col = 0; col < 4; col++)
#pragma unroll
    for (int h = 0; h < H; h++)
#pragma unroll
      for (int k = 0; k < OTM_NRHS; k++) fe[col][h][k] = fv[col][h][k];
  if (!zero_in) {
    // halo column x0 - 1 seen by the tile's first column: its T1; halo
    // column x0 + BX seen by the last column: its T0 (frozen values)
    __shared__ float hterm[2][BY * BZ][OTM_NRHS];
    for (int i = t; i < 2 * BY * BZ; i += T::NT) {
      const int side = i / (BY * BZ), r = i % (BY * BZ);
      const int ly = r % BY, lz = r / BY;
      const int hx = side ? HX - 1 : 0, kx = side ? BX : 0;
      const int c0 = hx + DY * (ly + 1) + DZ * (lz + 1);
#pragma unroll
      for (int k = 0; k < OTM_NRHS; k++) {
        const float* U = &su[k][0];
        float T = 0.f;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int ay = q & 1, az = q >> 1, sy = 2 * ay - 1, sz = 2 * az - 1;
          const float G = U[c0 + sy * DY] + U[c0 + sz * DZ] + U[c0 + sy * DY + sz * DZ];
          T += sk[kx + KX * ((ly + ay) + KY * (lz + az))] * G;
        }
        hterm[side][r][k] = T;
      }
    }
    __syncthreads();
#pragma unroll
    for (int col = 0; col < 4; col++)
#pragma unroll
      for (int h = 0; h < H; h++) {
        const int rr = ty + TR * h;
        const int ly = (col & 1) + 2 * (rr % (BY / 2)), lz = (col >> 1) + 2 * (rr / (BY / 2));
        const int r = ly + BY * lz;
#pragma unroll
        for (int k = 0; k < OTM_NRHS; k++) {
          if (lx == 0) fe[col][h][k] += c * hterm[0][r][k];
          if (lx == BX - 1) fe[col][h][k] += c * hterm[1][r][k];
        }
      }
  }
  for (int s = 0; s < sweeps; s++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int col = forward ? i : 3 - i;
#pragma unroll
      for (int h = 0; h < H; h++) {
        const int rr = ty + TR * h;
        const int ly = (col & 1) + 2 * (rr % (BY / 2)), lz = (col >> 1) + 2 * (rr / (BY / 2));
        // kappa of the elements (x-1+ax, y-1+ay, z-1+az): K0 for ax = 0, K1 for ax = 1
        const int kb = lx + KX * (ly + KY * lz);
        float K0[4], K1[4], W[4], K0s[4], K1s[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int o = KX * ((q & 1) + KY * (q >> 1));
          K0[q] = sk[kb + o];
          K1[q] = sk[kb + 1 + o];
          W[q] = K0[q] + K1[q];
          K1s[q] = send_left ? K1[q] : 0.f;
          K0s[q] = send_right ? K0[q] : 0.f;
        }
        const float ks = (W[0] + W[1]) + (W[2] + W[3]);
        const float inv = __builtin_amdgcn_rcpf(4.f * c * ks);
        // rows of the node's own column, addressed from the lowest one so
        // that every offset (load cases included) is an LDS immediate
        int cb = (lx + 1) + DY * (ly + 1) + DZ * (lz + 1) - LO;
        asm volatile("" : "+v"(cb));
#pragma unroll
        for (int k = 0; k < OTM_NRHS; k++) {
          const float* U = &su[k][0] + cb;
          const float ym = U[LO - DY], yp = U[LO + DY], zm = U[LO - DZ], zp = U[LO + DZ];
          const float D[4] = {U[0], U[2 * DY], U[2 * DZ], U[2 * DY + 2 * DZ]};  // (y-+1, z-+1)
          const float G[4] = {ym + zm + D[0], yp + zm + D[1], ym + zp + D[2], yp + zp + D[3]};
          float T1 = K1s[0] * G[0], T0 = K0s[0] * G[0], acc = W[0] * D[0];
#pragma unroll
          for (int q = 1; q < 4; q++) {
            T1 += K1s[q] * G[q];
            T0 += K0s[q] * G[q];
            acc += W[q] * D[q];
          }
          acc += from_left(T1);
          acc += from_right(T0);
          (&su[k][0] + cb)[LO] = (fe[col][h][k] + c * acc) * inv;
        }
      }
      __syncthreads();
    }
  if (dot) {
    double acc[OTM_NRHS] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int col = 0; col < 4; col++)
#pragma unroll
      for (int h = 0; h < H; h++) {
        const int rr = ty + TR * h;
        const int ly = (col & 1) + 2 * (rr % (BY / 2)), lz = (col >> 1) + 2 * (rr / (BY / 2));
        const int c0 = (lx + 1) + DY * (ly + 1) + DZ * (lz + 1);
#pragma unroll
        for (int k = 0; k < OTM_NRHS; k++) acc[k] += (double)fv[col][h][k] * (double)su[k][c0];
      }
    reduce_block_at<OTM_NRHS>(acc, RD, b);
  }
  static_assert((BY * BZ) % TR == 0, "thread rows must divide the tile's rows");
#pragma unroll
  for (int j = 0; j < BY * BZ / TR; j++) {
    unsigned off = 0u;
#pragma unroll
    for (int q = 0; q < 64 / BX; q++) {
      const int r = wr0 + q + TR * j;
      const unsigned oq = nxu * (unsigned)(oy + r % BY) + nxyu * (unsigned)(oz + r / BY);
      off = sub == q ? oq : off;
    }
    off = 4u * (off + (unsigned)(ox + tx));
    const int r = ty + TR * j;
    const int ly = r % BY, lz = r / BY;
    const int c0 = (tx + 1) + DY * (ly + 1) + DZ * (lz + 1);
#pragma unroll
    for (int k = 0; k < OTM_NRHS; k++) *(float*)((char*)(uout + (size_t)k * (size_t)g.nn) + off) = su[k][c0];
  }
}

}  // namespace otm
