// otm_coarse.hip -- coarse multigrid levels: stored 27-point stencils,
// transfers, Galerkin assembly and the coarsest dense FP64 solve.
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
#include "otm_kernels.hpp"

namespace otm {

static inline unsigned nblk(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

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
  const int x = (int)(t % g.nx);
  const long long r = t / g.nx;
  const int y = (int)(r % g.ny), z = (int)(r / g.ny);
  const int xs[3] = {x == 0 ? g.nx - 1 : x - 1, x, x == g.nx - 1 ? 0 : x + 1};
  const int ys[3] = {y == 0 ? g.ny - 1 : y - 1, y, y == g.ny - 1 ? 0 : y + 1};
  const int zs[3] = {z == 0 ? g.nz - 1 : z - 1, z, z == g.nz - 1 ? 0 : z + 1};
  float a[27];
  load_stencil<HALF>(S, g.nn, t, a);
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) {
    const float* uk = u + k * g.nn;
    float s = 0.f;
#pragma unroll
    for (int o = 0; o < 27; o++) s += a[o] * uk[lin(g, xs[o % 3], ys[(o / 3) % 3], zs[o / 9])];
    res[k * g.nn + t] = f[k * g.nn + t] - s;
  }
}

// f_c(I) = sum_{d in {-1,0,1}^3} w(d) res(2I + d), w = prod(d_i == 0 ? 1 : 1/2)
// (restrict_residual_kernel_H, grid.cu:811-878, weights (2-|d|)/2 per axis).
__device__ inline void restrict_node(const Geom& fg, const Geom& cg, const float* res, float* fc, long long t) {
  const int X = (int)(t % cg.nx);
  const long long r = t / cg.nx;
  const int Y = (int)(r % cg.ny), Z = (int)(r / cg.ny);
  const int x = 2 * X, y = 2 * Y, z = 2 * Z;
  const int xs[3] = {x == 0 ? fg.nx - 1 : x - 1, x, x + 1};
  const int ys[3] = {y == 0 ? fg.ny - 1 : y - 1, y, y + 1};
  const int zs[3] = {z == 0 ? fg.nz - 1 : z - 1, z, z + 1};
  const float w1[3] = {0.5f, 1.f, 0.5f};
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) {
    const float* rk = res + k * fg.nn;
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        float row = 0.f;
        const long long rb = (long long)fg.nx * ((long long)ys[b] + (long long)fg.ny * zs[c]);
#pragma unroll
        for (int a = 0; a < 3; a++) row += w1[a] * rk[rb + xs[a]];
        s += w1[b] * w1[c] * row;
      }
    fc[k * cg.nn + t] = s;
  }
}

// u_f(i) += sum_I P(i, I) u_c(I), trilinear (prolongate_correction_kernel_H,
// grid.cu:881-940).
__device__ inline void prolong_node(const Geom& fg, const Geom& cg, const float* uc, float* uf, long long t) {
  const int x = (int)(t % fg.nx);
  const long long r = t / fg.nx;
  const int y = (int)(r % fg.ny), z = (int)(r / fg.ny);
  int X[2], Y[2], Z[2];
  float wx[2], wy[2], wz[2];
  X[0] = x >> 1; X[1] = (x & 1) ? ((x >> 1) + 1 == cg.nx ? 0 : (x >> 1) + 1) : X[0];
  wx[0] = (x & 1) ? 0.5f : 1.f; wx[1] = (x & 1) ? 0.5f : 0.f;
  Y[0] = y >> 1; Y[1] = (y & 1) ? ((y >> 1) + 1 == cg.ny ? 0 : (y >> 1) + 1) : Y[0];
  wy[0] = (y & 1) ? 0.5f : 1.f; wy[1] = (y & 1) ? 0.5f : 0.f;
  Z[0] = z >> 1; Z[1] = (z & 1) ? ((z >> 1) + 1 == cg.nz ? 0 : (z >> 1) + 1) : Z[0];
  wz[0] = (z & 1) ? 0.5f : 1.f; wz[1] = (z & 1) ? 0.5f : 0.f;
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) {
    const float* ck = uc + k * cg.nn;
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int a = 0; a < 2; a++) {
          const float w = wx[a] * wy[b] * wz[c];
          if (w != 0.f) s += w * ck[lin(cg, X[a], Y[b], Z[c])];
        }
    uf[k * fg.nn + t] += s;
  }
}

// ---- per-level kernels ----
// GS update of one of 8 colours (parity of x, y, z), colour = cx + 2cy + 4cz.
template <bool HALF>
__global__ void __launch_bounds__(kBlock) k_st_gs_color(Geom g, StRef S, const float* __restrict__ f, float* u, int cx,
                                                        int cy, int cz) {
  const int hx = g.nx >> 1, hy = g.ny >> 1, hz = g.nz >> 1;
  const long long cnt = (long long)hx * hy * hz;
  const long long t = gtid();
  if (t >= cnt) return;
  const int x = 2 * (int)(t % hx) + cx;
  const long long r = t / hx;
  const int y = 2 * (int)(r % hy) + cy, z = 2 * (int)(r / hy) + cz;
  st_gs_node<HALF>(g, S, f, u, x, y, z);
}

template <bool HALF>
__global__ void __launch_bounds__(kBlock) k_st_residual(Geom g, StRef S, const float* __restrict__ f,
                                                        const float* __restrict__ u, float* __restrict__ res) {
  const long long t = gtid();
  if (t >= g.nn) return;
  st_resid_node<HALF>(g, S, f, u, res, t);
}

// Largest |coupling| of a level (as float bits: non-negative floats order
// like their bit patterns), for the FP16 scale.
__global__ void __launch_bounds__(kBlock) k_stencil_absmax(Geom g, const float* __restrict__ st,
                                                           unsigned int* __restrict__ bits) {
  float m = 0.f;
  for (long long v = gtid(); v < g.nn; v += (long long)gridDim.x * blockDim.x)
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
__global__ void __launch_bounds__(kBlock) k_stencil_pack(Geom g, const float* __restrict__ st,
                                                         const float* __restrict__ scale, _Float16* __restrict__ st16) {
  const long long t = gtid();
  if (t >= g.nn) return;
  const int x = (int)(t % g.nx);
  const long long r = t / g.nx;
  const int y = (int)(r % g.ny), z = (int)(r / g.ny);
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

__global__ void __launch_bounds__(kBlock) k_restrict(Geom fg, Geom cg, const float* __restrict__ res,
                                                     float* __restrict__ fc) {
  const long long t = gtid();
  if (t >= cg.nn) return;
  restrict_node(fg, cg, res, fc, t);
}

__global__ void __launch_bounds__(kBlock) k_prolong_add(Geom fg, Geom cg, const float* __restrict__ uc,
                                                        float* __restrict__ uf) {
  const long long t = gtid();
  if (t >= fg.nn) return;
  prolong_node(fg, cg, uc, uf, t);
}

// Level-1 element matrices from kappa: E(E) = sum_c kappa_c G[c].
__global__ void __launch_bounds__(kBlock) k_galerkin_first(Geom fg, Geom cg, GalerkinConsts G,
                                                           const float* __restrict__ kappa, float* __restrict__ E) {
  const long long t = gtid();
  if (t >= cg.nn) return;
  const int X = (int)(t % cg.nx);
  const long long r = t / cg.nx;
  const int Y = (int)(r % cg.ny), Z = (int)(r / cg.ny);
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
__global__ void __launch_bounds__(kBlock) k_galerkin_next(Geom fg, Geom cg, const float* __restrict__ Ein,
                                                          float* __restrict__ Eout) {
  const long long t = gtid();
  if (t >= cg.nn) return;
  const int X = (int)(t % cg.nx);
  const long long r = t / cg.nx;
  const int Y = (int)(r % cg.ny), Z = (int)(r / cg.ny);
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
__global__ void __launch_bounds__(kBlock) k_stencil_from_elements(Geom g, const float* __restrict__ E,
                                                                  float* __restrict__ st) {
  const long long t = gtid();
  if (t >= g.nn) return;
  const int x = (int)(t % g.nx);
  const long long r = t / g.nx;
  const int y = (int)(r % g.ny), z = (int)(r / g.ny);
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

// n <= 128: in-place Gauss-Jordan inversion in LDS. The matrix is SPD, so no
// pivoting is needed; every elimination step updates all n^2 entries in
// parallel between two barriers (2n barriers in total instead of the long
// dependent chains of a triangular inverse).
__global__ void __launch_bounds__(kFactorThreads) k_coarsest_invert_lds(Geom g, const float* __restrict__ st,
                                                                        double* __restrict__ Binv) {
  extern __shared__ double A[];
  __shared__ double colk[kInvertLdsMax], rowk[kInvertLdsMax];
  __shared__ double red[kFactorThreads / 64];
  __shared__ double alpha_sh;
  const int n = (int)g.nn;
  const int tid = threadIdx.x, bs = blockDim.x;
  coarsest_assemble(g, st, A, red, &alpha_sh);
  const int rstep = bs / n;  // rows updated per pass (n <= bs)
  const int j = tid % n, i0 = tid / n;
  for (int k = 0; k < n; k++) {
    const double ip = 1.0 / A[k * n + k];
    for (int i = tid; i < n; i += bs) {
      colk[i] = A[i * n + k];
      rowk[i] = (i == k ? 1.0 : A[k * n + i]) * ip;
    }
    __syncthreads();
    if (i0 < rstep)
      for (int i = i0; i < n; i += rstep) {
        const int q = i * n + j;
        A[q] = (i == k) ? rowk[j] : ((j == k ? 0.0 : A[q]) - colk[i] * rowk[j]);
      }
    __syncthreads();
  }
  for (int q = tid; q < n * n; q += bs) {
    const int a = q / n, b = q - a * n;
    Binv[q] = 0.5 * (A[a * n + b] + A[b * n + a]);
  }
}

// Larger coarsest levels (n <= 4096, work matrices in global memory):
// right-looking Cholesky B = L L^T, X = L^{-1} row by row, Binv = X^T X.
__global__ void __launch_bounds__(kFactorThreads) k_coarsest_factor(Geom g, const float* __restrict__ st, double* B,
                                                                    double* Xm, double* Binv) {
  __shared__ double red[kFactorThreads / 64];
  __shared__ double alpha_sh;
  const int n = (int)g.nn;
  const int tid = threadIdx.x, bs = blockDim.x;
  coarsest_assemble(g, st, B, red, &alpha_sh);
  for (int k = 0; k < n; k++) {
    if (tid == 0) B[k * n + k] = sqrt(B[k * n + k]);
    __syncthreads();
    const double dk = B[k * n + k];
    for (int i = k + 1 + tid; i < n; i += bs) B[i * n + k] /= dk;
    __syncthreads();
    const int m = n - k - 1;
    for (int q = tid; q < m * m; q += bs) {
      const int i = k + 1 + q / m, jj = k + 1 + q % m;
      if (jj <= i) B[i * n + jj] -= B[i * n + k] * B[jj * n + k];
    }
    __syncthreads();
  }
  for (int k = 0; k < n; k++) {
    for (int jj = tid; jj < n; jj += bs) {
      double s = (jj == k) ? 1.0 : 0.0;
      if (jj < k)
        for (int m2 = jj; m2 < k; m2++) s -= B[k * n + m2] * Xm[m2 * n + jj];
      Xm[k * n + jj] = jj <= k ? s / B[k * n + k] : 0.0;
    }
    __syncthreads();
  }
  for (int q = tid; q < n * n; q += bs) {
    const int i = q / n, jj = q - i * n;
    const int m0 = i > jj ? i : jj;
    double s = 0.0;
    for (int m2 = m0; m2 < n; m2++) s += Xm[m2 * n + i] * Xm[m2 * n + jj];
    Binv[q] = s;
  }
}

// u = Binv f per right-hand side, FP64 accumulation.
__global__ void __launch_bounds__(kBlock) k_coarsest_solve(Geom g, const double* __restrict__ Binv,
                                                           const float* __restrict__ f, float* __restrict__ u) {
  const int n = (int)g.nn;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= n) return;
  const float* fk = f + (long long)k * n;
  double s = 0.0;
  for (int j = 0; j < n; j++) s += Binv[(long long)i * n + j] * (double)fk[j];
  u[(long long)k * n + i] = (float)s;
}

// Brick smoother for stored-stencil levels whose edges are multiples of 8.
// One 512-thread workgroup owns an 8^3 brick: its nodes' stencils sit in
// registers, u of the brick plus a one-node halo sits in LDS. All colours
// of `sweeps` Gauss-Seidel sweeps run inside the brick with workgroup
// barriers; halo values stay frozen at their values on entry (block
// Gauss-Seidel across bricks, exact multicolour Gauss-Seidel inside). The
// result goes to a second buffer so bricks never read a neighbour's new
// values. The post-smoother runs the colours in reverse order, which is the
// adjoint of the pre-smoother: the V-cycle stays symmetric.
constexpr int kBrick = 8;
constexpr int kBrickThreads = kBrick * kBrick * kBrick;
constexpr int kHalo = kBrick + 2;
template <bool HALF>
__global__ void __launch_bounds__(kBrickThreads) k_st_smooth_brick(Geom g, StRef S, const float* __restrict__ f,
                                                                   const float* __restrict__ uin,
                                                                   float* __restrict__ uout, int sweeps,
                                                                   int forward, int zero_in) {
  __shared__ float su[OTM_NRHS][kHalo * kHalo * kHalo];
  const int bx = g.nx / kBrick, by = g.ny / kBrick;
  const int b = xcd_block();
  const int ox = (b % bx) * kBrick, oy = ((b / bx) % by) * kBrick, oz = (b / (bx * by)) * kBrick;
  const int t = threadIdx.x;
  const int lx = t & 7, ly = (t >> 3) & 7, lz = t >> 6;
  const int x = ox + lx, y = oy + ly, z = oz + lz;
  const int nn = (int)g.nn;
  const int v = x + g.nx * (y + g.ny * z);
  // halo load (or zero initial guess)
  for (int i = t; i < kHalo * kHalo * kHalo; i += kBrickThreads) {
    const int hx = i % kHalo, hy = (i / kHalo) % kHalo, hz = i / (kHalo * kHalo);
    if (zero_in) {
#pragma unroll
      for (int k = 0; k < OTM_NRHS; k++) su[k][i] = 0.f;
    } else {
      const int X = wrapm(ox + hx - 1, g.nx), Y = wrapm(oy + hy - 1, g.ny), Z = wrapm(oz + hz - 1, g.nz);
      const int w = X + g.nx * (Y + g.ny * Z);
#pragma unroll
      for (int k = 0; k < OTM_NRHS; k++) su[k][i] = uin[k * (long long)nn + w];
    }
  }
  float a[27];
  load_stencil<HALF>(S, nn, v, a);
  float fv[OTM_NRHS];
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) fv[k] = f[k * (long long)nn + v];
  const float inv = 1.f / a[13];
  const int mycol = (x & 1) + 2 * (y & 1) + 4 * (z & 1);
  const int c0 = (lx + 1) + kHalo * ((ly + 1) + kHalo * (lz + 1));
  __syncthreads();
  for (int s = 0; s < sweeps; s++)
    for (int i = 0; i < 8; i++) {
      const int col = forward ? i : 7 - i;
      if (col == mycol) {
#pragma unroll
        for (int k = 0; k < OTM_NRHS; k++) {
          float acc = 0.f;
#pragma unroll
          for (int o = 0; o < 27; o++) {
            if (o == 13) continue;
            const int d = (o % 3 - 1) + kHalo * (((o / 3) % 3 - 1) + kHalo * (o / 9 - 1));
            acc += a[o] * su[k][c0 + d];
          }
          su[k][c0] = (fv[k] - acc) * inv;
        }
      }
      __syncthreads();
    }
#pragma unroll
  for (int k = 0; k < OTM_NRHS; k++) uout[k * (long long)nn + v] = su[k][c0];
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
void launch_coarsest_factor(hipStream_t s, const Geom& g, const float* st, double* B, double* X, double* Binv) {
  const int n = (int)g.nn;
  if (n <= kInvertLdsMax) {
    // per device and cheap; set on every call (once per design update)
    (void)hipFuncSetAttribute((const void*)k_coarsest_invert_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(kInvertLdsMax * kInvertLdsMax * sizeof(double)));
    int nt = ((n * n + 63) / 64) * 64;
    if (nt > kFactorThreads) nt = kFactorThreads;
    hipLaunchKernelGGL(k_coarsest_invert_lds, dim3(1), dim3(nt), (size_t)n * n * sizeof(double), s, g, st, Binv);
  } else {
    hipLaunchKernelGGL(k_coarsest_factor, dim3(1), dim3(kFactorThreads), 0, s, g, st, B, X, Binv);
  }
}
bool brick_ok(const Geom& g) { return g.nx % kBrick == 0 && g.ny % kBrick == 0 && g.nz % kBrick == 0; }
void launch_st_smooth_brick(hipStream_t s, const Geom& g, const StRef& S, const float* f, const float* uin,
                            float* uout, int sweeps, bool forward, bool zero_in) {
  const unsigned nb = (unsigned)((g.nx / kBrick) * (g.ny / kBrick) * (g.nz / kBrick));
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
