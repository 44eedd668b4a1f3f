// otm_mma.hip -- method of moving asymptotes on the GPU, FP64.
//
// Restates mmasub_ker (mmaOpt.cu:834-1019) and its primal-dual interior
// point subsolv_g (mmaOpt.cu:429-830). The reference runs every vector
// expression as its own kernel and solves the (m+1) Newton system on the
// host each step (solveLinearHost, mmaOpt.cpp:48-73). Here each Newton step
// is three fused passes over the n design variables:
//   newton: plambda/qlambda, delx, diagx, GG and the reductions that build
//           the (m+1) system (gvec, GG.(delx/diagx), GG diag(1/diagx) GG^T);
//   dir:    dx, dxsi, deta and the fraction-to-boundary bound;
//   trial:  the trial point and its KKT residual norm/max;
// each followed by a one-block finaliser that runs the scalar logic (the
// (m+1) solve by Gaussian elimination with partial pivoting, step length,
// line-search bookkeeping) on the device. m is limited to kMmaMaxM.
#include <math.h>
#include <string.h>

#include "otm_kernels.hpp"

namespace otm {

struct MmaSmall {
  int m, itto;
  int accept, fault;
  double a0, epsi;
  double a[kMmaMaxM], b[kMmaMaxM], c[kMmaMaxM], d[kMmaMaxM];
  double y[kMmaMaxM], lam[kMmaMaxM], mu[kMmaMaxM], s[kMmaMaxM], z, zet;
  double dy[kMmaMaxM], dlam[kMmaMaxM], dmu[kMmaMaxM], ds[kMmaMaxM], dz, dzet;
  double yt[kMmaMaxM], lamt[kMmaMaxM], mut[kMmaMaxM], st[kMmaMaxM], zt, zett;
  double steg, stm_small;
  double residunorm, residumax, resinew, rmaxnew;
};

static inline unsigned nblk(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// sums and maxima of one block into partials[block * (NS + NM) + ...]
template <int NS, int NM>
__device__ inline void block_sum_max(double (&s)[NS], double (&mx)[NM], double* partials) {
  __shared__ double sh[(NS + NM) * kWaves];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    double v = wave_sum(s[i]);
    if (lane == 0) sh[i * kWaves + wid] = v;
  }
#pragma unroll
  for (int i = 0; i < NM; i++) {
    double v = mx[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    if (lane == 0) sh[(NS + i) * kWaves + wid] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < NS; i++) {
      double v = 0.0;
      for (int w = 0; w < kWaves; w++) v += sh[i * kWaves + w];
      partials[(size_t)blockIdx.x * (NS + NM) + i] = v;
    }
    for (int i = 0; i < NM; i++) {
      double v = -1e300;
      for (int w = 0; w < kWaves; w++) v = fmax(v, sh[(NS + i) * kWaves + w]);
      partials[(size_t)blockIdx.x * (NS + NM) + NS + i] = v;
    }
  }
}

// one-block reduction of the partials: out[0..NS) sums, out[NS..) maxima
template <int NS, int NM>
__device__ inline void final_sum_max(const double* partials, int nblocks, double* out) {
  __shared__ double sh[(NS + NM) * kWaves];
  double s[NS + NM];
#pragma unroll
  for (int i = 0; i < NS; i++) s[i] = 0.0;
#pragma unroll
  for (int i = 0; i < NM; i++) s[NS + i] = -1e300;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
    for (int i = 0; i < NS; i++) s[i] += partials[(size_t)b * (NS + NM) + i];
#pragma unroll
    for (int i = 0; i < NM; i++) s[NS + i] = fmax(s[NS + i], partials[(size_t)b * (NS + NM) + NS + i]);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NS + NM; i++) {
    double v = s[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double w = __shfl_down(v, o, 64);
      v = i < NS ? v + w : fmax(v, w);
    }
    if (lane == 0) sh[i * kWaves + wid] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < NS + NM; i++) {
      double v = i < NS ? 0.0 : -1e300;
      for (int w = 0; w < kWaves; w++) v = i < NS ? v + sh[i * kWaves + w] : fmax(v, sh[i * kWaves + w]);
      out[i] = v;
    }
  }
}

// ---- mmasub: asymptotes, bounds, approximation coefficients ---------------
template <int M>
__global__ void __launch_bounds__(kBlock) k_mma_prep(MmaArrays A, int itn, double move) {
  constexpr double raa0 = 1e-5, albefa = 0.1, asyinit = 0.5, asyincr = 1.2, asydecr = 0.7;
  double bs[M];
#pragma unroll
  for (int i = 0; i < M; i++) bs[i] = 0.0;
  double dummy[1] = {-1e300};
  const BlockRange br = block_range(A.n);
  for (long long j = br.b + threadIdx.x; j < br.e; j += blockDim.x) {
    const double x = A.xvar[j], dxm = A.xmax[j] - A.xmin[j];
    double lo, up;
    if (itn <= 2) {
      lo = x - asyinit * dxm;
      up = x + asyinit * dxm;
    } else {
      const double x1 = A.xold1[j], x2 = A.xold2[j];
      const double zzz = (x - x1) * (x1 - x2);
      const double f = zzz > 0 ? asyincr : (zzz < 0 ? asydecr : 1.0);
      lo = x - f * (x1 - A.low[j]);
      up = x + f * (A.upp[j] - x1);
      lo = fmin(fmax(lo, x - 10.0 * dxm), x - 0.01 * dxm);
      up = fmax(fmin(up, x + 10.0 * dxm), x + 0.01 * dxm);
    }
    A.low[j] = lo;
    A.upp[j] = up;
    const double al = fmax(fmax(lo + albefa * (x - lo), x - move * dxm), A.xmin[j]);
    const double be = fmin(fmin(up - albefa * (up - x), x + move * dxm), A.xmax[j]);
    A.alfa[j] = al;
    A.beta[j] = be;
    const double xmamiinv = 1.0 / fmax(dxm, 1e-5);
    const double ux1 = up - x, xl1 = x - lo, ux2 = ux1 * ux1, xl2 = xl1 * xl1;
    const double g0 = A.df0dx[j];
    const double dp = fmax(g0, 0.0), dq = fmax(-g0, 0.0);
    const double pq0 = 1e-3 * (dp + dq) + raa0 * xmamiinv;
    A.p0[j] = (dp + pq0) * ux2;
    A.q0[j] = (dq + pq0) * xl2;
#pragma unroll
    for (int i = 0; i < M; i++) {
      const double g = A.dgdx[i * A.n + j];
      const double pp = fmax(g, 0.0), qq = fmax(-g, 0.0);
      const double pq = 1e-3 * (pp + qq) + raa0 * xmamiinv;
      const double Pi = (pp + pq) * ux2, Qi = (qq + pq) * xl2;
      A.P[i * A.n + j] = Pi;
      A.Q[i * A.n + j] = Qi;
      bs[i] += Pi / ux1 + Qi / xl1;
    }
    // subsolv start point (mmaOpt.cu:437-452)
    const double xs = 0.5 * (al + be);
    A.xs[j] = xs;
    A.xsi[j] = fmax(1.0 / (xs - al), 1.0);
    A.eta[j] = fmax(1.0 / (be - xs), 1.0);
  }
  block_sum_max<M, 1>(bs, dummy, A.partials);
}

template <int M>
__global__ void __launch_bounds__(kBlock) k_mma_prep_final(MmaArrays A, MmaSmall* S, const double* gval, int nblocks,
                                                           double a0, double av, double cv, double dv) {
  __shared__ double out[M + 1];
  final_sum_max<M, 1>(A.partials, nblocks, out);
  if (threadIdx.x != 0) return;
  S->m = M;
  S->a0 = a0;
  S->epsi = 1.0;
  for (int i = 0; i < M; i++) {
    S->a[i] = av;
    S->c[i] = cv;
    S->d[i] = dv;
    S->b[i] = out[i] - gval[i];
    S->y[i] = 1.0;
    S->lam[i] = 1.0;
    S->s[i] = 1.0;
    S->mu[i] = fmax(0.5 * cv, 1.0);
  }
  S->z = 1.0;
  S->zet = 1.0;
  S->fault = 0;
}

// ---- KKT residual at the current point or at a trial point -----------------
// Small-vector part of the residual (mmaOpt.cu:514-522); adds to ss / mx.
__device__ inline void small_residual(const MmaSmall* S, const double* gvec, const double* y, double z,
                                      const double* lam, const double* mu, double zet, const double* s, double& ss,
                                      double& mx) {
  const int m = S->m;
  const double epsi = S->epsi;
  double alam = 0.0;
  for (int i = 0; i < m; i++) alam += S->a[i] * lam[i];
  const double rez = S->a0 - zet - alam, rezet = zet * z - epsi;
  ss += rez * rez + rezet * rezet;
  mx = fmax(mx, fmax(fabs(rez), fabs(rezet)));
  for (int i = 0; i < m; i++) {
    const double rey = S->c[i] + S->d[i] * y[i] - mu[i] - lam[i];
    const double relam = gvec[i] - S->a[i] * z - y[i] + s[i] - S->b[i];
    const double remu = mu[i] * y[i] - epsi;
    const double res = lam[i] * s[i] - epsi;
    ss += rey * rey + relam * relam + remu * remu + res * res;
    mx = fmax(mx, fmax(fmax(fabs(rey), fabs(relam)), fmax(fabs(remu), fabs(res))));
  }
}

// trial != 0: the point is old + steg * direction; it is written to the
// trial buffers and the small-vector trial values to S->*t.
template <int M>
__global__ void __launch_bounds__(kBlock) k_mma_point(MmaArrays A, const MmaSmall* __restrict__ S, int trial) {
  double lam[M];
  const double epsi = S->epsi, steg = S->steg;
#pragma unroll
  for (int i = 0; i < M; i++) lam[i] = trial ? S->lam[i] + steg * S->dlam[i] : S->lam[i];
  double sums[1 + M];
#pragma unroll
  for (int i = 0; i < 1 + M; i++) sums[i] = 0.0;
  double mx[1] = {0.0};
  const BlockRange br = block_range(A.n);
  for (long long j = br.b + threadIdx.x; j < br.e; j += blockDim.x) {
    const double al = A.alfa[j], be = A.beta[j];
    double x = A.xs[j], xsi = A.xsi[j], eta = A.eta[j];
    if (trial) {
      const double dx = A.dx[j];
      const double xa = x - al, bx = be - x;
      const double dxsi = -xsi + epsi / xa - xsi * dx / xa;
      const double deta = -eta + epsi / bx + eta * dx / bx;
      x += steg * dx;
      xsi += steg * dxsi;
      eta += steg * deta;
      A.xs_t[j] = x;
      A.xsi_t[j] = xsi;
      A.eta_t[j] = eta;
    }
    const double ux1 = A.upp[j] - x, xl1 = x - A.low[j];
    double plam = A.p0[j], qlam = A.q0[j];
#pragma unroll
    for (int i = 0; i < M; i++) {
      const double Pi = A.P[i * A.n + j], Qi = A.Q[i * A.n + j];
      plam += lam[i] * Pi;
      qlam += lam[i] * Qi;
      sums[1 + i] += Pi / ux1 + Qi / xl1;
    }
    const double rex = plam / (ux1 * ux1) - qlam / (xl1 * xl1) - xsi + eta;
    const double rexsi = xsi * (x - al) - epsi;
    const double reeta = eta * (be - x) - epsi;
    sums[0] += rex * rex + rexsi * rexsi + reeta * reeta;
    mx[0] = fmax(mx[0], fmax(fabs(rex), fmax(fabs(rexsi), fabs(reeta))));
  }
  block_sum_max<1 + M, 1>(sums, mx, A.partials);
}

template <int M>
__global__ void __launch_bounds__(kBlock) k_mma_point_final(MmaArrays A, MmaSmall* S, int nblocks, int trial) {
  __shared__ double out[M + 2];
  final_sum_max<1 + M, 1>(A.partials, nblocks, out);
  if (threadIdx.x != 0) return;
  double ss = out[0], mx = out[1 + M];
  const double steg = S->steg;
  if (!trial) {
    small_residual(S, out + 1, S->y, S->z, S->lam, S->mu, S->zet, S->s, ss, mx);
    S->residunorm = sqrt(ss);
    S->residumax = mx;
    return;
  }
  for (int i = 0; i < M; i++) {
    S->yt[i] = S->y[i] + steg * S->dy[i];
    S->lamt[i] = S->lam[i] + steg * S->dlam[i];
    S->mut[i] = S->mu[i] + steg * S->dmu[i];
    S->st[i] = S->s[i] + steg * S->ds[i];
  }
  S->zt = S->z + steg * S->dz;
  S->zett = S->zet + steg * S->dzet;
  small_residual(S, out + 1, S->yt, S->zt, S->lamt, S->mut, S->zett, S->st, ss, mx);
  S->resinew = sqrt(ss);
  S->rmaxnew = mx;
  S->itto += 1;
  // line search rule (mmaOpt.cu:718-805): accept when the norm does not
  // grow, or after 50 halvings
  S->accept = (!(S->resinew > S->residunorm) || S->itto >= 50) ? 1 : 0;
  if (!(S->resinew - S->resinew == 0.0)) S->fault = 1;
  S->steg = steg / 2;
}

// After acceptance the trial becomes the current point.
__global__ void k_mma_commit(MmaSmall* S) {
  for (int i = 0; i < S->m; i++) {
    S->y[i] = S->yt[i];
    S->lam[i] = S->lamt[i];
    S->mu[i] = S->mut[i];
    S->s[i] = S->st[i];
  }
  S->z = S->zt;
  S->zet = S->zett;
  S->residunorm = S->resinew;
  S->residumax = S->rmaxnew;
}

// ---- Newton system ---------------------------------------------------------------
template <int M>
struct NewtonNV {
  static constexpr int NS = 2 * M + (M * (M + 1)) / 2;
};

template <int M>
__global__ void __launch_bounds__(kBlock) k_mma_newton(MmaArrays A, const MmaSmall* __restrict__ S) {
  constexpr int NS = NewtonNV<M>::NS;
  double lam[M];
#pragma unroll
  for (int i = 0; i < M; i++) lam[i] = S->lam[i];
  const double epsi = S->epsi;
  double sums[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) sums[i] = 0.0;
  double dummy[1] = {-1e300};
  const BlockRange br = block_range(A.n);
  for (long long j = br.b + threadIdx.x; j < br.e; j += blockDim.x) {
    const double x = A.xs[j], xsi = A.xsi[j], eta = A.eta[j];
    const double xa = x - A.alfa[j], bx = A.beta[j] - x;
    const double ux1 = A.upp[j] - x, xl1 = x - A.low[j];
    const double ux2 = ux1 * ux1, xl2 = xl1 * xl1;
    double plam = A.p0[j], qlam = A.q0[j];
    double GG[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
      const double Pi = A.P[i * A.n + j], Qi = A.Q[i * A.n + j];
      plam += lam[i] * Pi;
      qlam += lam[i] * Qi;
      sums[i] += Pi / ux1 + Qi / xl1;  // gvec
      GG[i] = Pi / ux2 - Qi / xl2;
    }
    const double dpsidx = plam / ux2 - qlam / xl2;
    const double delx = dpsidx - epsi / xa + epsi / bx;
    const double diagx = 2.0 * (plam / (ux2 * ux1) + qlam / (xl2 * xl1)) + xsi / xa + eta / bx;
    const double inv = 1.0 / diagx;
    int t = 2 * M;
#pragma unroll
    for (int i = 0; i < M; i++) {
      sums[M + i] += GG[i] * delx * inv;
#pragma unroll
      for (int k = i; k < M; k++) sums[t++] += GG[i] * GG[k] * inv;
    }
  }
  block_sum_max<NS, 1>(sums, dummy, A.partials);
}

// Solves [Alam a; a^T -zet/z] [dlam; dz] = [blam; delz] and the small
// directions (mmaOpt.cu:595-678).
template <int M>
__global__ void __launch_bounds__(kBlock) k_mma_newton_final(MmaArrays A, MmaSmall* S, int nblocks) {
  constexpr int NS = NewtonNV<M>::NS;
  __shared__ double out[NS + 1];
  final_sum_max<NS, 1>(A.partials, nblocks, out);
  if (threadIdx.x != 0) return;
  const double epsi = S->epsi;
  const double* gvec = out;
  const double* GGdd = out + M;
  double Alam[M][M];
  int t = 2 * M;
  for (int i = 0; i < M; i++)
    for (int k = i; k < M; k++) {
      Alam[i][k] = out[t];
      Alam[k][i] = out[t];
      t++;
    }
  constexpr int K = M + 1;
  double AA[K][K], bb[K];
  double alam = 0.0;
  for (int i = 0; i < M; i++) alam += S->a[i] * S->lam[i];
  for (int i = 0; i < M; i++) {
    const double dely = S->c[i] + S->d[i] * S->y[i] - S->lam[i] - epsi / S->y[i];
    const double dellam = gvec[i] - S->a[i] * S->z - S->y[i] - S->b[i] + epsi / S->lam[i];
    const double diagy = S->d[i] + S->mu[i] / S->y[i];
    const double diaglamyi = S->s[i] / S->lam[i] + 1.0 / diagy;
    bb[i] = dellam + dely / diagy - GGdd[i];
    for (int k = 0; k < M; k++) AA[i][k] = Alam[i][k] + (i == k ? diaglamyi : 0.0);
    AA[i][M] = S->a[i];
    AA[M][i] = S->a[i];
  }
  AA[M][M] = -S->zet / S->z;
  bb[M] = S->a0 - alam - epsi / S->z;
  // Gaussian elimination with partial pivoting
  for (int c = 0; c < K; c++) {
    int p = c;
    for (int r = c + 1; r < K; r++)
      if (fabs(AA[r][c]) > fabs(AA[p][c])) p = r;
    if (p != c) {
      for (int j = 0; j < K; j++) {
        const double tmp = AA[c][j];
        AA[c][j] = AA[p][j];
        AA[p][j] = tmp;
      }
      const double tmp = bb[c];
      bb[c] = bb[p];
      bb[p] = tmp;
    }
    for (int r = c + 1; r < K; r++) {
      const double f = AA[r][c] / AA[c][c];
      for (int j = c; j < K; j++) AA[r][j] -= f * AA[c][j];
      bb[r] -= f * bb[c];
    }
  }
  for (int r = K - 1; r >= 0; r--) {
    double sacc = bb[r];
    for (int j = r + 1; j < K; j++) sacc -= AA[r][j] * bb[j];
    bb[r] = sacc / AA[r][r];
  }
  double stm = 1.0;
  for (int i = 0; i < M; i++) {
    const double dlam = bb[i];
    const double dely = S->c[i] + S->d[i] * S->y[i] - S->lam[i] - epsi / S->y[i];
    const double diagy = S->d[i] + S->mu[i] / S->y[i];
    const double dy = -dely / diagy + dlam / diagy;
    const double dmu = -S->mu[i] + epsi / S->y[i] - S->mu[i] * dy / S->y[i];
    const double ds = -S->s[i] + epsi / S->lam[i] - S->s[i] * dlam / S->lam[i];
    S->dlam[i] = dlam;
    S->dy[i] = dy;
    S->dmu[i] = dmu;
    S->ds[i] = ds;
    stm = fmax(stm, fmax(fmax(-1.01 * dy / S->y[i], -1.01 * dlam / S->lam[i]),
                         fmax(-1.01 * dmu / S->mu[i], -1.01 * ds / S->s[i])));
  }
  const double dz = bb[M];
  const double dzet = -S->zet + epsi / S->z - S->zet * dz / S->z;
  S->dz = dz;
  S->dzet = dzet;
  stm = fmax(stm, fmax(-1.01 * dz / S->z, -1.01 * dzet / S->zet));
  S->stm_small = stm;
  if (!(stm - stm == 0.0)) S->fault = 1;
}

// ---- direction and fraction-to-boundary bound ------------------------------------
template <int M>
__global__ void __launch_bounds__(kBlock) k_mma_dir(MmaArrays A, const MmaSmall* __restrict__ S) {
  double lam[M], dlam[M];
#pragma unroll
  for (int i = 0; i < M; i++) {
    lam[i] = S->lam[i];
    dlam[i] = S->dlam[i];
  }
  const double epsi = S->epsi;
  double dummy[1] = {0.0};
  double mx[1] = {-1e300};
  const BlockRange br = block_range(A.n);
  for (long long j = br.b + threadIdx.x; j < br.e; j += blockDim.x) {
    const double x = A.xs[j], xsi = A.xsi[j], eta = A.eta[j];
    const double xa = x - A.alfa[j], bx = A.beta[j] - x;
    const double ux1 = A.upp[j] - x, xl1 = x - A.low[j];
    const double ux2 = ux1 * ux1, xl2 = xl1 * xl1;
    double plam = A.p0[j], qlam = A.q0[j], gd = 0.0;
#pragma unroll
    for (int i = 0; i < M; i++) {
      const double Pi = A.P[i * A.n + j], Qi = A.Q[i * A.n + j];
      plam += lam[i] * Pi;
      qlam += lam[i] * Qi;
      gd += (Pi / ux2 - Qi / xl2) * dlam[i];
    }
    const double dpsidx = plam / ux2 - qlam / xl2;
    const double delx = dpsidx - epsi / xa + epsi / bx;
    const double diagx = 2.0 * (plam / (ux2 * ux1) + qlam / (xl2 * xl1)) + xsi / xa + eta / bx;
    const double dx = -delx / diagx - gd / diagx;
    A.dx[j] = dx;
    const double dxsi = -xsi + epsi / xa - xsi * dx / xa;
    const double deta = -eta + epsi / bx + eta * dx / bx;
    mx[0] = fmax(mx[0], fmax(fmax(-1.01 * dxsi / xsi, -1.01 * deta / eta), fmax(-1.01 * dx / xa, 1.01 * dx / bx)));
  }
  block_sum_max<1, 1>(dummy, mx, A.partials);
}

__global__ void __launch_bounds__(kBlock) k_mma_dir_final(MmaArrays A, MmaSmall* S, int nblocks) {
  __shared__ double out[2];
  final_sum_max<1, 1>(A.partials, nblocks, out);
  if (threadIdx.x != 0) return;
  const double stm = fmax(S->stm_small, out[1]);
  S->steg = 1.0 / stm;
  S->itto = 0;
  S->accept = 0;
}

__global__ void k_mma_epsi(MmaSmall* S) { S->epsi *= 0.1; }

// xold2 = xold1, xold1 = x, x = subsolver solution (mmaOpt.cu:1055-1057),
// also the FP32 copy for the caller
__global__ void __launch_bounds__(kBlock) k_mma_finish(MmaArrays A, float* xout) {
  const long long j = gtid();
  if (j >= A.n) return;
  const double xn = A.xs[j];
  A.xold2[j] = A.xold1[j];
  A.xold1[j] = A.xvar[j];
  A.xvar[j] = xn;
  if (xout) xout[j] = (float)xn;
}

// ---- host driver ---------------------------------------------------------------
template <int M>
static int mma_run(hipStream_t s, MmaArrays A, MmaSmall* S, MmaSmall* Sh, const double* gval_dev, int itn,
                   double move, double a0, double av, double cv, double dv, float* xout, double epsimin) {
  const unsigned nb = red_blocks_for_mma(A.n);
  auto read = [&]() -> int {
    if (hipMemcpyAsync(Sh, S, sizeof(MmaSmall), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    return Sh->fault ? -2 : 0;
  };
  hipLaunchKernelGGL(k_mma_prep<M>, dim3(nb), dim3(kBlock), 0, s, A, itn, move);
  hipLaunchKernelGGL(k_mma_prep_final<M>, dim3(1), dim3(kBlock), 0, s, A, S, gval_dev, (int)nb, a0, av, cv, dv);
  int newton = 0;
  double epsi = 1.0;
  while (epsi > epsimin) {
    hipLaunchKernelGGL(k_mma_point<M>, dim3(nb), dim3(kBlock), 0, s, A, S, 0);
    hipLaunchKernelGGL(k_mma_point_final<M>, dim3(1), dim3(kBlock), 0, s, A, S, (int)nb, 0);
    int rc = read();
    if (rc) return rc;
    int ittt = 0;
    while (Sh->residumax > 0.9 * epsi && ittt < 200) {
      ittt++;
      newton++;
      hipLaunchKernelGGL(k_mma_newton<M>, dim3(nb), dim3(kBlock), 0, s, A, S);
      hipLaunchKernelGGL(k_mma_newton_final<M>, dim3(1), dim3(kBlock), 0, s, A, S, (int)nb);
      hipLaunchKernelGGL(k_mma_dir<M>, dim3(nb), dim3(kBlock), 0, s, A, S);
      hipLaunchKernelGGL(k_mma_dir_final, dim3(1), dim3(kBlock), 0, s, A, S, (int)nb);
      do {
        hipLaunchKernelGGL(k_mma_point<M>, dim3(nb), dim3(kBlock), 0, s, A, S, 1);
        hipLaunchKernelGGL(k_mma_point_final<M>, dim3(1), dim3(kBlock), 0, s, A, S, (int)nb, 1);
        rc = read();
        if (rc) return rc;
      } while (!Sh->accept);
      hipLaunchKernelGGL(k_mma_commit, dim3(1), dim3(1), 0, s, S);
      std::swap(A.xs, A.xs_t);
      std::swap(A.xsi, A.xsi_t);
      std::swap(A.eta, A.eta_t);
      Sh->residumax = Sh->rmaxnew;
    }
    hipLaunchKernelGGL(k_mma_epsi, dim3(1), dim3(1), 0, s, S);
    epsi *= 0.1;
  }
  hipLaunchKernelGGL(k_mma_finish, dim3(nblk(A.n)), dim3(kBlock), 0, s, A, xout);
  if (hipGetLastError() != hipSuccess) return -1;
  return newton;
}

int mma_update(hipStream_t s, int m, MmaArrays& A, void* small_dev, void* small_host, const double* gval_dev,
               int itn, double move, double a0, double av, double cv, double dv, float* xout) {
  MmaSmall* S = (MmaSmall*)small_dev;
  MmaSmall* Sh = (MmaSmall*)small_host;
  MmaArrays B = A;  // the driver swaps the subsolver buffers locally
  int rc;
  switch (m) {
    case 1: rc = mma_run<1>(s, B, S, Sh, gval_dev, itn, move, a0, av, cv, dv, xout, 1e-7); break;
    case 2: rc = mma_run<2>(s, B, S, Sh, gval_dev, itn, move, a0, av, cv, dv, xout, 1e-7); break;
    case 3: rc = mma_run<3>(s, B, S, Sh, gval_dev, itn, move, a0, av, cv, dv, xout, 1e-7); break;
    case 4: rc = mma_run<4>(s, B, S, Sh, gval_dev, itn, move, a0, av, cv, dv, xout, 1e-7); break;
    default: rc = -3;
  }
  return rc;
}

size_t mma_small_bytes() { return sizeof(MmaSmall); }

}  // namespace otm
