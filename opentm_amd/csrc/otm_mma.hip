// otm_mma.hip -- method of moving asymptotes on the GPU, FP64.
//
// Restates mmasub_ker (mmaOpt.cu:834-1019) and its primal-dual interior
// point subsolv_g (mmaOpt.cu:429-830). The reference runs every vector
// expression as its own kernel, keeps the subproblem coefficients (alfa,
// beta, p0, q0, P, Q) as FP64 arrays and solves the (m+1) Newton system on
// the host each step (solveLinearHost, mmaOpt.cpp:48-73).
//
// Every pass over the n design variables is bound by HBM bandwidth, so its
// time is its bytes per element. The coefficients are therefore not stored:
// each pass recomputes them from the FP32 design x, the FP32 gradients (the
// reference casts both to FP64, which is exact) and the FP64 asymptotes, for
// about 60 FP64 operations per element instead of 8 (4 + 2m) bytes of
// reads. A pass reads 4 + 16 + 4 (1 + m) + 24 bytes per element (x, low and
// upp, gradients, the interior point; per-element bounds add 8). Each Newton
// step is two passes:
//   dir:   the Newton direction dx from the small system's dlam and the
//          fraction-to-boundary bound (dx is not stored);
//   point: the trial point (dx recomputed by the same function as in dir,
//          so the same bits), its KKT residual norm/max and, speculatively,
//          the sums that build the Newton system at that point (gvec,
//          GG.(delx/diagx), GG diag(1/diagx) GG^T). When the line search
//          accepts the point these are the next step's system; when it
//          rejects it they are dropped.
// Each pass is followed by a one-block finaliser for the scalar logic (the
// (m+1) solve by Gaussian elimination with partial pivoting, step length,
// line-search bookkeeping) on the device. m is limited to kMmaMaxM.
#include <math.h>

#include <algorithm>
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
  double lamtr[kMmaMaxM];  // lam + steg dlam of the next trial (uniform for its pass)
  int respec;              // the last trial's step bound was wrong: run it again at steg
  double residunorm, residumax, resinew, rmaxnew;
  // Newton-system sums at the current point and at the trial point:
  // gvec[m], GG.(delx/diagx)[m], upper triangle of GG diag(1/diagx) GG^T
  double ns[kMmaMaxNS], ns_t[kMmaMaxNS];
  // the trial point's residual and Newton sums at the next barrier level
  // (epsi / 10): the start of that level when the trial ends this one
  double ns_n[kMmaMaxNS], resn, rmaxn;
};

static inline unsigned nblk(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Sums and maxima of one block into the partials, value-major:
// partials[v * gridDim.x + block] for the NS sums then the NM maxima, so a
// finaliser reads each value's partials contiguously.
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
  const size_t nb = gridDim.x;
  if (threadIdx.x < NS) {
    double v = 0.0;
    for (int w = 0; w < kWaves; w++) v += sh[threadIdx.x * kWaves + w];
    partials[threadIdx.x * nb + blockIdx.x] = v;
  } else if (threadIdx.x < NS + NM) {
    double v = -1e300;
    for (int w = 0; w < kWaves; w++) v = fmax(v, sh[threadIdx.x * kWaves + w]);
    partials[threadIdx.x * nb + blockIdx.x] = v;
  }
}

// One-block reduction of the value-major partials of nblocks blocks:
// out[0..NS) sums (in block order per thread, then a fixed tree), out[NS..)
// maxima.
template <int NS, int NM>
__device__ inline void final_sum_max(const double* partials, int nblocks, double* out) {
  constexpr int NV = NS + NM;
  __shared__ double sh[NV * kWaves];
  double s[NV];
#pragma unroll
  for (int i = 0; i < NS; i++) s[i] = 0.0;
#pragma unroll
  for (int i = 0; i < NM; i++) s[NS + i] = -1e300;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
    for (int i = 0; i < NS; i++) s[i] += partials[(size_t)i * nblocks + b];
#pragma unroll
    for (int i = 0; i < NM; i++) s[NS + i] = fmax(s[NS + i], partials[(size_t)(NS + i) * nblocks + b]);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    const double v = wave_sum(s[i]);
    if (lane == 0) sh[i * kWaves + wid] = v;
  }
#pragma unroll
  for (int i = NS; i < NV; i++) {
    double v = s[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    if (lane == 0) sh[i * kWaves + wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int i = threadIdx.x;
    double v = i < NS ? 0.0 : -1e300;
    for (int w = 0; w < kWaves; w++) v = i < NS ? v + sh[i * kWaves + w] : fmax(v, sh[i * kWaves + w]);
    out[i] = v;
  }
  __syncthreads();
}

// ---- per-element subproblem data (mmaOpt.cu:870-1000) ----------------------
constexpr double kRaa0 = 1e-5, kAlbefa = 0.1, kAsyinit = 0.5, kAsyincr = 1.2, kAsydecr = 0.7;

// 1/v for finite, normal v: the hardware estimate refined by two
// Newton-Raphson steps, r += r (1 - v r) (error within an ulp). Every
// divisor of the passes is such a value (distances to the asymptotes and
// bounds, which the interior point keeps positive, and diagx > 0); an FP64
// division costs about twice as many instructions and the passes carry some
// ten of them per element.
__device__ inline double rcp64(double v) {
  double r = __builtin_amdgcn_rcp(v);
  r = fma(fma(-v, r, 1.0), r, r);
  r = fma(fma(-v, r, 1.0), r, r);
  return r;
}

template <int M>
struct MmaElem {
  double lo, up;      // asymptotes
  double al, be;      // bounds of the subproblem variable
  double p0, q0;      // objective approximation
  double P[M], Q[M];  // constraint approximations
};

// The values a pass reads for one element.
template <int M>
struct MmaRaw {
  float x, g0, g[M];     // design and gradients (FP32 as the caller holds them)
  double xmn, xmx;       // bounds
  double lo, up;         // asymptotes
  double xs, xsi, eta;   // interior point
};

template <int M>
__device__ inline void mma_fetch(const MmaArrays& A, unsigned j, MmaRaw<M>& r) {
  r.x = ld_stream(&A.xvar[j]);
  r.lo = ld_stream(&A.low[j]);
  r.up = ld_stream(&A.upp[j]);
  r.g0 = ld_stream(&A.df0dx[j]);
#pragma unroll
  for (int i = 0; i < M; i++) r.g[i] = ld_stream(&A.dgdx[(size_t)i * (size_t)A.n + j]);
  if (A.xmin) {
    r.xmn = (double)ld_stream(&A.xmin[j]);
    r.xmx = (double)ld_stream(&A.xmax[j]);
  } else {
    r.xmn = A.xmin_u;
    r.xmx = A.xmax_u;
  }
  r.xs = ld_stream(&A.xs[j]);
  r.xsi = ld_stream(&A.xsi[j]);
  r.eta = ld_stream(&A.eta[j]);
}

// alfa, beta, p0, q0, P, Q at the design x, given e.lo/e.up. Every pass
// calls this one function with contraction off, so every pass works with
// the same coefficient bits.
template <int M>
__device__ inline void mma_coefficients(double move, double x, double xmn, double xmx, double g0, const float (&g)[M],
                                        MmaElem<M>& e) {
#pragma clang fp contract(off)
  const double dxm = xmx - xmn;
  e.al = fmax(fmax(e.lo + kAlbefa * (x - e.lo), x - move * dxm), xmn);
  e.be = fmin(fmin(e.up - kAlbefa * (e.up - x), x + move * dxm), xmx);
  const double xmamiinv = rcp64(fmax(dxm, 1e-5));
  const double ux1 = e.up - x, xl1 = x - e.lo, ux2 = ux1 * ux1, xl2 = xl1 * xl1;
  const double reg = kRaa0 * xmamiinv;
  const double dp = fmax(g0, 0.0), dq = fmax(-g0, 0.0);
  const double pq0 = 1e-3 * (dp + dq) + reg;
  e.p0 = (dp + pq0) * ux2;
  e.q0 = (dq + pq0) * xl2;
#pragma unroll
  for (int i = 0; i < M; i++) {
    const double gi = (double)g[i];
    const double pp = fmax(gi, 0.0), qq = fmax(-gi, 0.0);
    const double pq = 1e-3 * (pp + qq) + reg;
    e.P[i] = (pp + pq) * ux2;
    e.Q[i] = (qq + pq) * xl2;
  }
}

template <int M>
__device__ inline void mma_elem(const MmaArrays& A, const MmaRaw<M>& r, MmaElem<M>& e) {
  e.lo = r.lo;
  e.up = r.up;
  mma_coefficients<M>(A.move, (double)r.x, r.xmn, r.xmx, (double)r.g0, r.g, e);
}

// f(j, raw) over this block's elements: the whole grid sweeps the arrays
// front to back (block b takes elements b*256 .. b*256+255 of every
// grid-wide stride), so the HBM sees the pass's 13 streams advancing
// together rather than one stream per block and array. Against a
// contiguous range per block the first trial pass went from 289 to 265 us
// at 2^24 elements, with 1024, 1280 or 2048 blocks alike. (32-bit element
// index: a design has fewer than 2^31 elements.)
template <int M, typename F>
__device__ inline void mma_for_each(const MmaArrays& A, F&& f) {
  const unsigned e = (unsigned)A.n, stride = gridDim.x * blockDim.x;
  for (unsigned j = blockIdx.x * blockDim.x + threadIdx.x; j < e; j += stride) {
    MmaRaw<M> r;
    mma_fetch<M>(A, j, r);
    f(j, r);
  }
}

// Newton-step terms of one element at the point (x, xsi, eta) for the
// multipliers lam (mmaOpt.cu:540-600, :680-700), with the reciprocals of
// the distances it divides by.
template <int M>
struct MmaNewtonElem {
  double rux1, rxl1, rxa, rbx;  // 1/(upp - x), 1/(x - low), 1/(x - alfa), 1/(beta - x)
  double dpsidx, delx, rdiagx, GG[M];
};

// The one term of the Newton system that depends on the barrier epsi.
template <int M>
__device__ inline double mma_delx(const MmaNewtonElem<M>& q, double epsi) {
#pragma clang fp contract(off)
  return q.dpsidx - epsi * q.rxa + epsi * q.rbx;
}

template <int M>
__device__ inline void mma_newton_elem(const MmaElem<M>& e, double x, double xsi, double eta, const double (&lam)[M],
                                       double epsi, MmaNewtonElem<M>& q) {
#pragma clang fp contract(off)
  q.rxa = rcp64(x - e.al);
  q.rbx = rcp64(e.be - x);
  q.rux1 = rcp64(e.up - x);
  q.rxl1 = rcp64(x - e.lo);
  const double rux2 = q.rux1 * q.rux1, rxl2 = q.rxl1 * q.rxl1;
  double plam = e.p0, qlam = e.q0;
#pragma unroll
  for (int i = 0; i < M; i++) {
    plam += lam[i] * e.P[i];
    qlam += lam[i] * e.Q[i];
    q.GG[i] = e.P[i] * rux2 - e.Q[i] * rxl2;
  }
  q.dpsidx = plam * rux2 - qlam * rxl2;
  q.delx = mma_delx<M>(q, epsi);
  const double diagx = 2.0 * (plam * (rux2 * q.rux1) + qlam * (rxl2 * q.rxl1)) + xsi * q.rxa + eta * q.rbx;
  q.rdiagx = rcp64(diagx);
}

// dx = -(delx + GG^T dlam) / diagx (mmaOpt.cu:680-690)
template <int M>
__device__ inline double mma_dx(const MmaNewtonElem<M>& q, const double (&dlam)[M]) {
#pragma clang fp contract(off)
  double gd = 0.0;
#pragma unroll
  for (int i = 0; i < M; i++) gd += q.GG[i] * dlam[i];
  return -(q.delx + gd) * q.rdiagx;
}

// ---- mmasub: asymptotes, right-hand sides, subsolver start ----------------------
template <int M>
__global__ void __launch_bounds__(kBlock) k_mma_prep(MmaArrays A, int itn) {
  double bs[M];
#pragma unroll
  for (int i = 0; i < M; i++) bs[i] = 0.0;
  double dummy[1] = {-1e300};
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < A.n; j += stride) {
    const double x = (double)A.xvar[j];
    const double xmn = A.xmin ? (double)A.xmin[j] : A.xmin_u;
    const double xmx = A.xmax ? (double)A.xmax[j] : A.xmax_u;
    const double dxm = xmx - xmn;
    MmaElem<M> e;
    if (itn <= 2) {
      e.lo = x - kAsyinit * dxm;
      e.up = x + kAsyinit * dxm;
    } else {
      const double x1 = (double)A.xold1[j], x2 = (double)A.xold2[j];
      const double zzz = (x - x1) * (x1 - x2);
      const double f = zzz > 0 ? kAsyincr : (zzz < 0 ? kAsydecr : 1.0);
      double lo = x - f * (x1 - A.low[j]);
      double up = x + f * (A.upp[j] - x1);
      lo = fmin(fmax(lo, x - 10.0 * dxm), x - 0.01 * dxm);
      up = fmax(fmin(up, x + 10.0 * dxm), x + 0.01 * dxm);
      e.lo = lo;
      e.up = up;
    }
    A.low[j] = e.lo;
    A.upp[j] = e.up;
    float g[M];
#pragma unroll
    for (int i = 0; i < M; i++) g[i] = A.dgdx[i * A.n + j];
    mma_coefficients<M>(A.move, x, xmn, xmx, (double)A.df0dx[j], g, e);
    const double ux1 = e.up - x, xl1 = x - e.lo;
#pragma unroll
    for (int i = 0; i < M; i++) bs[i] += e.P[i] / ux1 + e.Q[i] / xl1;
    // subsolv start point (mmaOpt.cu:437-452)
    const double xs = 0.5 * (e.al + e.be);
    A.xs[j] = xs;
    A.xsi[j] = fmax(1.0 / (xs - e.al), 1.0);
    A.eta[j] = fmax(1.0 / (e.be - xs), 1.0);
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
                                      double& mx, double epsi) {
  const int m = S->m;
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

// Sums of one point pass: [|residual|^2, gvec[M], GG.(delx/diagx)[M],
// upper triangle of GG diag(1/diagx) GG^T]; gvec is shared by the residual
// and the Newton system.
// A trial pass also sums, at the next barrier level epsi / 10, the residual
// and GG.(delx/diagx) (NX more sums, one more maximum): when the trial
// ends its level, these with the epsi-free sums are the next level's start
// and its point pass is not needed.
template <int M>
struct MmaPointNV {
  static constexpr int NN = 2 * M + (M * (M + 1)) / 2;  // Newton-system sums
  static constexpr int NS = 1 + NN;
  static constexpr int NX = 1 + M;
};
static_assert(MmaPointNV<kMmaMaxM>::NN <= kMmaMaxNS, "Newton sums exceed their slots");
static_assert(MmaPointNV<kMmaMaxM>::NS + MmaPointNV<kMmaMaxM>::NX + 2 <= 32, "point pass partials exceed their slots");

// TRIAL: the point is the current one + steg * (Newton direction); the
// direction is recomputed from the current point, the point written to the trial
// buffers and the small-vector trial values are formed by the finaliser.
// BOUND (the first trial of a Newton step): the step was taken as
// 1 / stm_small, the bound of the small variables alone, and the pass also
// forms the elements' fraction-to-boundary bound (mmaOpt.cu:700-716) as
// its third maximum; the finaliser redoes the trial if that bound is the
// larger one.
template <int M, bool TRIAL, bool BOUND = false>
__device__ inline void mma_point_body(MmaArrays& A, const MmaSmall* __restrict__ S) {
  constexpr int NS = MmaPointNV<M>::NS, NX = TRIAL ? MmaPointNV<M>::NX : 0, NM = TRIAL ? (BOUND ? 3 : 2) : 1;
  double lam[M], lamc[M], dlam[M];
  const double epsi = S->epsi, steg = S->steg;
  const double epsi2 = epsi * 0.1;  // the next level's, as k_mma_epsi forms it
#pragma unroll
  for (int i = 0; i < M; i++) {
    lamc[i] = S->lam[i];
    dlam[i] = S->dlam[i];
    lam[i] = TRIAL ? S->lamtr[i] : lamc[i];
  }
  double sums[NS + NX];
#pragma unroll
  for (int i = 0; i < NS + NX; i++) sums[i] = 0.0;
  double mx[NM];
#pragma unroll
  for (int i = 0; i < NM; i++) mx[i] = 0.0;
  if (BOUND) mx[NM - 1] = -1e300;
  mma_for_each<M>(A, [&](unsigned j, const MmaRaw<M>& r) {
    MmaElem<M> e;
    mma_elem<M>(A, r, e);
    double x = r.xs, xsi = r.xsi, eta = r.eta;
    if (TRIAL) {
      MmaNewtonElem<M> qc;
      mma_newton_elem<M>(e, x, xsi, eta, lamc, epsi, qc);
      const double dx = mma_dx<M>(qc, dlam);
      const double dxsi = -xsi + epsi * qc.rxa - xsi * dx * qc.rxa;
      const double deta = -eta + epsi * qc.rbx + eta * dx * qc.rbx;
      if (BOUND)
        mx[NM - 1] = fmax(mx[NM - 1], fmax(fmax(-1.01 * dxsi * rcp64(xsi), -1.01 * deta * rcp64(eta)),
                                           fmax(-1.01 * dx * qc.rxa, 1.01 * dx * qc.rbx)));
      x += steg * dx;
      xsi += steg * dxsi;
      eta += steg * deta;
      // non-temporal, like the loads: a pass streams more than the caches
      // hold, and with both marked the pass ran 2-3% faster (loads alone
      // 1%, stores alone not at all)
      __builtin_nontemporal_store(x, &A.xs_t[j]);
      __builtin_nontemporal_store(xsi, &A.xsi_t[j]);
      __builtin_nontemporal_store(eta, &A.eta_t[j]);
    }
    MmaNewtonElem<M> q;
    mma_newton_elem<M>(e, x, xsi, eta, lam, epsi, q);
#pragma unroll
    for (int i = 0; i < M; i++) sums[1 + i] += e.P[i] * q.rux1 + e.Q[i] * q.rxl1;  // gvec
    const double rex = q.dpsidx - xsi + eta;
    const double rexsi = xsi * (x - e.al) - epsi;
    const double reeta = eta * (e.be - x) - epsi;
    sums[0] += rex * rex + rexsi * rexsi + reeta * reeta;
    mx[0] = fmax(mx[0], fmax(fabs(rex), fmax(fabs(rexsi), fabs(reeta))));
    // the Newton system at this point (mmaOpt.cu:560-600)
    const double inv = q.rdiagx;
    int t = 1 + 2 * M;
#pragma unroll
    for (int i = 0; i < M; i++) {
      sums[1 + M + i] += q.GG[i] * q.delx * inv;
#pragma unroll
      for (int k = i; k < M; k++) sums[t++] += q.GG[i] * q.GG[k] * inv;
    }
    if (TRIAL) {
      const double rexsi2 = xsi * (x - e.al) - epsi2;
      const double reeta2 = eta * (e.be - x) - epsi2;
      sums[NS] += rex * rex + rexsi2 * rexsi2 + reeta2 * reeta2;
      mx[1] = fmax(mx[1], fmax(fabs(rex), fmax(fabs(rexsi2), fabs(reeta2))));
      const double delx2 = mma_delx<M>(q, epsi2);
#pragma unroll
      for (int i = 0; i < M; i++) sums[NS + 1 + i] += q.GG[i] * delx2 * inv;
    }
  });
  block_sum_max<NS + NX, NM>(sums, mx, A.partials);
}

// The start-of-level point pass (up to m = 3) asks the compiler for five
// waves per SIMD. The trial passes do not: in grid-stride order a trial
// pass with that request (96 VGPRs, 1280 workgroups) took 282 us at 2^24
// elements, without it (four waves, 1024 workgroups) 262 us; at m = 4 the
// request would spill.
template <int M, bool TRIAL, bool BOUND>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_mma_point(MmaArrays A, const MmaSmall* __restrict__ S) {
  mma_point_body<M, TRIAL, BOUND>(A, S);
}
template <int M, bool TRIAL, bool BOUND>
__global__ void __launch_bounds__(kBlock) k_mma_point_wide(MmaArrays A, const MmaSmall* __restrict__ S) {
  mma_point_body<M, TRIAL, BOUND>(A, S);
}
template <int M, bool TRIAL, bool BOUND = false>
static constexpr auto mma_point_kernel() {
  if constexpr (M <= 3 && !BOUND && !TRIAL) return k_mma_point<M, TRIAL, BOUND>;
  else return k_mma_point_wide<M, TRIAL, BOUND>;
}

// The host loop's inputs to the host block, then the tag (one thread).
__device__ inline void mma_publish(const MmaSmall* S, MmaHost* mh, unsigned tag) {
  mh->residumax = S->residumax;
  mh->rmaxnew = S->rmaxnew;
  mh->rmaxnext = S->rmaxn;
  mh->accept = S->accept;
  mh->respec = S->respec;
  mh->fault = S->fault;
  __hip_atomic_store(&mh->seq, tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <int M, bool TRIAL, bool BOUND = false>
__device__ inline void mma_point_final_body(MmaArrays& A, MmaSmall* S, int nblocks) {
  constexpr int NS = MmaPointNV<M>::NS, NN = MmaPointNV<M>::NN;
  constexpr int NX = TRIAL ? MmaPointNV<M>::NX : 0, NM = TRIAL ? (BOUND ? 3 : 2) : 1;
  __shared__ double out[NS + NX + NM];
  final_sum_max<NS + NX, NM>(A.partials, nblocks, out);
  if (threadIdx.x != 0) return;
  S->respec = 0;
  if (BOUND) {
    // step = 1 / max(stm_small, element bound) (mmaOpt.cu:700-716): when the
    // elements bound it, the trial taken at 1 / stm_small is discarded and
    // the line search starts at the right step
    const double stm_elem = out[NS + NX + NM - 1];
    if (stm_elem > S->stm_small) {
      S->steg = 1.0 / fmax(S->stm_small, stm_elem);
      for (int i = 0; i < M; i++) S->lamtr[i] = S->lam[i] + S->steg * S->dlam[i];
      S->accept = 0;
      S->respec = 1;
      return;
    }
  }
  double ss = out[0], mx = out[NS + NX];
  double* ns = TRIAL ? S->ns_t : S->ns;
  for (int i = 0; i < NN; i++) ns[i] = out[1 + i];
  if (!TRIAL) {
    small_residual(S, out + 1, S->y, S->z, S->lam, S->mu, S->zet, S->s, ss, mx, S->epsi);
    S->residunorm = sqrt(ss);
    S->residumax = mx;
    return;
  }
  const double steg = S->steg;
  for (int i = 0; i < M; i++) {
    S->yt[i] = S->y[i] + steg * S->dy[i];
    S->lamt[i] = S->lamtr[i];  // the multipliers the pass used
    S->mut[i] = S->mu[i] + steg * S->dmu[i];
    S->st[i] = S->s[i] + steg * S->ds[i];
  }
  S->zt = S->z + steg * S->dz;
  S->zett = S->zet + steg * S->dzet;
  small_residual(S, out + 1, S->yt, S->zt, S->lamt, S->mut, S->zett, S->st, ss, mx, S->epsi);
  S->resinew = sqrt(ss);
  S->rmaxnew = mx;
  // the same point at the next barrier level (used if this trial ends the level)
  const double epsi2 = S->epsi * 0.1;
  double ss2 = out[NS], mx2 = out[NS + NX + 1];
  for (int i = 0; i < NN; i++) S->ns_n[i] = out[1 + i];
  for (int i = 0; i < M; i++) S->ns_n[M + i] = out[NS + 1 + i];
  small_residual(S, out + 1, S->yt, S->zt, S->lamt, S->mut, S->zett, S->st, ss2, mx2, epsi2);
  S->resn = sqrt(ss2);
  S->rmaxn = mx2;
  S->itto += 1;
  // line search rule (mmaOpt.cu:718-805): accept when the norm does not
  // grow, or after 50 halvings
  S->accept = (!(S->resinew > S->residunorm) || S->itto >= 50) ? 1 : 0;
  if (!(S->resinew - S->resinew == 0.0)) S->fault = 1;
  S->steg = steg / 2;
  for (int i = 0; i < M; i++) S->lamtr[i] = S->lam[i] + S->steg * S->dlam[i];
}

template <int M, bool TRIAL, bool BOUND>
__global__ void __launch_bounds__(kBlock) k_mma_point_final(MmaArrays A, MmaSmall* S, int nblocks, MmaHost* mh,
                                                            unsigned tag) {
  mma_point_final_body<M, TRIAL, BOUND>(A, S, nblocks);
  if (threadIdx.x == 0) mma_publish(S, mh, tag);
}

// After acceptance the trial becomes the current point, with its Newton
// sums; done by the kernel that follows the accepted trial (the next Newton
// system or the next barrier level).
__device__ inline void mma_commit(MmaSmall* S) {
  for (int i = 0; i < S->m; i++) {
    S->y[i] = S->yt[i];
    S->lam[i] = S->lamt[i];
    S->mu[i] = S->mut[i];
    S->s[i] = S->st[i];
  }
  for (int i = 0; i < kMmaMaxNS; i++) S->ns[i] = S->ns_t[i];
  S->z = S->zt;
  S->zet = S->zett;
  S->residunorm = S->resinew;
  S->residumax = S->rmaxnew;
}

// ---- Newton system ---------------------------------------------------------------
// Solves [Alam a; a^T -zet/z] [dlam; dz] = [blam; delz] from the point's
// sums and the small directions (mmaOpt.cu:595-678). One thread.
template <int M>
__global__ void k_mma_newton_final(MmaSmall* S, int commit) {
  if (commit) mma_commit(S);
  const double epsi = S->epsi;
  const double* gvec = S->ns;
  const double* GGdd = S->ns + M;
  double Alam[M][M];
  int t = 2 * M;
  for (int i = 0; i < M; i++)
    for (int k = i; k < M; k++) {
      Alam[i][k] = S->ns[t];
      Alam[k][i] = S->ns[t];
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
  // the first trial takes the small variables' step; its pass forms the
  // elements' bound and the finaliser redoes it if that one is larger
  S->steg = 1.0 / stm;
  for (int i = 0; i < M; i++) S->lamtr[i] = S->lam[i] + S->steg * S->dlam[i];
  S->itto = 0;
  S->accept = 0;
}

__global__ void k_mma_epsi(MmaSmall* S, int commit, int spec) {
  if (commit) mma_commit(S);
  S->epsi *= 0.1;
  if (spec) {
    for (int i = 0; i < kMmaMaxNS; i++) S->ns[i] = S->ns_n[i];
    S->residunorm = S->resn;
    S->residumax = S->rmaxn;
  }
}

// xold2 = xold1, xold1 = x (mmaOpt.cu:1055-1057); the subsolver solution
// goes to the caller's FP32 design
__global__ void __launch_bounds__(kBlock) k_mma_finish(MmaArrays A, float* xout) {
  const long long j = stid();
  if (j >= A.n) return;
  A.xold2[j] = A.xold1[j];
  A.xold1[j] = A.xvar[j];
  xout[j] = (float)A.xs[j];
}

// ---- host driver ---------------------------------------------------------------
static hipEvent_t prof_event(MmaProf& P) {
  if (!P.pool.empty()) {
    hipEvent_t e = P.pool.back();
    P.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
// events of finished launches -> per-class time (after a stream sync)
static void prof_collect(MmaProf& P) {
  for (const auto& q : P.pending) {
    float ms = 0.f;
    if (hipEventSynchronize(q.b) == hipSuccess && hipEventElapsedTime(&ms, q.a, q.b) == hipSuccess) P.ms[q.cls] += ms;
    P.pool.push_back(q.a);
    P.pool.push_back(q.b);
  }
  P.pending.clear();
}
void mma_prof_release(MmaProf& P) {
  for (const auto& q : P.pending) {
    P.pool.push_back(q.a);
    P.pool.push_back(q.b);
  }
  P.pending.clear();
  for (hipEvent_t e : P.pool) (void)hipEventDestroy(e);
  P.pool.clear();
}

// Workgroups of a streaming pass: as many as the device holds at once at
// the kernel's occupancy (its register count decides it), so that every CU
// keeps that many waves' loads in flight through the whole pass; at most
// one per kBlock elements and kMmaMaxBlocks partial-sum slots.
template <typename K>
static unsigned mma_grid(K kernel, long long n) {
  int dev = 0, cus = 0, per_cu = 0;
  long long b = 1024;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) == hipSuccess && per_cu > 0 &&
      cus > 0)
    b = (long long)per_cu * cus;
  b = std::min<long long>(b, kMmaMaxBlocks);
  b = std::min<long long>(b, (n + kBlock - 1) / kBlock);
  return (unsigned)std::max<long long>(b, 1);
}

// Polls the host block for a tag; -1 when the stream finished without it
// or failed.
static int mma_wait(hipStream_t s, const MmaHost* mh, unsigned tag) {
  unsigned spins = 0;
  while (__atomic_load_n(&mh->seq, __ATOMIC_ACQUIRE) != tag) {
    if ((++spins & 255u) == 0) {
      const hipError_t e = hipStreamQuery(s);
      if (e == hipSuccess) return __atomic_load_n(&mh->seq, __ATOMIC_ACQUIRE) == tag ? 0 : -1;
      if (e != hipErrorNotReady) return -1;
    }
    __builtin_ia32_pause();
  }
  return 0;
}

template <int M>
static int mma_run(hipStream_t s, MmaArrays A, MmaSmall* S, MmaHost* mh, MmaHost* mh_dev, const double* gval_dev,
                   int itn, double a0, double av, double cv, double dv, float* xout, double epsimin, MmaProf* prof) {
  const unsigned nb_prep = mma_grid(k_mma_prep<M>, A.n);
  const unsigned nb_point = mma_grid(mma_point_kernel<M, false>(), A.n);
  const unsigned nb_trial = mma_grid(mma_point_kernel<M, true>(), A.n);
  const unsigned nb_bound = mma_grid(mma_point_kernel<M, true, true>(), A.n);
  // no kernel of this optimiser is in flight: the block holds the last tag
  unsigned tag = __atomic_load_n(&mh->seq, __ATOMIC_ACQUIRE);
  auto read = [&]() -> int {
    if (mma_wait(s, mh, tag) != 0) return -1;
    if (prof && prof->on) prof_collect(*prof);
    return mh->fault ? -2 : 0;
  };
  // algorithmic bytes per element: x, low/upp, gradients, interior point
  // (+ per-element bounds), and the trial point written
  const double rd = 4.0 + 16.0 + 4.0 * (1 + M) + 24.0 + (A.xmin ? 8.0 : 0.0);
  const double bytes_of[3] = {(rd + 24.0) * (double)A.n, (rd + 24.0) * (double)A.n, rd * (double)A.n};
  struct Timed {
    MmaProf* P;
    int cls;
    hipStream_t s;
    hipEvent_t a = nullptr;
    Timed(MmaProf* P_, int cls_, hipStream_t s_, const double* by) : P(P_), cls(cls_), s(s_) {
      if (!P || !P->on) return;
      a = prof_event(*P);
      if (a && hipEventRecord(a, s) != hipSuccess) a = nullptr;
      if (a) {
        P->launches[cls]++;
        P->bytes[cls] += by[cls];
      }
    }
    ~Timed() {
      if (!a) return;
      hipEvent_t b = prof_event(*P);
      if (b && hipEventRecord(b, s) == hipSuccess)
        P->pending.push_back({cls, a, b});
      else
        P->pool.push_back(a);
    }
  };
  hipLaunchKernelGGL(k_mma_prep<M>, dim3(nb_prep), dim3(kBlock), 0, s, A, itn);
  hipLaunchKernelGGL(k_mma_prep_final<M>, dim3(1), dim3(kBlock), 0, s, A, S, gval_dev, (int)nb_prep, a0, av, cv,
                     dv);
  int newton = 0;
  double epsi = 1.0;
  int commit = 0;   // an accepted trial waits to become the current point
  bool spec = false;  // that trial's sums at this level are known (its pass formed them)
  double residumax_next = 0.0;
  while (epsi > epsimin) {
    double residumax;
    if (spec) {
      residumax = residumax_next;
    } else {
      // residual and Newton system at the start point of this epsi level
      {
        Timed t(prof, 2, s, bytes_of);
        hipLaunchKernelGGL((mma_point_kernel<M, false>()), dim3(nb_point), dim3(kBlock), 0, s, A, S);
      }
      hipLaunchKernelGGL((k_mma_point_final<M, false, false>), dim3(1), dim3(kBlock), 0, s, A, S, (int)nb_point,
                         mh_dev, ++tag);
      const int rc = read();
      if (rc) return rc;
      residumax = mh->residumax;
    }
    spec = false;
    int rc = 0;
    int ittt = 0;
    while (residumax > 0.9 * epsi && ittt < 200) {
      ittt++;
      newton++;
      hipLaunchKernelGGL(k_mma_newton_final<M>, dim3(1), dim3(1), 0, s, S, commit);
      commit = 0;
      // line search; the first trial also checks the step it was given
      bool first = true;
      do {
        {
          Timed t(prof, first ? 1 : 0, s, bytes_of);
          if (first)
            hipLaunchKernelGGL((mma_point_kernel<M, true, true>()), dim3(nb_bound), dim3(kBlock), 0, s, A, S);
          else
            hipLaunchKernelGGL((mma_point_kernel<M, true>()), dim3(nb_trial), dim3(kBlock), 0, s, A, S);
        }
        if (first)
          hipLaunchKernelGGL((k_mma_point_final<M, true, true>), dim3(1), dim3(kBlock), 0, s, A, S, (int)nb_bound,
                             mh_dev, ++tag);
        else
          hipLaunchKernelGGL((k_mma_point_final<M, true, false>), dim3(1), dim3(kBlock), 0, s, A, S, (int)nb_trial,
                             mh_dev, ++tag);
        first = false;
        rc = read();
        if (rc) return rc;
      } while (!mh->accept);
      commit = 1;
      std::swap(A.xs, A.xs_t);
      std::swap(A.xsi, A.xsi_t);
      std::swap(A.eta, A.eta_t);
      residumax = mh->rmaxnew;
      residumax_next = mh->rmaxnext;
      spec = true;
    }
    hipLaunchKernelGGL(k_mma_epsi, dim3(1), dim3(1), 0, s, S, commit, spec ? 1 : 0);
    commit = 0;
    epsi *= 0.1;
  }
  hipLaunchKernelGGL(k_mma_finish, dim3(nblk(A.n)), dim3(kBlock), 0, s, A, xout);
  if (hipGetLastError() != hipSuccess) return -1;
  return newton;
}

int mma_update(hipStream_t s, int m, MmaArrays& A, void* small_dev, MmaHost* mh, MmaHost* mh_dev,
               const double* gval_dev, int itn, double a0, double av, double cv, double dv, float* xout,
               MmaProf* prof) {
  MmaSmall* S = (MmaSmall*)small_dev;
  MmaArrays B = A;  // the driver swaps the subsolver buffers locally
  int rc;
  switch (m) {
    case 1: rc = mma_run<1>(s, B, S, mh, mh_dev, gval_dev, itn, a0, av, cv, dv, xout, 1e-7, prof); break;
    case 2: rc = mma_run<2>(s, B, S, mh, mh_dev, gval_dev, itn, a0, av, cv, dv, xout, 1e-7, prof); break;
    case 3: rc = mma_run<3>(s, B, S, mh, mh_dev, gval_dev, itn, a0, av, cv, dv, xout, 1e-7, prof); break;
    case 4: rc = mma_run<4>(s, B, S, mh, mh_dev, gval_dev, itn, a0, av, cv, dv, xout, 1e-7, prof); break;
    default: rc = -3;
  }
  return rc;
}

size_t mma_small_bytes() { return sizeof(MmaSmall); }

}  // namespace otm
