// otm_opt.hip -- density filter, SIMP map and OC update kernels.
//
// Filter: renormalised radial convolution (TensorExpression.cuh:303-384):
// out_i = sum_j w_ij in_j / W_i over taps inside the cell (all taps when
// periodic), adjoint out_j = sum_i w_ij in_i / W_i. OC: density-weighted
// periodic sensitivity filter (OCOptimizer.cu:76-104) and the bisection
// update rule (OCOptimizer.cu:12-34).
#include <algorithm>
#include <cstdlib>
#include <stdexcept>

#include "otm_kernels.hpp"

namespace otm {

static inline unsigned nblk(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ---- filters, LDS-tiled ----
// A block computes a 32 x 8 x 4 box of outputs (thread (x, y), four z) from
// an LDS copy of the box grown by the taps' reach h (at most 3) on every
// side, so each input comes from HBM about twice (halo) instead of once per
// tap. Positions outside a non-periodic cell hold 0, which leaves the sums
// as if the tap were skipped. Taps become offsets into that copy.
constexpr int kFtX = 32, kFtY = 8, kFtZ = 4, kFtH = 3;
constexpr int kFtLds = (kFtX + 2 * kFtH) * (kFtY + 2 * kFtH) * (kFtZ + 2 * kFtH);
static_assert(kFtX * kFtY == kBlock, "one thread per (x, y) column of a filter tile");
enum { FILT_FWD = 0, FILT_ADJ = 1, FILT_SENS = 2 };

// FILT_FWD: out = sum w in / W and W = sum of the weights of the taps
// inside the cell (all taps when periodic; W may be null).
// FILT_ADJ: out_j = sum_i w in_i / W_i over the reversed taps (b = W).
// FILT_SENS: out = sum w (sens rho) / (W rho), periodic (a = sens, b = rho).
// wall is the sum of all tap weights in tap order (the W of an output whose
// taps are all inside).
// w_known (FILT_FWD): W already holds the weight sums (a previous call with
// the same taps wrote them); nodes near a non-periodic cell's faces read
// theirs instead of summing the taps again.
template <int MODE>
__global__ void __launch_bounds__(kBlock) k_filter_tile(Geom g, Taps T, int h, float wall, const float* __restrict__ a,
                                                        const float* __restrict__ b, float* __restrict__ out,
                                                        float* __restrict__ W, int w_known) {
  __shared__ float tile[kFtLds];
  __shared__ int toff[kMaxTaps];
  __shared__ float tw[kMaxTaps];
  const int SX = kFtX + 2 * h, SY = kFtY + 2 * h, SZ = kFtZ + 2 * h;
  const int ntx = (g.nx + kFtX - 1) / kFtX, nty = (g.ny + kFtY - 1) / kFtY;
  const int bx = (int)blockIdx.x % ntx, byz = (int)blockIdx.x / ntx;
  const int x0 = bx * kFtX, y0 = (byz % nty) * kFtY, z0 = (byz / nty) * kFtZ;
  const bool periodic = MODE == FILT_SENS || T.periodic;
  for (int i = threadIdx.x; i < T.n; i += kBlock) {
    const int o = T.dx[i] + SX * (T.dy[i] + SY * T.dz[i]);
    toff[i] = MODE == FILT_ADJ ? -o : o;
    tw[i] = T.w[i];
  }
  for (int i = threadIdx.x; i < SX * SY * SZ; i += kBlock) {
    const int lx = i % SX, rest = i / SX;
    int X = x0 - h + lx, Y = y0 - h + rest % SY, Z = z0 - h + rest / SY;
    float v = 0.f;
    // positions past the cell by more than h feed no output of this tile
    if (X < g.nx + h && Y < g.ny + h && Z < g.nz + h) {
      const bool inside = X >= 0 && X < g.nx && Y >= 0 && Y < g.ny && Z >= 0 && Z < g.nz;
      if (inside || periodic) {
        X = wrapm(X, g.nx);
        Y = wrapm(Y, g.ny);
        Z = wrapm(Z, g.nz);
        const long long j = (long long)X + (long long)g.nx * ((long long)Y + (long long)g.ny * Z);
        v = MODE == FILT_FWD ? a[j] : (MODE == FILT_ADJ ? a[j] / b[j] : a[j] * b[j]);
      }
    }
    tile[i] = v;
  }
  __syncthreads();
  const int tx = threadIdx.x % kFtX, ty = threadIdx.x / kFtX;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= g.nx || y >= g.ny) return;
  const int plane = SX * SY;
  const int c0 = (tx + h) + SX * ((ty + h) + SY * h);
  float acc[kFtZ] = {0.f, 0.f, 0.f, 0.f};
  for (int i = 0; i < T.n; i++) {
    const int o = c0 + toff[i];
    const float w = tw[i];
#pragma unroll
    for (int k = 0; k < kFtZ; k++) acc[k] = fmaf(w, tile[o + k * plane], acc[k]);
  }
#pragma unroll
  for (int k = 0; k < kFtZ; k++) {
    const int z = z0 + k;
    if (z >= g.nz) break;
    const long long t = (long long)x + (long long)g.nx * ((long long)y + (long long)g.ny * z);
    if (MODE == FILT_ADJ) {
      out[t] = acc[k];
    } else if (MODE == FILT_SENS) {
      out[t] = acc[k] / (wall * b[t]);
    } else {
      float ws = wall;
      const bool interior = x >= h && x + h < g.nx && y >= h && y + h < g.ny && z >= h && z + h < g.nz;
      if (!periodic && !interior) {
        if (w_known) {
          ws = W[t];
        } else {
          ws = 0.f;
          for (int i = 0; i < T.n; i++) {
            const int X = x + T.dx[i], Y = y + T.dy[i], Z = z + T.dz[i];
            if (X >= 0 && X < g.nx && Y >= 0 && Y < g.ny && Z >= 0 && Z < g.nz) ws += tw[i];
          }
        }
      }
      out[t] = acc[k] / ws;
      if (W && !w_known) W[t] = ws;
    }
  }
}

// kappa = k1 + (k0 - k1) rho^3 (+floor), dkappa = 3 (k0 - k1) rho^2
// (Framework.cu:163 with penalty 3).
__global__ void __launch_bounds__(kBlock) k_simp_fwd(long long n, const float* __restrict__ rho, float k0, float k1,
                                                     float floor_, float* __restrict__ kappa,
                                                     float* __restrict__ dkappa) {
  const long long t = stid();
  if (t >= n) return;
  const float p = rho[t];
  const float d = k0 - k1;
  kappa[t] = k1 + d * p * p * p + floor_;
  if (dkappa) dkappa[t] = 3.f * d * p * p;
}

__global__ void __launch_bounds__(kBlock) k_mul(long long n, const float* __restrict__ a, const float* __restrict__ b,
                                                float* __restrict__ out) {
  const long long t = stid();
  if (t >= n) return;
  out[t] = a[t] * b[t];
}

__global__ void __launch_bounds__(kBlock) k_maxabs(long long n, const float* __restrict__ a, unsigned int* out_bits) {
  float m = 0.f;
  for (long long i = stream_range(n).b + threadIdx.x, i_end = stream_range(n).e; i < i_end; i += blockDim.x)
    m = fmaxf(m, fabsf(a[i]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_down(m, o, 64));
  __shared__ float wm[kWaves];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kWaves; w++) m = fmaxf(m, wm[w]);
    atomicMax(out_bits, __float_as_uint(m));
  }
}


// Starts the bisection from [0, max|sens|] and clears the maximum for the
// next call (the word is zeroed once when the optimiser is created).
__global__ void k_oc_init(unsigned int* __restrict__ bits, OcState* st) {
  const unsigned int b = *bits;
  *bits = 0u;
  st->lo = 0.f;
  st->hi = fabsf(__uint_as_float(b));
  st->g = 0.f;
  st->done = 0;
  st->steps = 0;
  st->vol = 0.0;
}

// Speculative bisection: one pass over the design evaluates the midpoints of
// the next kOcTree bisection levels at once (heap order: node i has children
// 2i+1, "volume too small, hi = g", and 2i+2, "volume too large, lo = g"),
// and a one-block kernel then walks the tree with the bracket rule of
// OCOptimizer.cu:50-61. The midpoints are the FP32 midpoints the sequential
// bisection would take, and each volume is the same FP64 reduction, so the
// walk takes the sequential bisection's path kOcTree levels per pass.
//
// Update rule of OCOptimizer.cu:12-34 per element: B = -s/g (0.01 when
// negative), rho' = clamp(B^damp * rho, rho -+ step, [min_rho, 1]). B^damp is
// evaluated as 2^(damp (log2(-s) - log2 g)) so the logarithm of the
// sensitivity is taken once per element for all midpoints.
// Midpoint of heap node i: the path from the root is the binary expansion
// of i + 1 after its leading one (0: left child, hi = g; 1: right, lo = g),
// with the bracket halved in FP32 exactly as the sequential bisection does.
__device__ inline float oc_tree_mid(const OcState* st, int i) {
  float lo = st->lo, hi = st->hi;
  const unsigned p = (unsigned)i + 1u;
  const int depth = 31 - __builtin_clz(p);
  for (int j = depth - 1; j >= 0; j--) {
    const float g = 0.5f * (hi + lo);
    if ((p >> j) & 1u)
      lo = g;
    else
      hi = g;
  }
  return 0.5f * (hi + lo);
}

__device__ inline float oc_rule_log(float s, float ls, float rho, float lg, float c01, float min_rho, float step,
                                    float damp) {
  float v = (s > 0.f ? c01 : __builtin_amdgcn_exp2f(damp * (ls - lg))) * rho;
  if (v - rho < -step) v = rho - step;
  if (v - rho > step) v = rho + step;
  if (v < min_rho) v = min_rho;
  if (v > 1.f) v = 1.f;
  return v;
}

// The bisection's volume sums factor the rule's power: (-s/g)^damp =
// 2^(damp log2(-s)) * 2^(-damp log2 g), one exp2 per element and one per
// candidate instead of one per (element, candidate). The final update
// (k_oc_apply) evaluates the rule unfactored with the chosen multiplier.
__device__ inline float oc_rule_fac(float s, float e, float kc, float rho, float c01, float min_rho, float step) {
  float v = (s > 0.f ? c01 : e * kc) * rho;
  if (v - rho < -step) v = rho - step;
  if (v - rho > step) v = rho + step;
  if (v < min_rho) v = min_rho;
  if (v > 1.f) v = 1.f;
  return v;
}

// Adds one element's updated density at every candidate multiplier. With
// the element's factors er = 2^(damp log2(-s)) rho and its clamp interval
// [max(rho - step, min_rho), min(rho + step, 1)], each candidate costs a
// multiply, a median-of-three and an add; an element with s > 0 takes the
// multiplier-free value c01 rho (clamped) at every candidate (interval of
// one point).
__device__ inline void oc_tree_add(float sv, float r, float damp, float c01, float min_rho, float step,
                                   const float (&kc)[kOcCand], float (&accf)[kOcCand]) {
  float er = 0.f, lo, hi;
  if (sv > 0.f) {
    lo = hi = oc_rule_fac(sv, 0.f, 0.f, r, c01, min_rho, step);
  } else {
    er = __builtin_amdgcn_exp2f(damp * __builtin_amdgcn_logf(-sv)) * r;
    lo = fmaxf(r - step, min_rho);
    hi = fminf(r + step, 1.f);
  }
#pragma unroll
  for (int c = 0; c < kOcCand; c++) accf[c] += __builtin_amdgcn_fmed3f(er * kc[c], lo, hi);
}

// Block stage of the candidate sums. The threads' FP32 sums go through LDS
// (rows padded to 33 words: conflict-free writes and reads), thread t adds
// candidate t % 32 over the 32 threads of group t / 32 in FP64, and the
// eight group sums are added in group order. Wave 0 stores the block's
// partials at the block's range index, so the grid stage adds them in item
// order whatever the launch shape. (A shuffle tree per candidate costs six
// cross-lane rounds for each of the 31 sums.)
__device__ inline void oc_block_partials(const float (&accf)[kOcCand], const Reduce& R) {
  constexpr int kRow = 33, kGroups = kBlock / 32;
  __shared__ float rows[kBlock * kRow];
  __shared__ double part[kGroups][32];
#pragma unroll
  for (int c = 0; c < kOcCand; c++) rows[threadIdx.x * kRow + c] = accf[c];
  __syncthreads();
  const int c = threadIdx.x & 31, j = threadIdx.x >> 5;
  double a = 0.0;
  if (c < kOcCand) {
#pragma unroll 8
    for (int k = 0; k < 32; k++) a += (double)rows[(j * 32 + k) * kRow + c];
  }
  part[j][c] = a;
  __syncthreads();
  if (threadIdx.x < kOcCand) {
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < kGroups; q++) sum += part[q][threadIdx.x];
    R.partials[(size_t)blockIdx.x * kOcCand + threadIdx.x] = sum;
  }
}

// Per-thread sums stay in FP32 over the thread's few elements (each term is
// at most 1) and go to FP64 for the block and grid stages. Four elements
// per thread are loaded together so that their latencies overlap.
__device__ inline void oc_tree_sum_body(long long n, const float* __restrict__ sens, const float* __restrict__ rho,
                                        const OcState* __restrict__ st, float min_rho, float step, float damp,
                                        float c01, const Reduce& R) {
  __shared__ float sk[kOcCand];
  if (threadIdx.x < kOcCand)
    sk[threadIdx.x] = __builtin_amdgcn_exp2f(-damp * __builtin_amdgcn_logf(oc_tree_mid(st, threadIdx.x)));
  __syncthreads();
  float kc[kOcCand], accf[kOcCand];
#pragma unroll
  for (int c = 0; c < kOcCand; c++) {
    kc[c] = sk[c];
    accf[c] = 0.f;
  }
  const BlockRange rg = stream_range(n);
  long long i = rg.b + threadIdx.x;
  for (; i + 3 * kBlock < rg.e; i += 4 * kBlock) {
    float sv[4], r[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      sv[k] = sens[i + k * kBlock];
      r[k] = rho[i + k * kBlock];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) oc_tree_add(sv[k], r[k], damp, c01, min_rho, step, kc, accf);
  }
  for (; i < rg.e; i += kBlock) oc_tree_add(sens[i], rho[i], damp, c01, min_rho, step, kc, accf);
  oc_block_partials(accf, R);
}

__global__ void __launch_bounds__(kBlock) k_oc_tree_sum(long long n, const float* __restrict__ sens,
                                                        const float* __restrict__ rho,
                                                        const OcState* __restrict__ st, float min_rho, float step,
                                                        float damp, float c01, int max_steps, Reduce R) {
  if (st->done || st->steps >= max_steps) return;
  oc_tree_sum_body(n, sens, rho, st, min_rho, step, damp, c01, R);
}

// Sums the partials of the kOcCand midpoints and walks the bisection tree
// with them (thread 0): the state's bracket, multiplier and step count
// advance by up to kOcTree levels. Thread t sums candidate t % 32 over the
// blocks t / 32, t / 32 + 8, ... (coalesced reads, eight in flight), and the
// eight groups are added in group order: a fixed order for every launch.
__device__ inline void oc_tree_final_body(OcState* st, const Reduce& R, int nblocks, double inv_n, double volratio,
                                          double tol, int max_steps) {
  constexpr int kGroups = kBlock / 32;
  __shared__ double part[kGroups][32];
  __shared__ double vol_sum[kOcCand];
  __shared__ float g[kOcCand];
  const int c = threadIdx.x & 31, j = threadIdx.x >> 5;
  if (threadIdx.x < kOcCand) g[threadIdx.x] = oc_tree_mid(st, threadIdx.x);
  double a = 0.0;
  if (c < kOcCand) {
    const double* src = R.partials + c;
    int b = j;
    for (; b + 7 * kGroups < nblocks; b += 8 * kGroups) {
      double v[8];
#pragma unroll
      for (int k = 0; k < 8; k++) v[k] = src[(size_t)(b + k * kGroups) * kOcCand];
#pragma unroll
      for (int k = 0; k < 8; k++) a += v[k];
    }
    for (; b < nblocks; b += kGroups) a += src[(size_t)b * kOcCand];
  }
  part[j][c] = a;
  __syncthreads();
  if (threadIdx.x < kOcCand) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kGroups; q++) s += part[q][threadIdx.x];
    vol_sum[threadIdx.x] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  float lo = st->lo, hi = st->hi;
  int node = 0, steps = st->steps;
  for (int lev = 0; lev < kOcTree && steps < max_steps; lev++) {
    const float gm = g[node];
    const double vol = vol_sum[node] * inv_n;
    st->g = gm;
    st->vol = vol;
    steps++;
    if (vol < volratio - tol) {
      hi = gm;
      node = 2 * node + 1;
    } else if (vol > volratio + tol) {
      lo = gm;
      node = 2 * node + 2;
    } else {
      st->done = 1;
      break;
    }
  }
  st->lo = lo;
  st->hi = hi;
  st->steps = steps;
}

__global__ void __launch_bounds__(kBlock) k_oc_tree_final(OcState* st, Reduce R, int nblocks, double inv_n,
                                                          double volratio, double tol, int max_steps) {
  if (st->done || st->steps >= max_steps) return;
  oc_tree_final_body(st, R, nblocks, inv_n, volratio, tol, max_steps);
}

// rho <- update rule at the final multiplier (in place).
__global__ void __launch_bounds__(kBlock) k_oc_apply(long long n, const float* __restrict__ sens, float* rho,
                                                     const OcState* __restrict__ st, float min_rho, float step,
                                                     float damp, float c01, int max_steps, int force, OcHost* hm,
                                                     unsigned tag) {
  const bool apply = st->steps > 0 && (st->done || st->steps >= max_steps || force);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    hm->st = *st;
    hm->applied = apply ? 1 : 0;
    __hip_atomic_store(&hm->seq, tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  if (!apply) return;
  const float lg = __builtin_amdgcn_logf(st->g);
  for (long long i = stid(); i < n; i += (long long)gridDim.x * blockDim.x) {
    const float sv = sens[i];
    rho[i] = oc_rule_log(sv, __builtin_amdgcn_logf(-sv), rho[i], lg, c01, min_rho, step, damp);
  }
}

// Density initialisers (Framework.cu:61-145). tanproj(v, 20, 0.5) is the
// smooth projection of lib.cuh:1372; TPMS fields are clamped to
// [0.001, 1], then everything to [0.1, 1] (Framework.cu:144).
__device__ inline float tanproj(float v, float beta) {
  const float t = tanhf(beta * 0.5f);
  return (t + tanhf(beta * (v - 0.5f))) / (t + tanhf(beta * 0.5f));
}

__global__ void __launch_bounds__(kBlock) k_init_density(Geom g, int kind, InitModes M, float* __restrict__ rho) {
  const long long t = gtid();
  if (t >= g.nn) return;
  int i, j, k;
  split_node(g.nx, g.ny, t, i, j, k);
  const float pi = 3.1415926f;
  const float x = float(i) / g.nx, y = float(j) / g.ny, z = float(k) / g.nz;
  float v = 0.f;
  switch (kind) {
    case 6: {  // P
      const float val = cosf(2 * pi * x) + cosf(2 * pi * y) + cosf(2 * pi * z);
      v = fmaxf(fminf(tanproj(-val, 20.f), 1.f), 0.001f);
      break;
    }
    case 7: {  // G
      const float val = sinf(2 * pi * x) * cosf(2 * pi * y) + sinf(2 * pi * z) * cosf(2 * pi * x) +
                        sinf(2 * pi * y) * cosf(2 * pi * z);
      v = fmaxf(fminf(tanproj(val, 20.f), 1.f), 0.001f);
      break;
    }
    case 8: {  // D
      const float val = cosf(2 * pi * x) * cosf(2 * pi * y) * cosf(2 * pi * z) -
                        sinf(2 * pi * x) * sinf(2 * pi * y) * sinf(2 * pi * z);
      v = fmaxf(fminf(tanproj(val, 20.f), 1.f), 0.001f);
      break;
    }
    case 9: {  // IWP
      const float val = 2 * (cosf(2 * pi * x) * cosf(2 * pi * y) + cosf(2 * pi * y) * cosf(2 * pi * z) +
                             cosf(2 * pi * z) * cosf(2 * pi * x)) -
                        (cosf(4 * pi * x) + cosf(4 * pi * y) + cosf(4 * pi * z));
      v = fmaxf(fminf(tanproj(val, 20.f), 1.f), 0.001f);
      break;
    }
    case 10: {  // example: a void sphere of radius min(reso)/6 in a 0.3 matrix
      const float dx = i - g.nx / 2.f + 0.5f, dy = j - g.ny / 2.f + 0.5f, dz = k - g.nz / 2.f + 0.5f;
      const float rad = fminf(fminf((float)g.nx, (float)g.ny), (float)g.nz) / 6.f;
      v = sqrtf(dx * dx + dy * dy + dz * dz) < rad ? 0.f : 0.3f;
      break;
    }
    default: {  // seeded periodic cosine modes around the volume ratio
      float s = 0.f;
      for (int m = 0; m < M.count; m++)
        s += M.amp[m] * cosf(2 * pi * (M.k[m][0] * x + M.k[m][1] * y + M.k[m][2] * z) + M.phase[m]);
      v = M.mean + s;
    }
  }
  rho[t] = fminf(fmaxf(v, 0.1f), 1.f);
}

void launch_init_density(hipStream_t s, const Geom& g, int kind, const InitModes& M, float* rho) {
  hipLaunchKernelGGL(k_init_density, dim3(nblk(g.nn)), dim3(kBlock), 0, s, g, kind, M, rho);
}

// ---- random trigonometric start (randTri, Framework.cuh:238-310) -------------
// Node value s = sum_i c_i phi_i + sum_{j,k} c_{6P + 6P j + k} psi_j psi_k over
// the basis of randTribase_sincos_kernel (Framework.cuh:106-214): with
// p = node / (edge - 1) per axis (optionally centred and rotated),
// C[a][f] = cos(2 pi (f+1) p_a), S[a][f] = sin(2 pi (f+1) p_a);
// first order: c_i C[i%3][i/3] for i < 3P, c_{3P+i} S[i%3][i/3];
// second order: psi_j = C[j/P][j%P] for j < 3P, S[j/P-3][j%P] otherwise.
template <int P>
__global__ void __launch_bounds__(kBlock) k_rand_tri(Geom g, int centered, RandTriRot R,
                                                     const float* __restrict__ coeffs, float* __restrict__ out) {
  constexpr int NB1 = 6 * P, NB = 6 * P + 36 * P * P;
  __shared__ float cs[NB];
  for (int i = threadIdx.x; i < NB; i += blockDim.x) cs[i] = coeffs[i];
  __syncthreads();
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nn) return;
  int ix, iy, iz;
  split_node(g.nx, g.ny, t, ix, iy, iz);
  float p[3] = {(float)ix / ((float)g.nx - 1.f), (float)iy / ((float)g.ny - 1.f), (float)iz / ((float)g.nz - 1.f)};
  if (centered) {
    const float v0 = p[0] - 0.5f, v1 = p[1] - 0.5f, v2 = p[2] - 0.5f;
    p[0] = R.m[0] * v0 + R.m[1] * v1 + R.m[2] * v2;
    p[1] = R.m[3] * v0 + R.m[4] * v1 + R.m[5] * v2;
    p[2] = R.m[6] * v0 + R.m[7] * v1 + R.m[8] * v2;
  }
  const float pi2 = 3.1415926f * 2.f;
  float psi[NB1];  // psi[a*P + f] = C[a][f], psi[3P + a*P + f] = S[a][f]
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int f = 0; f < P; f++) {
      psi[a * P + f] = cosf(pi2 * (float)(f + 1) * p[a]);
      psi[3 * P + a * P + f] = sinf(pi2 * (float)(f + 1) * p[a]);
    }
  float s = 0.f;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int f = 0; f < P; f++) {
      s += cs[3 * f + a] * psi[a * P + f];
      s += cs[3 * P + 3 * f + a] * psi[3 * P + a * P + f];
    }
#pragma unroll
  for (int j = 0; j < NB1; j++) {
    const float* row = cs + NB1 + NB1 * j;
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < NB1; k++) acc += row[k] * psi[k];
    s += acc * psi[j];
  }
  out[t] = s;
}

__device__ inline float level_map(float val, float t, float hi) {
  // sigmoid(val, 15, t) (lib.cuh:1381) scaled to [0.001, hi] (Framework.cuh:287-293)
  const float v = 1.f / (1.f + expf(-15.f * (val - t)));
  return v * (hi - 0.001f) + 0.001f;
}

__global__ void __launch_bounds__(kBlock) k_level_sum(long long n, const float* __restrict__ val, float t, float hi,
                                                      double* __restrict__ part) {
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    acc += (double)level_map(val[i], t, hi);
  __shared__ double sh[kBlock];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kBlock) k_level_apply(long long n, float* __restrict__ val, float t, float hi) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) val[i] = level_map(val[i], t, hi);
}

// Counter-based uniform in [0, 1): the top 24 bits of splitmix64 of
// (seed << 32) + index (one output per node, reproducible on any device).
__device__ inline float hash_uniform(unsigned long long seed, unsigned long long i) {
  unsigned long long z = (seed << 32) + i + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (float)(z >> 40) * (1.f / 16777216.f);
}

__global__ void __launch_bounds__(kBlock) k_noise(long long n, unsigned long long seed, float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = hash_uniform(seed, (unsigned long long)i);
}

// projectTensor_kernel (TensorExpression.cu:106-115) with beta 20, tau 0.5:
// tanproj, then below 0.5 -> 1e-9, above -> 1.
__global__ void __launch_bounds__(kBlock) k_project_binary(long long n, float* __restrict__ v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float r = tanproj(v[i], 20.f);
  if (r < 0.5f) r = 1e-9f;
  if (r > 0.5f) r = 1.f;
  v[i] = r;
}

// rho = clamp(rho * s, lo, hi)
__global__ void __launch_bounds__(kBlock) k_scale_clamp(long long n, float* __restrict__ v, float s, float lo,
                                                        float hi) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float r = v[i] * s;
  if (r < lo) r = lo;
  if (r >= hi) r = hi;
  v[i] = r;
}

int rand_tri_max_period() { return 10; }

void launch_rand_tri(hipStream_t s, const Geom& g, int P, int centered, const RandTriRot& R, const float* coeffs,
                     float* out) {
  const dim3 grid(nblk(g.nn)), block(kBlock);
  switch (P) {
    case 1: hipLaunchKernelGGL(k_rand_tri<1>, grid, block, 0, s, g, centered, R, coeffs, out); break;
    case 2: hipLaunchKernelGGL(k_rand_tri<2>, grid, block, 0, s, g, centered, R, coeffs, out); break;
    case 3: hipLaunchKernelGGL(k_rand_tri<3>, grid, block, 0, s, g, centered, R, coeffs, out); break;
    case 4: hipLaunchKernelGGL(k_rand_tri<4>, grid, block, 0, s, g, centered, R, coeffs, out); break;
    case 5: hipLaunchKernelGGL(k_rand_tri<5>, grid, block, 0, s, g, centered, R, coeffs, out); break;
    case 6: hipLaunchKernelGGL(k_rand_tri<6>, grid, block, 0, s, g, centered, R, coeffs, out); break;
    case 7: hipLaunchKernelGGL(k_rand_tri<7>, grid, block, 0, s, g, centered, R, coeffs, out); break;
    case 8: hipLaunchKernelGGL(k_rand_tri<8>, grid, block, 0, s, g, centered, R, coeffs, out); break;
    case 9: hipLaunchKernelGGL(k_rand_tri<9>, grid, block, 0, s, g, centered, R, coeffs, out); break;
    case 10: hipLaunchKernelGGL(k_rand_tri<10>, grid, block, 0, s, g, centered, R, coeffs, out); break;
    default: throw std::invalid_argument("randTri period must be 1..10");
  }
}
unsigned level_sum_blocks(long long n) {
  const long long b = (n + kBlock - 1) / kBlock;
  return (unsigned)(b > 512 ? 512 : (b < 1 ? 1 : b));
}
void launch_level_sum(hipStream_t s, long long n, const float* val, float t, float hi, double* part) {
  hipLaunchKernelGGL(k_level_sum, dim3(level_sum_blocks(n)), dim3(kBlock), 0, s, n, val, t, hi, part);
}
void launch_level_apply(hipStream_t s, long long n, float* val, float t, float hi) {
  hipLaunchKernelGGL(k_level_apply, dim3(nblk(n)), dim3(kBlock), 0, s, n, val, t, hi);
}
void launch_noise(hipStream_t s, long long n, unsigned long long seed, float* out) {
  hipLaunchKernelGGL(k_noise, dim3(nblk(n)), dim3(kBlock), 0, s, n, seed, out);
}
void launch_project_binary(hipStream_t s, long long n, float* v) {
  hipLaunchKernelGGL(k_project_binary, dim3(nblk(n)), dim3(kBlock), 0, s, n, v);
}
void launch_scale_clamp(hipStream_t s, long long n, float* v, float scale, float lo, float hi) {
  hipLaunchKernelGGL(k_scale_clamp, dim3(nblk(n)), dim3(kBlock), 0, s, n, v, scale, lo, hi);
}

// Reach of the taps (largest offset along any axis) and the sum of all
// weights in tap order, in FP32 as the kernel adds them.
static void tap_stats(const Taps& T, int& h, float& wall) {
  h = 0;
  wall = 0.f;
  for (int i = 0; i < T.n; i++) {
    h = std::max(h, std::max(std::abs((int)T.dx[i]), std::max(std::abs((int)T.dy[i]), std::abs((int)T.dz[i]))));
    wall += T.w[i];
  }
}

template <int MODE>
static void launch_filter_tile(hipStream_t s, const Geom& g, const Taps& T, const float* a, const float* b,
                               float* out, float* W, bool w_known = false) {
  int h;
  float wall;
  tap_stats(T, h, wall);
  if (h > kFtH || T.n > kMaxTaps) throw std::invalid_argument("filter taps reach beyond 3 cells");
  if (h > g.nx || h > g.ny || h > g.nz) throw std::invalid_argument("filter reach exceeds the cell");
  const unsigned nb = (unsigned)(((g.nx + kFtX - 1) / kFtX) * ((g.ny + kFtY - 1) / kFtY) * ((g.nz + kFtZ - 1) / kFtZ));
  hipLaunchKernelGGL(k_filter_tile<MODE>, dim3(nb), dim3(kBlock), 0, s, g, T, h, wall, a, b, out, W,
                     (int)(w_known && W));
}

void launch_filter_fwd(hipStream_t s, const Geom& g, const Taps& T, const float* in, float* out, float* W,
                       bool w_known) {
  launch_filter_tile<FILT_FWD>(s, g, T, in, nullptr, out, W, w_known);
}
void launch_filter_adj(hipStream_t s, const Geom& g, const Taps& T, const float* in, const float* W, float* out) {
  launch_filter_tile<FILT_ADJ>(s, g, T, in, W, out, nullptr);
}
void launch_simp_fwd(hipStream_t s, long long n, const float* rho, float k0, float k1, float floor_, float* kappa,
                     float* dkappa) {
  hipLaunchKernelGGL(k_simp_fwd, dim3(nblk(n)), dim3(kBlock), 0, s, n, rho, k0, k1, floor_, kappa, dkappa);
}
void launch_mul(hipStream_t s, long long n, const float* a, const float* b, float* out) {
  hipLaunchKernelGGL(k_mul, dim3(nblk(n)), dim3(kBlock), 0, s, n, a, b, out);
}
void launch_oc_filter_sens(hipStream_t s, const Geom& g, const Taps& T, const float* sens, const float* rho,
                           float* out) {
  launch_filter_tile<FILT_SENS>(s, g, T, sens, rho, out, nullptr);
}
void launch_maxabs(hipStream_t s, long long n, const float* a, unsigned int* out_bits) {
  unsigned nb = nblk(n);
  if (nb > 512) nb = 512;
  hipLaunchKernelGGL(k_maxabs, dim3(nb), dim3(kBlock), 0, s, n, a, out_bits);
}
void launch_oc_begin(hipStream_t s, long long n, const float* sens, unsigned int* bits, OcState* st) {
  launch_maxabs(s, n, sens, bits);
  hipLaunchKernelGGL(k_oc_init, dim3(1), dim3(1), 0, s, bits, st);
}
void launch_oc_tree_pass(hipStream_t s, long long n, const float* sens, const float* rho, OcState* st,
                         float min_rho, float step, float damp, double volratio, double vol_tol, int max_steps,
                         const Reduce& R, unsigned nblocks) {
  const float c01 = powf(0.01f, damp);
  hipLaunchKernelGGL(k_oc_tree_sum, dim3(nblocks), dim3(kBlock), 0, s, n, sens, rho, st, min_rho, step, damp, c01,
                     max_steps, R);
  hipLaunchKernelGGL(k_oc_tree_final, dim3(1), dim3(kBlock), 0, s, st, R, (int)nblocks, 1.0 / (double)n, volratio,
                     vol_tol, max_steps);
}
void launch_oc_apply(hipStream_t s, long long n, const float* sens, float* rho, const OcState* st, float min_rho,
                     float step, float damp, int max_steps, int force, OcHost* hm, unsigned tag) {
  const float c01 = powf(0.01f, damp);
  unsigned nb = nblk(n);
  if (nb > 8192) nb = 8192;
  hipLaunchKernelGGL(k_oc_apply, dim3(nb), dim3(kBlock), 0, s, n, sens, rho, st, min_rho, step, damp, c01, max_steps,
                     force, hm, tag);
}

// Orbit average under a symmetry group (symetrize_tensor_kernel,
// TensorExpression.cuh:70-300, driven by symmetrizeField, Framework.cuh:221-235):
// out[v] = mean of in over the orbit of v. Every member of an orbit sums
// the orbit in the same order (group order applied to the orbit's smallest
// index), so the members get bitwise the same value; reading `in` and
// writing `out` keeps the elements independent.
__global__ void __launch_bounds__(kBlock) k_symmetrize(Geom g, SymGroup G, const float* __restrict__ in,
                                                       float* __restrict__ out) {
  const long long t = gtid();
  if (t >= g.nn) return;
  const int dims[3] = {g.nx, g.ny, g.nz};
  const int c[3] = {(int)(t % g.nx), (int)((t / g.nx) % g.ny), (int)(t / ((long long)g.nx * g.ny))};
  auto image = [&](int i, const int* p, int* q) {
    for (int j = 0; j < 3; j++) {
      const int v = p[G.perm[i][j]];
      q[j] = ((G.flip[i] >> j) & 1) ? dims[j] - 1 - v : v;
    }
  };
  auto lin3 = [&](const int* q) { return (long long)q[0] + (long long)g.nx * ((long long)q[1] + (long long)g.ny * q[2]); };
  // canonical representative: the orbit member with the smallest index
  int r[3] = {c[0], c[1], c[2]};
  long long rbest = t;
  for (int i = 0; i < G.n; i++) {
    int q[3];
    image(i, c, q);
    const long long id = lin3(q);
    if (id < rbest) {
      rbest = id;
      r[0] = q[0];
      r[1] = q[1];
      r[2] = q[2];
    }
  }
  float sum = 0.f;
  for (int i = 0; i < G.n; i++) {
    int q[3];
    image(i, r, q);
    sum += in[lin3(q)];
  }
  out[t] = sum / (float)G.n;
}

void launch_symmetrize(hipStream_t s, const Geom& g, const SymGroup& G, const float* in, float* out) {
  hipLaunchKernelGGL(k_symmetrize, dim3(nblk(g.nn)), dim3(kBlock), 0, s, g, G, in, out);
}

}  // namespace otm

namespace otm {

// ---- full reductions of a field (reduce_tsexp_t, TensorExpression.h:1342-1402) ----
// Block partials in the slot of the block's item range (deterministic
// order), FP64 accumulation; the host sums the partials in order.
__global__ void __launch_bounds__(kBlock) k_field_max(long long n, const float* __restrict__ x,
                                                      double* __restrict__ part) {
  __shared__ double sh[kWaves];
  double m = -INFINITY;
  const BlockRange br = stream_range(n);
  for (long long i = br.b + threadIdx.x; i < br.e; i += blockDim.x) m = fmax(m, (double)x[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kWaves; w++) m = fmax(m, sh[w]);
    part[(int)blockIdx.x] = m;
  }
}

template <int KIND>
__global__ void __launch_bounds__(kBlock) k_field_sum(long long n, const float* __restrict__ x, double p,
                                                      double shift, double* __restrict__ part) {
  __shared__ double sh[kWaves];
  double v[1] = {0.0};
  const BlockRange br = stream_range(n);
  for (long long i = br.b + threadIdx.x; i < br.e; i += blockDim.x) {
    const double xi = (double)x[i];
    v[0] += KIND == kReduceSum ? xi : (KIND == kReducePnorm ? pow(xi, p) : exp(xi - shift));
  }
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) part[(int)blockIdx.x] = v[0];
}

// grad_i = upstream * d value / d x_i: 1, (x_i / value)^(p-1), exp(x_i - value)
// (AssociativeOp::diff, TensorExpression.h:1220-1229)
template <int KIND>
__global__ void __launch_bounds__(kBlock) k_field_grad(long long n, const float* __restrict__ x, double p,
                                                       double value, double upstream, float* __restrict__ g) {
  const long long i = stid();
  if (i >= n) return;
  const double xi = (double)x[i];
  const double d = KIND == kReduceSum ? 1.0
                                          : (KIND == kReducePnorm ? pow(xi / value, p - 1.0) : exp(xi - value));
  g[i] = (float)(upstream * d);
}

void launch_field_max(hipStream_t s, long long n, const float* x, double* part, unsigned nb) {
  hipLaunchKernelGGL(k_field_max, dim3(nb), dim3(kBlock), 0, s, n, x, part);
}
void launch_field_sum(hipStream_t s, int kind, long long n, const float* x, double p, double shift, double* part,
                      unsigned nb) {
  if (kind == kReduceSum)
    hipLaunchKernelGGL(k_field_sum<kReduceSum>, dim3(nb), dim3(kBlock), 0, s, n, x, p, shift, part);
  else if (kind == kReducePnorm)
    hipLaunchKernelGGL(k_field_sum<kReducePnorm>, dim3(nb), dim3(kBlock), 0, s, n, x, p, shift, part);
  else
    hipLaunchKernelGGL(k_field_sum<kReduceLogSumExp>, dim3(nb), dim3(kBlock), 0, s, n, x, p, shift, part);
}
void launch_field_grad(hipStream_t s, int kind, long long n, const float* x, double p, double value, double upstream,
                       float* g) {
  const unsigned nb = (unsigned)((n + kBlock - 1) / kBlock);
  if (kind == kReduceSum)
    hipLaunchKernelGGL(k_field_grad<kReduceSum>, dim3(nb), dim3(kBlock), 0, s, n, x, p, value, upstream, g);
  else if (kind == kReducePnorm)
    hipLaunchKernelGGL(k_field_grad<kReducePnorm>, dim3(nb), dim3(kBlock), 0, s, n, x, p, value, upstream, g);
  else
    hipLaunchKernelGGL(k_field_grad<kReduceLogSumExp>, dim3(nb), dim3(kBlock), 0, s, n, x, p, value, upstream,
                       g);
}

}  // namespace otm
