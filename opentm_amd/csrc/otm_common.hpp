// otm_common.hpp -- shared device/host definitions for the gfx950 engine.
//
// Layout of every level: a periodic nx*ny*nz lattice, lexicographic with x
// fastest (index = x + nx*(y + ny*z)). Nodes and elements share the index:
// element (x,y,z) spans nodes (x..x+1, y..y+1, z..z+1) modulo the size. The
// three load cases are stored structure-of-arrays: right-hand side k of a
// level vector lives at base + k*nn.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define OTM_NRHS 3

namespace otm {

// Level geometry passed by value to every kernel (no global device state).
struct Geom {
  int nx, ny, nz;
  long long nn;  // nx*ny*nz
};

__host__ __device__ inline int wrapm(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// Value at byte offset `off` from a wave-uniform base: the load takes the
// base in scalar registers and one 32-bit offset register (a lattice has at
// most 2^30 nodes, so a field's byte offsets fit 32 bits).
__device__ inline float ld_off(const float* __restrict__ base, unsigned off) {
  return *(const float*)((const char*)base + off);
}

// (x, y, z) of lexicographic index t on an nx x ny x * grid, x fastest.
// Indices stay below 2^31 (every axis is at most 1024), so two 32-bit
// divisions replace the much longer 64-bit division sequences.
__device__ inline void split_node(int nx, int ny, long long t, int& x, int& y, int& z) {
  const unsigned u = (unsigned)t, r = u / (unsigned)nx, q = r / (unsigned)ny;
  x = (int)(u - r * (unsigned)nx);
  y = (int)(r - q * (unsigned)ny);
  z = (int)q;
}

// Trilinear prolongation from a coarse grid of half the size
// (prolongate_correction_kernel_H, grid.cu:881-940): fine node (x, y, z)
// takes coarse nodes X[a], Y[b], Z[c] with weight wx[a] wy[b] wz[c]; on an
// even coordinate the second weight is zero.
struct ProlongW {
  int X[2], Y[2], Z[2];
  float wx[2], wy[2], wz[2];
};
__device__ inline ProlongW prolong_weights(const Geom& cg, int x, int y, int z) {
  ProlongW P;
  P.X[0] = x >> 1;
  P.X[1] = (x & 1) ? ((x >> 1) + 1 == cg.nx ? 0 : (x >> 1) + 1) : P.X[0];
  P.wx[0] = (x & 1) ? 0.5f : 1.f;
  P.wx[1] = (x & 1) ? 0.5f : 0.f;
  P.Y[0] = y >> 1;
  P.Y[1] = (y & 1) ? ((y >> 1) + 1 == cg.ny ? 0 : (y >> 1) + 1) : P.Y[0];
  P.wy[0] = (y & 1) ? 0.5f : 1.f;
  P.wy[1] = (y & 1) ? 0.5f : 0.f;
  P.Z[0] = z >> 1;
  P.Z[1] = (z & 1) ? ((z >> 1) + 1 == cg.nz ? 0 : (z >> 1) + 1) : P.Z[0];
  P.wz[0] = (z & 1) ? 0.5f : 1.f;
  P.wz[1] = (z & 1) ? 0.5f : 0.f;
  return P;
}
// sum_I P(i, I) u_c(I) for one right-hand side ck of the coarse grid.
__device__ inline float prolong_sum(const Geom& cg, const float* __restrict__ ck, const ProlongW& P) {
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const float w = P.wx[a] * P.wy[b] * P.wz[c];
        if (w != 0.f)
          s += w * ld_off(ck, 4u * ((unsigned)P.X[a] + (unsigned)cg.nx * ((unsigned)P.Y[b] + (unsigned)cg.ny * (unsigned)P.Z[c])));
      }
  return s;
}

// Element constants of the fine level (closed forms of the reference's
// quadrature, templateMatrix.cpp:21-87): Ke = c * P with P = 4 on the
// diagonal, 0 between nodes that differ along one axis, -1 otherwise,
// c = h/12. disp / feMu are the reference's 8x3 tables (local node, load).
struct FineConsts {
  float c;            // h/12
  float disp[8][3];   // element-local macro temperature per load case
  float femu[8][3];   // element load vectors
};

// Coarse-level Galerkin constants: G[c][t] = (Q_c^T Ke Q_c) packed upper
// triangle (t over a<=b) for child position c of a 2x coarsened element.
struct GalerkinConsts {
  float G[8][36];
};

// Packed upper-triangle index of (a,b), a<=b, of an 8x8 symmetric matrix.
__host__ __device__ constexpr int tri8(int a, int b) {
  return a <= b ? (a * 8 - (a * (a - 1)) / 2 + (b - a)) : (b * 8 - (b * (b - 1)) / 2 + (a - b));
}

// Scalars of the three lock-step conjugate-gradient recurrences (FP64).
struct PcgScalars {
  double rz[OTM_NRHS];     // r.z of the current iteration
  double pq[OTM_NRHS];     // p.Kp
  double rr[OTM_NRHS];     // r.r
  double bb[OTM_NRHS];     // b.b (reference norm)
  double alpha[OTM_NRHS];
  double beta[OTM_NRHS];
  double mean[OTM_NRHS];   // scratch for mean removal
  double thr[OTM_NRHS];    // r.r below which the current recurrence stops
  int active[OTM_NRHS];    // 1 while the right-hand side is iterating
  int fault;               // bit 0: NaN/Inf in a reduction; bit 1: preconditioner not positive
};

// Reduction finalisers run by the last block of a reduction kernel.
enum FinalMode : int {
  FIN_STORE = 0,     // store sums into out[0..nv)
  FIN_BB_RR = 1,     // sums = {b.b x3, r.r x3}: initial norms
  FIN_RZ_FIRST = 2,  // sums = r.z x3: rz = sum, beta = 0
  FIN_RZ = 3,        // sums = r.z x3: beta = sum/rz, rz = sum
  FIN_PQ = 4,        // sums = p.q x3: alpha = rz/pq
  FIN_RR = 5,        // sums = r.r x3: update activity against thr
  FIN_MEAN = 6,      // sums = sum(x) x3: mean = sum / nn
};

// Results the host waits for, in coherent (fine-grained) pinned host
// memory: the finaliser copies them here and then stores its tag, so the
// host learns them by polling the tag instead of a copy plus a stream
// synchronisation (about half the round trip on this GPU).
struct HostMirror {
  PcgScalars pcg;    // PCG modes
  double out[16];    // FIN_STORE values
  unsigned seq;      // tag of the last finaliser that wrote this block
};

struct Reduce {
  double* partials;      // [nblocks * nv]
  double* out;           // FIN_STORE target
  PcgScalars* pcg;       // for PCG modes
  double rtol2;          // squared relative tolerance for FIN_RR
  double inv_nn;         // 1/nn for FIN_MEAN
  double inner2 = 0.0;   // FIN_BB_RR: also stop once r.r falls by this factor (refinement)
  int mode;
  HostMirror* hm = nullptr;  // when set, the results also go to the host with tag
  unsigned tag = 0;
};

constexpr int kBlock = 256;       // threads per block for streaming kernels
constexpr int kMaxReduce = 9;     // values per reduction
constexpr int kMaxReduceBlocks = 2048;  // partial-sum slots of a reduction (the largest grid that reduces)
constexpr int kWaves = kBlock / 64;

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, and
// each XCD has its own L2. The logical block index below gives every XCD a
// contiguous range of blocks, so neighbouring tiles -- and the halo cache
// lines they share -- stay on one L2. The mapping is a bijection for any
// grid size.
constexpr int kXcds = 8;
__device__ inline int xcd_block() {
  const int b = blockIdx.x, nb = gridDim.x;
  const int q = nb / kXcds, r = nb % kXcds;
  const int x = b % kXcds, local = b / kXcds;
  const int start = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  return start + local;
}

// Global thread index of one-item-per-thread kernels, XCD-ordered.
__device__ inline long long gtid() { return (long long)xcd_block() * blockDim.x + threadIdx.x; }

// Contiguous item range [b, e) of block lb of nb blocks for fixed-grid
// streaming loops (block_range: this launch's XCD-ordered block).
struct BlockRange {
  long long b, e;
};
__device__ inline BlockRange block_range_v(long long n, int lb, int nb) {
  const long long per = ((n + nb - 1) / nb + blockDim.x - 1) / blockDim.x * blockDim.x;
  long long s = (long long)lb * per, e = s + per;
  if (s > n) s = n;
  if (e > n) e = n;
  return {s, e};
}
__device__ inline BlockRange block_range(long long n) { return block_range_v(n, xcd_block(), gridDim.x); }

// Streaming kernels (every item read and written once, no data shared
// between blocks) index in the hardware's dispatch order instead: the
// blocks running at one time then cover one region of each array together.
// XCD-contiguous ranges measured 13% lower HBM bandwidth on a read of ten
// FP32/FP64 streams (4.8 against 5.5 TB/s, 1024-2048 blocks).
__device__ inline long long stid() { return (long long)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ inline BlockRange stream_range(long long n) { return block_range_v(n, blockIdx.x, gridDim.x); }

// A load of a value this pass reads once, marked non-temporal (the caches
// do not keep it for reuse).
template <typename T>
__device__ inline T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Sum over aligned groups of 8 lanes; every lane of a group receives the
// same sum (IEEE addition commutes). Cross-lane moves through DPP: swap
// pairs (quad_perm 1,0,3,2), swap pairs of pairs (quad_perm 2,3,0,1), then
// the mirrored half row (lane i of 8 takes lane 7 - i).
__device__ inline float group8_sum(float v) {
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));
  return v;
}

// Lane i's value moved through a DPP pattern (both halves of the double).
template <int CTRL>
__device__ inline double dpp_f64(double v) {
  const unsigned long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ inline double readlane_f64(double v, int lane) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Wave64 sum (every lane gets it; a fixed order). Inside each row of 16
// lanes through DPP: swap pairs (quad_perm 1,0,3,2), pairs of pairs
// (quad_perm 2,3,0,1), the mirrored half row, the mirrored row; then the
// four row sums are read from lanes 0, 16, 32 and 48. (Cross-lane shuffles
// went through the LDS crossbar, one round trip per step.)
__device__ inline double wave_sum(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  v += dpp_f64<0x140>(v);
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// Block-wide sum of NV values; result valid in thread 0. `sh` holds
// NV*kWaves doubles.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double s = wave_sum(v[i]);
    if (lane == 0) sh[i * kWaves + wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; i++) {
      double s = 0;
      for (int w = 0; w < kWaves; w++) s += sh[i * kWaves + w];
      v[i] = s;
    }
  }
}

// Deterministic two-level reduction, stage 1: every block stores its NV
// partial sums (FP64). Stage 2 (k_reduce_final, launched right after on
// the same stream) sums the partials in block order and applies the
// finaliser; the kernel boundary makes the partials visible without
// device-scope fences inside the producing kernel (on this GPU an
// agent-scope release writes back the XCD's L2, which measured several
// times the cost of the separate one-block launch).
template <int NV>
__device__ inline void reduce_block_at(double (&v)[NV], const Reduce& R, int slot) {
  __shared__ double sh[NV * kWaves];
  block_sum<NV>(v, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; i++) R.partials[(size_t)slot * NV + i] = v[i];
  }
}
// The partial's slot is the block's range index, so the finaliser adds the
// partials in item order whatever the launch shape (stream_range order).
template <int NV>
__device__ inline void reduce_block(double (&v)[NV], const Reduce& R) {
  reduce_block_at<NV>(v, R, (int)blockIdx.x);
}

// Copies a finaliser's results to the host block and publishes the tag
// (one thread, after finalize). The system-scope release orders the copy
// before the tag for a host that polls it.
__device__ inline void publish_host(const double* acc, int nv, const Reduce& R) {
  if (!R.hm) return;
  if (R.mode == FIN_STORE) {
    for (int i = 0; i < nv && i < 16; i++) R.hm->out[i] = acc[i];
  } else {
    R.hm->pcg = *R.pcg;
  }
  __hip_atomic_store(&R.hm->seq, R.tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Applies a finaliser to the summed values (one thread).
__device__ inline void finalize(const double* acc, int nv, const Reduce& R) {
  PcgScalars* P = R.pcg;
  switch (R.mode) {
    case FIN_STORE:
      for (int i = 0; i < nv; i++) R.out[i] = acc[i];
      break;
    case FIN_BB_RR: {
      // sums = {b.b x3, r.r x3, sum(r) x3}. The mean of r is projected out
      // afterwards (a periodic cell's system is consistent only for
      // mean-free right-hand sides), so r.r is reported for the projected r.
      // A load vector that is zero up to rounding (e.g. a gradient along a
      // laminate) is measured against the largest of the three: its
      // reference norm is floored at 1e-3 of the largest norm (the
      // reference floors the norm with + 1e-10, MG.cpp:100).
      // A (re)start of the PCG begins here, so the fault flags start over
      // from this residual alone.
      double bmax = 0.0;
      int fault = 0;
      for (int k = 0; k < OTM_NRHS; k++) bmax = acc[k] > bmax ? acc[k] : bmax;
      for (int k = 0; k < OTM_NRHS; k++) {
        double bb = acc[k] > 1e-6 * bmax ? acc[k] : 1e-6 * bmax;
        if (!(bb > 0.0)) bb = 1.0;
        const double mean = acc[2 * OTM_NRHS + k] * R.inv_nn;
        double rr = acc[OTM_NRHS + k] - mean * acc[2 * OTM_NRHS + k];
        if (rr < 0.0) rr = 0.0;
        P->mean[k] = mean;
        P->bb[k] = bb;
        P->rr[k] = rr;
        if (!(acc[OTM_NRHS + k] - acc[OTM_NRHS + k] == 0.0)) fault = 1;
        P->active[k] = (rr > R.rtol2 * bb) ? 1 : 0;
        // an inner solve of the FP64 iterative refinement stops at a fixed
        // reduction of its own starting residual; a plain solve at rtol
        const double t_abs = R.rtol2 * bb, t_rel = R.inner2 * rr;
        P->thr[k] = t_abs > t_rel ? t_abs : t_rel;
        P->alpha[k] = 0.0;
        P->beta[k] = 0.0;
      }
      P->fault = fault;
      break;
    }
    case FIN_RZ_FIRST:
      for (int k = 0; k < OTM_NRHS; k++) {
        P->rz[k] = acc[k];
        P->beta[k] = 0.0;
        if (P->active[k] && !(acc[k] > 0.0)) P->fault |= 2;  // preconditioner not positive
      }
      break;
    case FIN_RZ:
      for (int k = 0; k < OTM_NRHS; k++) {
        P->beta[k] = (P->active[k] && P->rz[k] != 0.0) ? acc[k] / P->rz[k] : 0.0;
        P->rz[k] = acc[k];
        if (P->active[k] && !(acc[k] > 0.0)) P->fault |= 2;
      }
      break;
    case FIN_PQ:
      for (int k = 0; k < OTM_NRHS; k++) {
        P->pq[k] = acc[k];
        P->alpha[k] = (P->active[k] && acc[k] > 0.0) ? P->rz[k] / acc[k] : 0.0;
        if (P->active[k] && !(acc[k] > 0.0)) P->fault |= 2;
      }
      break;
    case FIN_RR:
      for (int k = 0; k < OTM_NRHS; k++) {
        P->rr[k] = acc[k];
        if (!(acc[k] - acc[k] == 0.0)) P->fault |= 1;  // NaN or Inf
        if (!(acc[k] > P->thr[k])) P->active[k] = 0;
      }
      break;
    case FIN_MEAN:
      for (int k = 0; k < OTM_NRHS; k++) P->mean[k] = acc[k] * R.inv_nn;
      break;
  }
}

}  // namespace otm
