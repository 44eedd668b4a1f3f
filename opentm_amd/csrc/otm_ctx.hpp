// otm_ctx.hpp -- engine-internal definitions shared by the single-cell
// engine (otm_engine.hip) and the batched runtime (otm_batch.hip): error
// plumbing, device allocation helpers, the multigrid level record and the
// context of one periodic cell. Not part of the public interface.
#pragma once

#include <string>
#include <vector>

#include "../../include/opentm.h"
#include "otm_coarse_dev.hpp"
#include "otm_fine_dev.hpp"
#include "otm_kernels.hpp"

namespace otm {

struct OtmError {
  int code;
  std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw OtmError{code, msg}; }

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(OTM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Message of the last failed call on this thread (otm_last_error).
void set_last_error(const std::string& msg);

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return OTM_OK;
  } catch (const OtmError& e) {
    set_last_error(e.msg);
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return OTM_E_ARG;
  }
}

template <typename T>
T* dalloc(size_t count) {
  void* p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fail(OTM_E_OOM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
  }
  return static_cast<T*>(p);
}

inline void dfree(void* p) {
  if (p) (void)hipFree(p);
}

inline hipMemcpyKind kind_of(int dst_where, int src_where) {
  if (dst_where == OTM_DEVICE && src_where == OTM_DEVICE) return hipMemcpyDeviceToDevice;
  if (dst_where == OTM_DEVICE) return hipMemcpyHostToDevice;
  if (src_where == OTM_DEVICE) return hipMemcpyDeviceToHost;
  return hipMemcpyHostToHost;
}

inline void check_where(int w) {
  if (w != OTM_HOST && w != OTM_DEVICE) fail(OTM_E_ARG, "where must be OTM_HOST or OTM_DEVICE");
}

// Polls a tag that a kernel on stream s publishes to coherent host memory
// with a system-scope release, until it equals `tag`. The stream keeps
// running what is queued after that kernel. A stream that drains or fails
// without the tag is an error (e.g. a kernel fault).
inline void wait_tag(hipStream_t s, const unsigned* seq, unsigned tag) {
  unsigned spins = 0;
  while (__atomic_load_n(seq, __ATOMIC_ACQUIRE) != tag) {
    if ((++spins & 255u) == 0) {
      const hipError_t e = hipStreamQuery(s);
      if (e == hipSuccess) {
        if (__atomic_load_n(seq, __ATOMIC_ACQUIRE) == tag) return;
        fail(OTM_E_HIP, "kernel results were not published");
      }
      if (e != hipErrorNotReady) hip_check(e, "hipStreamQuery");
    }
    __builtin_ia32_pause();
  }
}

// Workgroups of a reducing streaming kernel: one per kBlock items, at most
// 1024 (four waves per SIMD), or the reduction's 2048 slots for 8M items and
// more. Measured: 2048 gave the 256^3 precision loop 0.7% (more loads in
// flight per CU), and cost the 128^3 OC loop 1% (twice the partials to sum
// in every final for 8 items per thread).
__host__ __device__ inline unsigned red_blocks_for(long long n) {
  long long b = (n + kBlock - 1) / kBlock;
  const long long cap = n >= (1ll << 23) ? kMaxReduceBlocks : 1024;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (unsigned)b;
}

struct Level {
  Geom g;
  float* u = nullptr;    // correction (level 0: the PCG z vector)
  float* f = nullptr;    // right-hand side (level 0: the PCG residual)
  float* res = nullptr;  // residual before restriction
  float* st = nullptr;   // 27-point stencil (levels >= 1)
  float* E = nullptr;    // element matrices, 36 per element (levels >= 1)
  float* u2 = nullptr;   // second correction buffer for the brick smoother
  _Float16* st16 = nullptr;  // FP16 couplings (fp16_stencils)
  float* scale = nullptr;    // their per-level scale (device scalar)
  bool brick = false;    // smoothed with LDS bricks (edges multiples of 8)
  bool own_uf = true;
};

inline StRef stref(const Level& L) { return StRef{L.st, L.st16, L.scale}; }

}  // namespace otm

using namespace otm;  // internal header: the context below is written in the engine namespace's types

struct otm_ctx {
  otm_config cfg;
  int device = 0;
  hipStream_t s = nullptr;
  Geom g0;
  FineConsts fc;
  GalerkinConsts G;
  std::vector<Level> lv;
  // fine-level PCG state, 3 right-hand sides each
  float* kappa = nullptr;  // per element, with the 1e-9 floor
  float* x = nullptr;      // chi (warm start); the FP32 correction while refining
  double* x64 = nullptr;   // FP64 chi of the iterative refinement (fp64_refine)
  float* xlo = nullptr;    // low part of chi held as an FP32 pair (x + xlo) when not refining
  float *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  float* z2 = nullptr;  // second fine correction buffer (fine brick smoother)
  // coarsest dense solve
  double *cB = nullptr, *cBinv = nullptr;  // Gauss-Jordan work matrix, explicit inverse
  // reductions
  double* partials = nullptr;
  double* tile_partials = nullptr;  // the fine post-smoother's per-tile r.z
  double* red_out = nullptr;
  PcgScalars* pcg = nullptr;
  PcgScalars* pcg_host = nullptr;  // pinned
  double* out_host = nullptr;      // pinned, 8 doubles
  HostMirror* hm = nullptr;        // coherent pinned block the finalisers publish to
  HostMirror* hm_dev = nullptr;    // its device address
  unsigned hm_tag = 0;             // tag of the last finaliser asked to publish
  bool hm_pending = false;         // that finaliser's results not yet taken
  // design pipeline
  float *rho = nullptr, *rho_f = nullptr, *W = nullptr, *kraw = nullptr, *dkappa = nullptr;
  float *tmp = nullptr, *drho = nullptr;
  Taps taps;
  int taps_kind = -2;
  double taps_r = 0;
  int taps_periodic = -1;
  bool w_valid = false;  // W holds the weight sums of these taps (they depend on the taps alone)
  // live kernel timing (otm_profile_enable)
  bool prof_on = false;
  int prof_mask = 0;
  std::vector<hipEvent_t> ev_pool;
  struct Pending {
    int cls;
    double bytes;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  bool bricks_on = true;   // coarse levels use the LDS brick smoother
  int tail = -1;           // first level of the one-workgroup tail V-cycle
  int fine_tile = -1;      // fine brick smoother tile variant (kFineTileShapes), -1: none
  int post_dot_mode = -1;  // >= 0: the next V-cycle's fine post-smoother also reduces r.z (that finaliser)
  bool post_dot_done = false;
  int fallbacks = 0;       // solves that fell back to exact Gauss-Seidel
  long long prof_launches[OTM_PROF_NCLASSES] = {};
  double prof_ms[OTM_PROF_NCLASSES] = {};
  double prof_bytes[OTM_PROF_NCLASSES] = {};
  // status
  bool has_kappa = false, has_solution = false, has_drho = false;
  // Element energies of the current fluctuation fields, written by the
  // kappa^H kernel; every sensitivity of that solve is then a contraction
  // of them. energy_ok is cleared wherever the iterate changes.
  float* energy = nullptr;  // 6 x nn
  bool energy_ok = false;
  // Start residual of the next solve launched on s_side while the coarse
  // operators are rebuilt on s (otm_eval); s waits for it through ev_side.
  hipStream_t s_side = nullptr;
  hipEvent_t ev_main = nullptr, ev_side = nullptr;
  bool start_res_ready = false;
  // V-cycles captured as HIP graphs, keyed by the launch-relevant state (the
  // brick smoothers' buffer parities, the fused r.z mode, bricks on/off)
  struct VGraph {
    unsigned long long key;
    hipGraphExec_t exec;
  };
  std::vector<VGraph> vgraphs;
  bool graphs_failed = false;  // capture unavailable: run eagerly
  int iters[3] = {0, 0, 0};
  double relres[3] = {0, 0, 0};
  double C[9];

  // the fluctuation fields as the energy kernels read them: the FP64
  // iterate when refining, else the FP32 pair
  ChiRef chi_ref() const { return ChiRef{x, xlo, x64}; }
  Reduce reducer(int mode, double rtol2 = 0.0, double inv_nn = 0.0) const {
    Reduce R;
    R.partials = partials;
    R.out = red_out;
    R.pcg = pcg;
    R.rtol2 = rtol2;
    R.inv_nn = inv_nn;
    R.mode = mode;
    return R;
  }
  void sync() {
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (!pending.empty()) collect();
  }
  // A reducer whose finaliser also publishes its results to the host block
  // (taken by wait_host); one such reduction in flight at a time.
  Reduce reducer_host(int mode, double rtol2 = 0.0, double inv_nn = 0.0) {
    Reduce R = reducer(mode, rtol2, inv_nn);
    R.hm = hm_dev;
    R.tag = ++hm_tag;
    hm_pending = true;
    return R;
  }
  // Waits until the last publishing finaliser has run (wait_tag).
  void wait_host() {
    wait_tag(s, &hm->seq, hm_tag);
    hm_pending = false;
    // timing events are read at stream synchronisations; bound their number
    if (pending.size() > 512) sync();
  }
  hipEvent_t take_event() {
    if (ev_pool.empty()) {
      hipEvent_t e;
      hip_check(hipEventCreate(&e), "hipEventCreate");
      return e;
    }
    hipEvent_t e = ev_pool.back();
    ev_pool.pop_back();
    return e;
  }
  void collect() {
    for (auto& p : pending) {
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
      prof_ms[p.cls] += ms;
      prof_launches[p.cls] += 1;
      prof_bytes[p.cls] += p.bytes;
      ev_pool.push_back(p.a);
      ev_pool.push_back(p.b);
    }
    pending.clear();
  }
  // PCG scalars after the last reduction: from the host block when that
  // reduction published them, else by a copy and a synchronisation.
  void read_pcg() {
    if (hm_pending) {
      wait_host();
      *pcg_host = hm->pcg;
      return;
    }
    hip_check(hipMemcpyAsync(pcg_host, pcg, sizeof(PcgScalars), hipMemcpyDeviceToHost, s), "read pcg scalars");
    sync();
  }
};

// Records a start/stop event pair around the launches in its scope when
// timing is on; `bytes` is the algorithmic traffic of those launches.
struct ProfScope {
  otm_ctx* c;
  int cls;
  double bytes;
  hipEvent_t a = nullptr;
  ProfScope(otm_ctx* c_, int cls_, double bytes_) : c(c_), cls(cls_), bytes(bytes_) {
    if (c->prof_on && ((c->prof_mask >> cls) & 1)) {
      a = c->take_event();
      hip_check(hipEventRecord(a, c->s), "hipEventRecord");
    }
  }
  ~ProfScope() {
    if (a) {
      hipEvent_t b = c->take_event();
      if (hipEventRecord(b, c->s) == hipSuccess) c->pending.push_back({cls, bytes, a, b});
    }
  }
};

namespace otm {
// heatMatrix with exact multicolour Gauss-Seidel on every level; returns
// false when the solve ended above the relative residual target.
bool ctx_solve_exact(otm_ctx* c, double C[9]);
// Normalised radial filter taps (kind 0 Spline4, 1 Linear; radius r).
Taps make_taps(int kind, double r, int periodic);
// Level sizes (fine to coarse) of the hierarchy of a configuration; throws
// OTM_E_UNSUPPORTED for shapes the engine does not solve.
std::vector<Geom> plan_levels(const otm_config& cfg);
}  // namespace otm
