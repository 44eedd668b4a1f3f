// otm_batch.hip -- many cells solved together (batched and served forms).
//
// A batch runs the lock-step PCG of every member cell (each an ordinary
// otm_ctx with its own hierarchy and buffers) as one sequence of launches:
// each launch covers the same operation for all member cells -- one fine
// smoother launch for every cell's fine level, one coarse smoother launch
// for every cell's level l, one reduction launch for every cell's dot
// product -- so small cells fill the GPU together instead of one
// latency-bound launch at a time. A launch carries a table of (cell, first
// block) entries; a block finds its cell by binary search and runs the same
// device body as the single-cell kernel, with the cell's own block count
// and reduction order, so a cell's arithmetic is identical whatever else is
// in the batch.
//
// Cells join and leave at step boundaries. One step is one PCG iteration
// for every member: the cells that start or restart compute their FP64
// residual while the others apply the operator and update x and r; one
// host synchronisation reads every cell's scalars; then all active cells
// run their V-cycle and the r.z / p updates together.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "otm_coarse_dev.hpp"
#include "otm_ctx.hpp"
#include "otm_fine_dev.hpp"
#include "otm_vec_dev.hpp"

namespace otm {

constexpr int kMaxLevels = 10;  // 1024 -> 2 is ten levels

// Device view of one member cell.
struct CellDev {
  Geom g[kMaxLevels];
  int nlev;
  int tail;  // first level of the one-workgroup tail V-cycle
  int fine_bx;  // 32 or 16: fine brick width; 0: exact colour sweeps
  int tile_bx;  // 32 or 16: LDS tile width of the fine residual and apply; 0: pointwise kernels
  int sweeps_fine, sweeps_coarse;
  FineConsts fc;
  const float* kappa;
  float *x, *r, *p, *q;
  float* xlo;  // low part of chi as an FP32 pair (plain cells)
  double* x64;
  float* ua[kMaxLevels];  // correction buffer pair per level (level 0: z, z2)
  float* ub[kMaxLevels];
  float* f[kMaxLevels];
  float* res[kMaxLevels];
  StRef st[kMaxLevels];
  bool brick[kMaxLevels];
  const double* Binv;
  double* partials;
  double* red_out;
  PcgScalars* pcg;
  double rtol2, inv_nn, inner2;
  // sensitivity request
  float sw[6];
  float* sens;
  int sens_pitch;
  // setup (otm_batch_update): conductivity source and writable operators
  const float* ksrc;
  int kpitch;
  float* kappa_w;
  GalerkinConsts G;
  float* E[kMaxLevels];
  float* st_w[kMaxLevels];
  _Float16* st16_w[kMaxLevels];
  float* scale_w[kMaxLevels];
  double* cB;
  double* Binv_w;
};

// One entry of a launch table: blocks [block0, next entry's block0) work
// on cell `cell`; `par` holds the cell's buffer parity (bit l: level l's
// correction currently lives in ub), `aux` an index into a gather array.
struct Ent {
  int cell;
  int block0;
  int par;
  int aux;
};

__device__ inline float* ucur(const CellDev& C, int par, int l) { return ((par >> l) & 1) ? C.ub[l] : C.ua[l]; }
__device__ inline float* uoth(const CellDev& C, int par, int l) { return ((par >> l) & 1) ? C.ua[l] : C.ub[l]; }

__device__ inline Reduce cell_reduce(const CellDev& C, int mode) {
  Reduce R;
  R.partials = C.partials;
  R.out = C.red_out;
  R.pcg = C.pcg;
  R.rtol2 = C.rtol2;
  R.inv_nn = C.inv_nn;
  R.inner2 = C.inner2;    // used by FIN_BB_RR only, as in the single-cell solve
  R.mode = mode;
  return R;
}

__device__ inline int find_ent(const Ent* __restrict__ ents, int nents, int b) {
  int lo = 0, hi = nents - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ents[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <class Op>
__global__ void __launch_bounds__(Op::kThreads) kb_run(const Ent* __restrict__ ents, int nents,
                                                      const CellDev* __restrict__ cells, Op op) {
  const int b = xcd_block();
  const int i = find_ent(ents, nents, b);
  const Ent e = ents[i];
  const int end = i + 1 < nents ? ents[i + 1].block0 : (int)gridDim.x;
  op(cells[e.cell], e, b - e.block0, end - e.block0);
}

// Fine bricks keep the single-cell kernel's occupancy hint.
template <int BX>
__global__ void __launch_bounds__(BX * kFineTR) __attribute__((amdgpu_waves_per_eu(4, 8)))
kb_fine_brick(const Ent* __restrict__ ents, int nents, const CellDev* __restrict__ cells, int post) {
  const int b = xcd_block();
  const int i = find_ent(ents, nents, b);
  const Ent e = ents[i];
  const CellDev& C = cells[e.cell];
  const int sw = C.sweeps_fine;
  if (post)
    fine_brick_body<BX>(C.g[0], C.fc, C.kappa, C.f[0], ucur(C, e.par, 0), uoth(C, e.par, 0), sw, 0, 0,
                        b - e.block0);
  else
    fine_brick_body<BX>(C.g[0], C.fc, C.kappa, C.f[0], nullptr, ucur(C, e.par, 0), sw, 1, 1, b - e.block0);
}

// ---- operations (one functor per batched launch) ---------------------------
struct OpStBrick {
  static constexpr int kThreads = kBrickThreads;
  int l, post;
  __device__ void operator()(const CellDev& C, const Ent& e, int lb, int) const {
    const float* uin = post ? ucur(C, e.par, l) : nullptr;
    float* uout = post ? uoth(C, e.par, l) : ucur(C, e.par, l);
    if (C.st[l].f16)
      st_brick_body<true>(C.g[l], C.st[l], C.f[l], uin, uout, C.sweeps_coarse, post ? 0 : 1, post ? 0 : 1, lb);
    else
      st_brick_body<false>(C.g[l], C.st[l], C.f[l], uin, uout, C.sweeps_coarse, post ? 0 : 1, post ? 0 : 1, lb);
  }
};

struct OpFineGsColor {
  static constexpr int kThreads = kBlock;
  int cy, cz;
  __device__ void operator()(const CellDev& C, const Ent& e, int lb, int) const {
    fine_gs_color_body(C.g[0], C.fc, C.kappa, C.f[0], ucur(C, e.par, 0), cy, cz,
                       (long long)lb * kBlock + threadIdx.x);
  }
};

struct OpFillU {  // zero the current correction of level l (exact colour sweeps start from zero)
  static constexpr int kThreads = kBlock;
  int l;
  __device__ void operator()(const CellDev& C, const Ent& e, int lb, int nb) const {
    fill_body(ucur(C, e.par, l), 3 * C.g[l].nn, 0.f, lb, nb);
  }
};

struct OpFineResid {  // one LDS tile per block when the cell tiles
  static constexpr int kThreads = kBlock;
  __device__ void operator()(const CellDev& C, const Ent& e, int lb, int) const {
    if (C.tile_bx == 32)
      fine_residual_tile_body<32, kBlock / 32>(C.g[0], C.fc, C.kappa, C.f[0], ucur(C, e.par, 0), C.res[0], lb);
    else if (C.tile_bx == 16)
      fine_residual_tile_body<16, kBlock / 16>(C.g[0], C.fc, C.kappa, C.f[0], ucur(C, e.par, 0), C.res[0], lb);
    else
      fine_residual_body(C.g[0], C.fc, C.kappa, C.f[0], ucur(C, e.par, 0), C.res[0],
                         (long long)lb * kBlock + threadIdx.x);
  }
};

struct OpStResid {
  static constexpr int kThreads = kBlock;
  int l;
  __device__ void operator()(const CellDev& C, const Ent& e, int lb, int) const {
    const long long t = (long long)lb * kBlock + threadIdx.x;
    if (C.st[l].f16)
      st_residual_body<true>(C.g[l], C.st[l], C.f[l], ucur(C, e.par, l), C.res[l], t);
    else if (st_resid8(C.g[l], false))  // 8 lanes per node, as the single-cell launch (8 x the blocks)
      st_residual8_body(C.g[l], C.st[l].f32, C.f[l], ucur(C, e.par, l), C.res[l], t);
    else
      st_residual_body<false>(C.g[l], C.st[l], C.f[l], ucur(C, e.par, l), C.res[l], t);
  }
};

struct OpRestrict {
  static constexpr int kThreads = kBlock;
  int l;
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int) const {
    const long long t = (long long)lb * kBlock + threadIdx.x;
    if (restrict8(C.g[l + 1]))  // as the single-cell launch (8 x the blocks)
      restrict8_body(C.g[l], C.g[l + 1], C.res[l], C.f[l + 1], t);
    else
      restrict_body(C.g[l], C.g[l + 1], C.res[l], C.f[l + 1], t);
  }
};

struct OpProlong {
  static constexpr int kThreads = kBlock;
  int l;
  __device__ void operator()(const CellDev& C, const Ent& e, int lb, int) const {
    prolong_body(C.g[l], C.g[l + 1], ucur(C, e.par, l + 1), ucur(C, e.par, l), (long long)lb * kBlock + threadIdx.x);
  }
};

struct OpTail {  // one workgroup per cell: the tail V-cycle from level C.tail down
  static constexpr int kThreads = kTailThreads;
  __device__ void operator()(const CellDev& C, const Ent& e, int, int) const {
    __shared__ TailLevels T;
    if (threadIdx.x == 0) {
      T.n = C.nlev - C.tail;
      T.sweeps = C.sweeps_coarse;
      for (int i = 0; i < T.n; i++) {
        const int l = C.tail + i;
        T.g[i] = C.g[l];
        T.f[i] = C.f[l];
        T.res[i] = C.res[l];
        T.st[i] = C.st[l];
      }
      T.f0 = C.f[C.tail];
      T.Binv = C.Binv;
      T.uout = ucur(C, e.par, C.tail);
    }
    __syncthreads();
    tail_vcycle_body(T);
  }
};

template <typename XT>
struct OpInitResid {
  static constexpr int kThreads = kBlock;
  double rtol2;  // < 0: the cell's own
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int nb) const {
    Reduce R = cell_reduce(C, FIN_BB_RR);
    if (rtol2 >= 0.0) R.rtol2 = rtol2;
    const XT* x;
    if constexpr (sizeof(XT) == 8) x = C.x64;
    else x = C.x;
    const float* xlo = sizeof(XT) == 4 ? C.xlo : nullptr;
    // the single-cell launch's form for this row length, one load case after
    // another: the same r and partial sums
    if (!fine_rows_ok(C.g[0].nx)) {
      fine_init_residual_body<XT>(C.g[0], C.fc, C.kappa, x, xlo, C.r, R, lb, nb);
      return;
    }
    for (int k = 0; k < OTM_NRHS; k++)
      fine_init_residual_rows_body<XT>(C.g[0], C.fc, C.kappa, x, xlo, C.r, R, lb, nb, k);
  }
};

struct OpFinal {  // one block per cell
  static constexpr int kThreads = kBlock;
  int mode, nv;
  double rtol2;   // < 0: the cell's own
  int tiled = 0;  // partials of the tiled operator apply: its own count
  __device__ void operator()(const CellDev& C, const Ent&, int, int) const {
    Reduce R = cell_reduce(C, mode);
    if (rtol2 >= 0.0) R.rtol2 = rtol2;
    const unsigned nb = (tiled && C.tile_bx) ? fine_apply_tile_blocks(C.g[0], C.tile_bx)
                                              : red_blocks_for(C.g[0].nn);
    reduce_final_body(R, nv, (int)nb);
  }
};

struct OpApplyDot {
  static constexpr int kThreads = kBlock;
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int nb) const {
    if (C.tile_bx == 32)
      fine_apply_tile_body<32, kBlock / 32>(C.g[0], C.fc, C.kappa, C.p, C.q, cell_reduce(C, FIN_PQ), lb, nb);
    else if (C.tile_bx == 16)
      fine_apply_tile_body<16, kBlock / 16>(C.g[0], C.fc, C.kappa, C.p, C.q, cell_reduce(C, FIN_PQ), lb, nb);
    else
      fine_apply_dot_body(C.g[0], C.fc, C.kappa, C.p, C.q, cell_reduce(C, FIN_PQ), lb, nb);
  }
};

struct OpXr {
  static constexpr int kThreads = kBlock;
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int nb) const {
    pcg_xr_body(C.g[0].nn, C.x, C.xlo, C.r, C.p, C.q, C.pcg, cell_reduce(C, FIN_RR), lb, nb);
  }
};

struct OpDot3 {
  static constexpr int kThreads = kBlock;
  __device__ void operator()(const CellDev& C, const Ent& e, int lb, int nb) const {
    dot3_body(C.g[0].nn, C.r, ucur(C, e.par, 0), cell_reduce(C, FIN_RZ), lb, nb);
  }
};

struct OpPcgP {
  static constexpr int kThreads = kBlock;
  __device__ void operator()(const CellDev& C, const Ent& e, int lb, int nb) const {
    pcg_p_body(C.g[0].nn, C.p, ucur(C, e.par, 0), C.pcg, lb, nb);
  }
};

struct OpFillVec {  // which: 0 = p, 1 = x
  static constexpr int kThreads = kBlock;
  int which;
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int nb) const {
    fill_body(which ? C.x : C.p, 3 * C.g[0].nn, 0.f, lb, nb);
  }
};

struct OpSubMeanR {
  static constexpr int kThreads = kBlock;
  int which;  // 0 = r, 1 = x
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int nb) const {
    sub_mean3_body(C.g[0].nn, which ? C.x : C.r, C.pcg, lb, nb);
  }
};

struct OpRefineAdd {
  static constexpr int kThreads = kBlock;
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int nb) const {
    refine_add_body(3 * C.g[0].nn, C.x64, C.x, lb, nb);
  }
};

template <typename XT>
struct OpSum3 {
  static constexpr int kThreads = kBlock;
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int nb) const {
    Reduce R = cell_reduce(C, FIN_MEAN);
    const XT* x;
    if constexpr (sizeof(XT) == 8) x = C.x64;
    else x = C.x;
    sum3_body<XT>(C.g[0].nn, x, R, lb, nb);
  }
};

struct OpSubMeanCast {
  static constexpr int kThreads = kBlock;
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int nb) const {
    sub_mean_cast_body(C.g[0].nn, C.x64, C.x, C.pcg, lb, nb);
  }
};

struct OpKappaH {
  static constexpr int kThreads = kBlock;
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int nb) const {
    fine_kappa_h_body(C.g[0], C.fc, C.kappa, ChiRef{C.x, C.xlo, C.x64}, cell_reduce(C, FIN_STORE), lb, nb);
  }
};

struct OpSensitivity {
  static constexpr int kThreads = kBlock;
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int) const {
    fine_sensitivity_body(C.g[0], C.fc, ChiRef{C.x, C.xlo, C.x64}, C.sw, C.sens, C.sens_pitch, (long long)lb * kBlock + threadIdx.x);
  }
};

// ---- setup operations -------------------------------------------------------
struct OpCopyKappa {  // kappa_eff = rhoPenalMin + kappa (grid.h:429, grid.cu:405)
  static constexpr int kThreads = kBlock;
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int) const {
    const long long t = (long long)lb * kBlock + threadIdx.x;
    const Geom& g = C.g[0];
    if (t >= g.nn) return;
    const long long x = t % g.nx, r = t / g.nx;
    C.kappa_w[t] = C.ksrc[x + (long long)C.kpitch * r] + 1e-9f;
  }
};

struct OpGalerkin {  // level l's element matrices and stencil
  static constexpr int kThreads = kBlock;
  int l, what;  // what 0: element matrices, 1: node stencil
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int) const {
    const long long t = (long long)lb * kBlock + threadIdx.x;
    if (what == 1)
      stencil_from_elements_body(C.g[l], C.E[l], C.st_w[l], t);
    else if (l == 1)
      galerkin_first_body(C.g[0], C.g[1], C.G, C.kappa, C.E[1], t);
    else
      galerkin_next_body(C.g[l - 1], C.g[l], C.E[l - 1], C.E[l], t);
  }
};

struct OpPackFp16 {  // what 0: clear scale, 1: |coupling| max, 2: pack
  static constexpr int kThreads = kBlock;
  int l, what;
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int nb) const {
    if (what == 0) {
      if (lb == 0 && threadIdx.x == 0) *C.scale_w[l] = 0.f;
    } else if (what == 1) {
      stencil_absmax_body(C.g[l], C.st_w[l], reinterpret_cast<unsigned int*>(C.scale_w[l]), lb, nb);
    } else {
      stencil_pack_body(C.g[l], C.st_w[l], C.scale_w[l], C.st16_w[l], (long long)lb * kBlock + threadIdx.x);
    }
  }
};

struct OpCoarsestInvert {  // one workgroup per cell, work matrix in dynamic LDS (n <= kInvertLdsMax)
  static constexpr int kThreads = kFactorThreads;
  __device__ void operator()(const CellDev& C, const Ent&, int, int) const {
    extern __shared__ double A[];
    coarsest_invert_lds_body(C.g[C.nlev - 1], C.st_w[C.nlev - 1], A, C.Binv_w);
  }
};

// Per-cell scalars read back once per step: PCG scalars and kappa^H sums.
struct CellReadback {
  PcgScalars pcg;
  double out[6];
};

struct OpGather {  // one block per cell
  static constexpr int kThreads = 64;
  CellReadback* dst;
  __device__ void operator()(const CellDev& C, const Ent& e, int, int) const {
    if (threadIdx.x == 0) {
      dst[e.aux].pcg = *C.pcg;
      for (int i = 0; i < 6; i++) dst[e.aux].out[i] = C.red_out[i];
    }
  }
};

}  // namespace otm

using namespace otm;

namespace {

constexpr double kRefineInnerB = 1e-4;  // inner reduction per refinement step (as the single-cell solve)

inline unsigned nblk(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

struct Member {
  otm_ctx* c = nullptr;
  int slot = -1;   // index in the device CellDev array
  int out = -1;    // caller's index
  int par = 0;     // buffer parity bits
  int nlev = 0;
  enum Phase { RESTART, ITER, FINISH, DONE, FALLBACK, FAILED } phase = RESTART;
  int restart = 0, it = 0, its[3] = {0, 0, 0};
  double prev_worst = 1e300;
  bool pending = false;  // refinement correction not yet added
  bool refine = false;
  bool want_sens = false;
  int fail_code = 0;
  std::string fail_msg;
  double C[9];
  double relres[3];
};

}  // namespace

struct otm_batch {
  int device = 0;
  hipStream_t s = nullptr;
  CellDev* dcells = nullptr;
  int dcap = 0;
  std::vector<CellDev> hcells;
  Ent* dents[2] = {nullptr, nullptr};
  Ent* hents[2] = {nullptr, nullptr};  // pinned staging
  size_t ent_cap = 0;
  CellReadback* dread = nullptr;
  CellReadback* hread = nullptr;  // pinned
  int read_cap = 0;
  std::vector<Member> members;
  int flip = 0;  // which staging half the next plan uses
  // one planned launch
  struct Launch {
    std::function<void(const Ent*)> go;
    size_t off;
  };
  std::vector<Launch> plan;
  std::vector<Ent> plan_ents;
  long long launches = 0;
  int steps = 0;
};

namespace {

void ensure_cells(otm_batch* b, int n) {
  if (n <= b->dcap) return;
  if (b->s) hip_check(hipStreamSynchronize(b->s), "hipStreamSynchronize");  // no launch still reads them
  int cap = std::max(n, 2 * b->dcap);
  CellDev* d = dalloc<CellDev>(cap);
  dfree(b->dcells);
  b->dcells = d;
  b->dcap = cap;
  CellReadback* r = dalloc<CellReadback>(cap);
  dfree(b->dread);
  b->dread = r;
  if (b->hread) (void)hipHostFree(b->hread);
  b->hread = nullptr;
  hip_check(hipHostMalloc((void**)&b->hread, sizeof(CellReadback) * cap, hipHostMallocDefault), "hipHostMalloc");
  b->read_cap = cap;
}

void ensure_ents(otm_batch* b, size_t n) {
  if (n <= b->ent_cap) return;
  if (b->s) hip_check(hipStreamSynchronize(b->s), "hipStreamSynchronize");  // no launch or copy still uses them
  size_t cap = std::max(n, 2 * b->ent_cap);
  for (int h = 0; h < 2; h++) {
    dfree(b->dents[h]);
    if (b->hents[h]) (void)hipHostFree(b->hents[h]);
    b->dents[h] = dalloc<Ent>(cap);
    b->hents[h] = nullptr;
    hip_check(hipHostMalloc((void**)&b->hents[h], sizeof(Ent) * cap, hipHostMallocDefault), "hipHostMalloc");
  }
  b->ent_cap = cap;
}

CellDev describe(const Member& m) {
  const otm_ctx* c = m.c;
  CellDev d;
  memset(&d, 0, sizeof(d));
  d.nlev = (int)c->lv.size();
  d.tail = c->tail;
  for (int l = 0; l < d.nlev; l++) {
    const Level& L = c->lv[l];
    d.g[l] = L.g;
    d.f[l] = L.f;
    d.res[l] = L.res;
    d.st[l] = stref(L);
    d.brick[l] = L.brick;
    d.ua[l] = L.u;
    d.ub[l] = L.u2 ? L.u2 : L.u;
  }
  // the batched launches run the default fine tile shape
  if (c->lv[0].brick && c->fine_tile != fine_tile_default(c->g0))
    fail(OTM_E_UNSUPPORTED, "the batched form runs the default fine tile shape (otm_config.fine_tile = -1)");
  d.fine_bx = c->lv[0].brick ? fine_brick_x(c->g0) : 0;
  d.tile_bx = fine_brick_x(c->g0);
  d.sweeps_fine = c->cfg.pre_sweeps;
  d.sweeps_coarse = c->cfg.coarse_sweeps > 0 ? c->cfg.coarse_sweeps : c->cfg.pre_sweeps;
  d.fc = c->fc;
  d.kappa = c->kappa;
  d.x = c->x;
  d.xlo = c->xlo;
  d.r = c->r;
  d.p = c->p;
  d.q = c->q;
  d.x64 = c->x64;
  d.Binv = c->cBinv;
  d.partials = c->partials;
  d.red_out = c->red_out;
  d.pcg = c->pcg;
  d.rtol2 = c->cfg.fem_rel_thres * c->cfg.fem_rel_thres;
  d.inv_nn = 1.0 / (double)c->g0.nn;
  d.inner2 = c->x64 ? kRefineInnerB * kRefineInnerB : 0.0;
  d.kappa_w = c->kappa;
  d.G = c->G;
  for (int l = 0; l < d.nlev; l++) {
    const Level& L = c->lv[l];
    d.E[l] = L.E;
    d.st_w[l] = L.st;
    d.st16_w[l] = L.st16;
    d.scale_w[l] = L.scale;
  }
  d.cB = c->cB;
  d.Binv_w = c->cBinv;
  return d;
}

// ---- planning -----------------------------------------------------------------
// A launch over a set of members: per member a block count; the entries go
// to the staging buffer and the launch is issued after one upload.
template <class F>
void add_launch(otm_batch* b, const std::vector<Member*>& set, F blocks_of, int aux_from_slot,
                std::function<void(const Ent*, int, unsigned)> issue) {
  if (set.empty()) return;
  const size_t off = b->plan_ents.size();
  int block0 = 0;
  for (Member* m : set) {
    const int nb = (int)blocks_of(*m);
    if (nb <= 0) continue;
    b->plan_ents.push_back(Ent{m->slot, block0, m->par, aux_from_slot ? m->slot : 0});
    block0 += nb;
  }
  const int nents = (int)(b->plan_ents.size() - off);
  if (nents == 0) return;
  const unsigned grid = (unsigned)block0;
  b->plan.push_back({[issue, nents, grid](const Ent* base) { issue(base, nents, grid); }, off});
}

template <class Op>
void add_op(otm_batch* b, const std::vector<Member*>& set, Op op, std::function<long long(const Member&)> blocks) {
  CellDev* cells = b->dcells;
  hipStream_t s = b->s;
  add_launch(b, set, blocks, 1, [cells, s, op](const Ent* e, int n, unsigned grid) {
    hipLaunchKernelGGL(kb_run<Op>, dim3(grid), dim3(Op::kThreads), 0, s, e, n, cells, op);
  });
}

template <class Op>
void add_op_lds(otm_batch* b, const std::vector<Member*>& set, Op op, std::function<long long(const Member&)> blocks,
                size_t lds) {
  CellDev* cells = b->dcells;
  hipStream_t s = b->s;
  add_launch(b, set, blocks, 1, [cells, s, op, lds](const Ent* e, int n, unsigned grid) {
    hipLaunchKernelGGL(kb_run<Op>, dim3(grid), dim3(Op::kThreads), lds, s, e, n, cells, op);
  });
}

void flush(otm_batch* b) {
  if (b->plan.empty()) {
    b->plan_ents.clear();
    return;
  }
  ensure_ents(b, b->plan_ents.size());
  const int h = b->flip;
  b->flip ^= 1;
  memcpy(b->hents[h], b->plan_ents.data(), sizeof(Ent) * b->plan_ents.size());
  hip_check(hipMemcpyAsync(b->dents[h], b->hents[h], sizeof(Ent) * b->plan_ents.size(), hipMemcpyHostToDevice, b->s),
            "upload launch tables");
  for (auto& L : b->plan) L.go(b->dents[h] + L.off);
  b->launches += (long long)b->plan.size();
  hip_check(hipGetLastError(), "batched launch");
  b->plan.clear();
  b->plan_ents.clear();
}

long long nodes0(const Member& m) { return m.c->g0.nn; }
long long redb(const Member& m) { return red_blocks_for(m.c->g0.nn); }
// blocks of the operator apply and of the fine residual (LDS tiles when the
// cell tiles)
long long applyb(const Member& m) {
  const int bx = fine_brick_x(m.c->g0);
  return bx ? (long long)fine_apply_tile_blocks(m.c->g0, bx) : (long long)red_blocks_for(m.c->g0.nn);
}
long long residb(const Member& m) {
  const int bx = fine_brick_x(m.c->g0);
  return bx ? (long long)fine_tiles(m.c->g0, bx) : (long long)nblk(m.c->g0.nn);
}
long long one(const Member&) { return 1; }

// V-cycle for the member set V (all levels, every member at its own depth).
void plan_vcycle(otm_batch* b, std::vector<Member*>& V) {
  // levels above each member's tail run as separate launches (all members
  // per launch); every member's tail V-cycle runs in one launch at the bottom
  int maxlev = 0;
  for (Member* m : V) maxlev = std::max(maxlev, m->c->tail + 1);
  auto at = [&](int l, bool) {
    std::vector<Member*> s;
    for (Member* m : V)
      if (l < m->c->tail) s.push_back(m);
    return s;
  };
  // down
  for (int l = 0; l + 1 < maxlev; l++) {
    std::vector<Member*> S = at(l, false);
    if (S.empty()) continue;
    if (l == 0) {
      std::vector<Member*> b32, b16, col;
      for (Member* m : S) {
        const int bx = m->c->lv[0].brick ? fine_brick_x(m->c->g0) : 0;
        (bx == 32 ? b32 : bx == 16 ? b16 : col).push_back(m);
      }
      CellDev* cells = b->dcells;
      hipStream_t s = b->s;
      auto bricks = [](const Member& m) {
        return (long long)(m.c->g0.nn / (fine_brick_x(m.c->g0) * kFineBY * kFineBZ));
      };
      add_launch(b, b32, bricks, 0, [cells, s](const Ent* e, int n, unsigned grid) {
        hipLaunchKernelGGL(kb_fine_brick<32>, dim3(grid), dim3(32 * kFineTR), 0, s, e, n, cells, 0);
      });
      add_launch(b, b16, bricks, 0, [cells, s](const Ent* e, int n, unsigned grid) {
        hipLaunchKernelGGL(kb_fine_brick<16>, dim3(grid), dim3(16 * kFineTR), 0, s, e, n, cells, 0);
      });
      if (!col.empty()) {
        add_op(b, col, OpFillU{0}, [](const Member& m) { return (long long)nblk(3 * m.c->g0.nn); });
        for (int sw = 0; sw < col.front()->c->cfg.pre_sweeps; sw++)
          for (int cc = 0; cc < 4; cc++)
            add_op(b, col, OpFineGsColor{cc & 1, cc >> 1},
                   [](const Member& m) { return (long long)nblk(m.c->g0.nx * (m.c->g0.ny / 2) * (m.c->g0.nz / 2)); });
      }
      add_op(b, S, OpFineResid{}, residb);
    } else {
      add_op(b, S, OpStBrick{l, 0}, [l](const Member& m) { return (long long)brick_count(m.c->lv[l].g); });
      add_op(b, S, OpStResid{l}, [l](const Member& m) {
        const Level& L = m.c->lv[l];
        return (long long)nblk(st_resid8(L.g, L.st16 != nullptr) ? 8 * L.g.nn : L.g.nn);
      });
    }
    add_op(b, S, OpRestrict{l}, [l](const Member& m) {
      const Geom& cg = m.c->lv[l + 1].g;
      return (long long)nblk(restrict8(cg) ? 8 * cg.nn : cg.nn);
    });
  }
  // bottom: every member's tail
  add_op(b, V, OpTail{}, one);
  // up
  for (int l = maxlev - 2; l >= 0; l--) {
    std::vector<Member*> S = at(l, false);
    if (S.empty()) continue;
    add_op(b, S, OpProlong{l}, [l](const Member& m) { return (long long)nblk(m.c->lv[l + 1].g.nn); });
    if (l == 0) {
      std::vector<Member*> b32, b16, col;
      for (Member* m : S) {
        const int bx = m->c->lv[0].brick ? fine_brick_x(m->c->g0) : 0;
        (bx == 32 ? b32 : bx == 16 ? b16 : col).push_back(m);
      }
      CellDev* cells = b->dcells;
      hipStream_t s = b->s;
      auto bricks = [](const Member& m) {
        return (long long)(m.c->g0.nn / (fine_brick_x(m.c->g0) * kFineBY * kFineBZ));
      };
      add_launch(b, b32, bricks, 0, [cells, s](const Ent* e, int n, unsigned grid) {
        hipLaunchKernelGGL(kb_fine_brick<32>, dim3(grid), dim3(32 * kFineTR), 0, s, e, n, cells, 1);
      });
      add_launch(b, b16, bricks, 0, [cells, s](const Ent* e, int n, unsigned grid) {
        hipLaunchKernelGGL(kb_fine_brick<16>, dim3(grid), dim3(16 * kFineTR), 0, s, e, n, cells, 1);
      });
      for (Member* m : b32) m->par ^= 1;
      for (Member* m : b16) m->par ^= 1;
      if (!col.empty())
        for (int sw = 0; sw < col.front()->c->cfg.post_sweeps; sw++)
          for (int cc = 3; cc >= 0; cc--)
            add_op(b, col, OpFineGsColor{cc & 1, cc >> 1},
                   [](const Member& m) { return (long long)nblk(m.c->g0.nx * (m.c->g0.ny / 2) * (m.c->g0.nz / 2)); });
    } else {
      add_op(b, S, OpStBrick{l, 1}, [l](const Member& m) { return (long long)brick_count(m.c->lv[l].g); });
      for (Member* m : S) m->par ^= (1 << l);
    }
  }
}

void read_back(otm_batch* b, const std::vector<Member*>& set) {
  if (set.empty()) return;
  add_op(b, set, OpGather{b->dread}, one);
  flush(b);
  int maxslot = 0;
  for (Member* m : set) maxslot = std::max(maxslot, m->slot);
  hip_check(hipMemcpyAsync(b->hread, b->dread, sizeof(CellReadback) * (maxslot + 1), hipMemcpyDeviceToHost, b->s),
            "read cell scalars");
  hip_check(hipStreamSynchronize(b->s), "hipStreamSynchronize");
}

bool any_active(const PcgScalars& P) { return P.active[0] || P.active[1] || P.active[2]; }

// Applies the batch's buffer parity to the context (its pointers then name
// the buffers that hold the current corrections, as after single-cell
// solves).
void write_back_parity(Member& m) {
  otm_ctx* c = m.c;
  for (size_t l = 0; l < c->lv.size(); l++)
    if ((m.par >> l) & 1) {
      Level& L = c->lv[l];
      if (!L.u2) continue;
      std::swap(L.u, L.u2);
      if (l == 0) std::swap(c->z, c->z2);
    }
  m.par = 0;
}

// One step of the batch state machine (see the file comment).
void step(otm_batch* b) {
  std::vector<Member*> A, B;
  for (Member& m : b->members) {
    if (m.phase == Member::RESTART) A.push_back(&m);
    else if (m.phase == Member::ITER) B.push_back(&m);
  }
  // ---- phase 1: restarts compute the FP64 residual; the others apply K
  std::vector<Member*> Aadd, Afill, A32, A64;
  for (Member* m : A) {
    if (m->refine) (m->pending ? Aadd : Afill).push_back(m);
    (m->refine ? A64 : A32).push_back(m);
  }
  add_op(b, Aadd, OpRefineAdd{}, [](const Member& m) { return (long long)std::min<long long>(8192, nblk(3 * m.c->g0.nn)); });
  add_op(b, Afill, OpFillVec{1}, [](const Member& m) { return (long long)std::min<long long>(8192, nblk(3 * m.c->g0.nn)); });
  for (Member* m : Aadd) m->pending = false;
  add_op(b, A32, OpInitResid<float>{-1.0}, redb);
  add_op(b, A64, OpInitResid<double>{-1.0}, redb);
  add_op(b, A, OpFinal{FIN_BB_RR, 3 * OTM_NRHS, -1.0}, one);
  add_op(b, B, OpApplyDot{}, applyb);
  add_op(b, B, OpFinal{FIN_PQ, OTM_NRHS, -1.0, 1}, one);
  add_op(b, B, OpXr{}, redb);
  add_op(b, B, OpFinal{FIN_RR, OTM_NRHS, -1.0}, one);
  std::vector<Member*> AB(A);
  AB.insert(AB.end(), B.begin(), B.end());
  read_back(b, AB);
  // ---- decisions
  std::vector<Member*> A2, B2;
  for (Member* m : A) {
    const PcgScalars& P = b->hread[m->slot].pcg;
    const int max_restarts = m->refine ? 8 : 3;
    if (P.fault & 1) {
      m->phase = Member::FAILED;
      m->fail_code = OTM_E_NUMERIC;
      m->fail_msg = "non-finite residual (MG numerical explode)";
      continue;
    }
    double worst = 0.0;
    for (int k = 0; k < 3; k++) worst = std::max(worst, P.rr[k] / P.bb[k]);
    if (!any_active(P) || m->it >= m->c->cfg.max_cg_iter || (m->restart > 0 && worst > 0.25 * m->prev_worst) ||
        m->restart >= max_restarts) {
      // the true residual was just evaluated at the final iterate
      for (int k = 0; k < 3; k++) m->relres[k] = sqrt(P.rr[k] / P.bb[k]);
      m->phase = Member::FINISH;
      continue;
    }
    m->prev_worst = worst;
    m->restart++;
    m->pending = m->refine;
    A2.push_back(m);
  }
  for (Member* m : B) {
    const PcgScalars& P = b->hread[m->slot].pcg;
    m->it++;
    for (int k = 0; k < 3; k++)
      if (P.active[k]) m->its[k]++;
    if (P.fault & 1) {
      m->phase = Member::FAILED;
      m->fail_code = OTM_E_NUMERIC;
      m->fail_msg = "non-finite residual (MG numerical explode)";
      continue;
    }
    if (P.fault & 2) {
      m->phase = Member::FALLBACK;
      continue;
    }
    if (!any_active(P) || m->it >= m->c->cfg.max_cg_iter) {
      m->phase = Member::RESTART;
      continue;
    }
    B2.push_back(m);
  }
  // ---- phase 2: V-cycle and direction update for every active cell
  std::vector<Member*> V(A2);
  V.insert(V.end(), B2.begin(), B2.end());
  if (!V.empty()) {
    add_op(b, A2, OpSubMeanR{0}, [](const Member& m) { return (long long)std::min<long long>(8192, nblk(m.c->g0.nn)); });
    plan_vcycle(b, V);
    add_op(b, V, OpDot3{}, redb);
    add_op(b, A2, OpFinal{FIN_RZ_FIRST, OTM_NRHS, -1.0}, one);
    add_op(b, B2, OpFinal{FIN_RZ, OTM_NRHS, -1.0}, one);
    add_op(b, A2, OpFillVec{0}, [](const Member& m) { return (long long)std::min<long long>(8192, nblk(3 * m.c->g0.nn)); });
    add_op(b, V, OpPcgP{}, redb);
    flush(b);
    for (Member* m : V) m->phase = Member::ITER;
  }
  b->steps++;
}

// Mean removal and kappa^H for the members that finished. A member finishes
// right after its restart phase evaluated the true residual at its final
// iterate (pending refinement corrections are added before that
// evaluation), so its relative residuals are already recorded.
void finish(otm_batch* b) {
  std::vector<Member*> D, D32, D64;
  for (Member& m : b->members)
    if (m.phase == Member::FINISH) {
      D.push_back(&m);
      (m.refine ? D64 : D32).push_back(&m);
    }
  if (D.empty()) return;
  add_op(b, D32, OpSum3<float>{}, redb);
  add_op(b, D64, OpSum3<double>{}, redb);
  add_op(b, D, OpFinal{FIN_MEAN, OTM_NRHS, -1.0}, one);
  add_op(b, D32, OpSubMeanR{1}, [](const Member& m) { return (long long)std::min<long long>(8192, nblk(m.c->g0.nn)); });
  add_op(b, D64, OpSubMeanCast{}, [](const Member& m) { return (long long)std::min<long long>(8192, nblk(m.c->g0.nn)); });
  add_op(b, D, OpKappaH{}, redb);
  add_op(b, D, OpFinal{FIN_STORE, 6, -1.0}, one);
  read_back(b, D);
  for (Member* m : D) {
    if (m->phase == Member::FAILED) continue;
    const double* s = b->hread[m->slot].out;
    double* C = m->C;
    C[0] = s[0]; C[1] = s[1]; C[2] = s[2];
    C[3] = s[1]; C[4] = s[3]; C[5] = s[4];
    C[6] = s[2]; C[7] = s[4]; C[8] = s[5];
    m->phase = Member::DONE;
  }
}

// Plans the setup of the contexts in `tmp` (slots already assigned):
// conductivity copy from src[i] (row pitch `pitch`, <= 0: reso[0]), every
// Galerkin level, FP16 packing and the coarsest inverse, all cells per
// launch; `cold` also zeroes the iterates. The descriptors are uploaded to
// the members' slots. Leaves b->members alone, so it can run between the
// steps of a live batch.
struct OpZeroX {
  static constexpr int kThreads = kBlock;
  __device__ void operator()(const CellDev& C, const Ent&, int lb, int nb) const {
    fill_body(C.x, 3 * C.g[0].nn, 0.f, lb, nb);
    if (C.xlo) fill_body(C.xlo, 3 * C.g[0].nn, 0.f, lb, nb);
    if (C.x64) {
      const BlockRange br = block_range_v(3 * C.g[0].nn, lb, nb);
      for (long long i = br.b + threadIdx.x; i < br.e; i += blockDim.x) C.x64[i] = 0.0;
    }
  }
};

void plan_setup(otm_batch* b, std::vector<Member>& tmp, const std::vector<const float*>& src, int pitch, bool cold) {
  std::vector<Member*> all;
  int maxlev = 0;
  size_t lds_max = 0;
  std::vector<Member*> inv_lds, inv_glob;
  for (size_t i = 0; i < tmp.size(); i++) {
    Member& m = tmp[i];
    otm_ctx* c = m.c;
    CellDev d = describe(m);
    d.ksrc = src[i];
    d.kpitch = pitch > 0 ? pitch : c->g0.nx;
    if ((int)b->hcells.size() <= m.slot) b->hcells.resize(m.slot + 1);
    b->hcells[m.slot] = d;
    hip_check(hipMemcpyAsync(b->dcells + m.slot, &b->hcells[m.slot], sizeof(CellDev), hipMemcpyHostToDevice, b->s),
              "upload cell");
    all.push_back(&m);
    maxlev = std::max(maxlev, m.nlev);
    const long long nc = c->lv.back().g.nn;
    if (nc <= kInvertLdsMax) {
      inv_lds.push_back(&m);
      lds_max = std::max(lds_max, (size_t)(nc * nc) * sizeof(double));
    } else {
      inv_glob.push_back(&m);
    }
  }
  add_op(b, all, OpCopyKappa{}, [](const Member& m) { return (long long)nblk(m.c->g0.nn); });
  if (cold) add_op(b, all, OpZeroX{}, [](const Member& m) { return std::min<long long>(8192, nblk(3 * m.c->g0.nn)); });
  for (int l = 1; l < maxlev; l++) {
    std::vector<Member*> S, P;
    for (Member* m : all)
      if (l < m->nlev) {
        S.push_back(m);
        if (m->c->lv[l].st16) P.push_back(m);
      }
    add_op(b, S, OpGalerkin{l, 0}, [l](const Member& m) { return (long long)nblk(m.c->lv[l].g.nn); });
    add_op(b, S, OpGalerkin{l, 1}, [l](const Member& m) { return (long long)nblk(m.c->lv[l].g.nn); });
    add_op(b, P, OpPackFp16{l, 0}, one);
    add_op(b, P, OpPackFp16{l, 1},
           [l](const Member& m) { return std::min<long long>(1024, (long long)nblk(m.c->lv[l].g.nn)); });
    add_op(b, P, OpPackFp16{l, 2}, [l](const Member& m) { return (long long)nblk(m.c->lv[l].g.nn); });
  }
  if (!inv_lds.empty())
    hip_check(hipFuncSetAttribute((const void*)kb_run<OpCoarsestInvert>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(kInvertLdsMax * kInvertLdsMax * sizeof(double))),
              "hipFuncSetAttribute");
  add_op_lds(b, inv_lds, OpCoarsestInvert{}, one, lds_max);
  // larger coarsest levels: the blocked multi-workgroup inverse, per cell
  for (Member* m : inv_glob) {
    const otm_ctx* c = m->c;
    const Geom g = c->lv.back().g;
    const float* st = c->lv.back().st;
    double* B = c->cB;
    double* Binv = c->cBinv;
    hipStream_t s = b->s;
    b->plan.push_back({[s, g, st, B, Binv](const Ent*) { launch_coarsest_blocked(s, g, st, B, Binv); }, 0});
  }
}

// Adds a context as a member in slot `slot` (its CellDev is uploaded at
// once); the caller keeps slots unique and within the capacity.
Member& add_member(otm_batch* b, otm_ctx* c, int slot, int out) {
  Member m;
  m.c = c;
  c->energy_ok = false;  // the batched solve changes its iterate
  m.slot = slot;
  m.out = out;
  m.nlev = (int)c->lv.size();
  m.refine = c->x64 != nullptr;
  // a context already on exact sweeps (or without a brick smoother on some
  // coarse level) is solved alone
  bool bricks = c->bricks_on;
  for (size_t l = 1; l + 1 < c->lv.size(); l++) bricks = bricks && c->lv[l].brick;
  if (!bricks) m.phase = Member::FALLBACK;
  if ((int)b->hcells.size() <= slot) b->hcells.resize(slot + 1);
  CellDev d = describe(m);
  d.ksrc = b->hcells[slot].ksrc;  // keep what a pending setup launch reads
  d.kpitch = b->hcells[slot].kpitch;
  b->hcells[slot] = d;
  hip_check(hipMemcpyAsync(b->dcells + slot, &b->hcells[slot], sizeof(CellDev), hipMemcpyHostToDevice, b->s),
            "upload cell");
  hip_check(hipMemsetAsync(&c->pcg->fault, 0, sizeof(int), b->s), "clear fault");
  b->members.push_back(m);
  return b->members.back();
}

bool live(const Member& m) { return m.phase == Member::RESTART || m.phase == Member::ITER; }

// Settles a member that left the state machine: buffer parity back into
// the context, the exact-sweep solve for a fallback, solve statistics into
// the context. Returns the member's status.
int conclude(Member& m) {
  write_back_parity(m);
  otm_ctx* c = m.c;
  if (m.phase == Member::FALLBACK) {
    c->fallbacks++;
    try {
      (void)ctx_solve_exact(c, m.C);  // the target is checked below, as for every member
      for (int k = 0; k < 3; k++) m.relres[k] = c->relres[k];
      m.phase = Member::DONE;
    } catch (const OtmError& e) {
      m.phase = Member::FAILED;
      m.fail_code = e.code;
      m.fail_msg = e.msg;
    }
  } else if (m.phase == Member::DONE) {
    for (int k = 0; k < 3; k++) {
      c->relres[k] = m.relres[k];
      c->iters[k] = m.its[k] + (m.its[k] < m.it ? 1 : 0);
    }
    memcpy(c->C, m.C, sizeof(c->C));
    c->has_solution = true;
  }
  if (m.phase == Member::DONE) {
    // C and the statistics are published either way; a solve that ended
    // above its target (iteration cap, stalled residual) says so
    for (int k = 0; k < 3; k++)
      if (!(m.relres[k] <= c->cfg.fem_rel_thres)) {
        char buf[160];
        snprintf(buf, sizeof(buf), "solve ended above the relative residual target %.3g: (%.3g, %.3g, %.3g)",
                 c->cfg.fem_rel_thres, m.relres[0], m.relres[1], m.relres[2]);
        m.fail_msg = buf;
        return OTM_E_NOCONV;
      }
    return OTM_OK;
  }
  return m.fail_code ? m.fail_code : OTM_E_NUMERIC;
}

}  // namespace

extern "C" {

OTM_API int otm_batch_create(int device, otm_batch** out) {
  return guarded([&] {
    if (!out) fail(OTM_E_ARG, "null argument");
    *out = nullptr;
    otm_batch* b = new otm_batch();
    b->device = device;
    try {
      hip_check(hipSetDevice(device), "hipSetDevice");
      hip_check(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking), "hipStreamCreate");
      ensure_cells(b, 16);
      ensure_ents(b, 1024);
    } catch (...) {
      if (b->s) (void)hipStreamDestroy(b->s);
      delete b;
      throw;
    }
    *out = b;
  });
}

OTM_API int otm_batch_destroy(otm_batch* b) {
  return guarded([&] {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->s) (void)hipStreamSynchronize(b->s);
    dfree(b->dcells);
    dfree(b->dread);
    for (int h = 0; h < 2; h++) {
      dfree(b->dents[h]);
      if (b->hents[h]) (void)hipHostFree(b->hents[h]);
    }
    if (b->hread) (void)hipHostFree(b->hread);
    if (b->s) (void)hipStreamDestroy(b->s);
    delete b;
  });
}

OTM_API int otm_batch_heat_matrix(otm_batch* b, otm_ctx* const* cells, int n, double* C, int* status) {
  return guarded([&] {
    if (!b || (!cells && n > 0) || n < 0 || (!C && n > 0)) fail(OTM_E_ARG, "null argument");
    hip_check(hipSetDevice(b->device), "hipSetDevice");
    for (int i = 0; i < n; i++) {
      if (!cells[i]) fail(OTM_E_ARG, "null context");
      if (cells[i]->device != b->device) fail(OTM_E_ARG, "context on another device than the batch");
      if (!cells[i]->has_kappa) fail(OTM_E_STATE, "otm_batch_heat_matrix before otm_update");
      for (int j = 0; j < i; j++)
        if (cells[j] == cells[i]) fail(OTM_E_ARG, "context listed twice");
      hip_check(hipStreamSynchronize(cells[i]->s), "hipStreamSynchronize");
    }
    ensure_cells(b, n);
    b->members.clear();
    b->members.reserve(n);
    for (int i = 0; i < n; i++) add_member(b, cells[i], i, i);
    while (std::any_of(b->members.begin(), b->members.end(), live)) {
      step(b);
      finish(b);
    }
    hip_check(hipStreamSynchronize(b->s), "hipStreamSynchronize");
    int first_fail = OTM_OK;
    std::string msg;
    for (Member& m : b->members) {
      const int st = conclude(m);
      if (status) status[m.out] = st;
      if (st == OTM_OK || st == OTM_E_NOCONV)  // published also when the target was missed
        memcpy(C + 9 * m.out, m.C, sizeof(m.C));
      if (st != OTM_OK) {
        if (st != OTM_E_NOCONV)
          for (int q = 0; q < 9; q++) C[9 * m.out + q] = NAN;
        if (first_fail == OTM_OK) {
          first_fail = st;
          msg = m.fail_msg;
        }
      }
    }
    b->members.clear();
    if (first_fail != OTM_OK && !status) fail(first_fail, msg.empty() ? "a batched solve failed" : msg);
  });
}

OTM_API int otm_batch_sensitivity(otm_batch* b, otm_ctx* const* cells, int n, const double* dC, float* const* sens,
                                  int pitch, int where) {
  return guarded([&] {
    if (!b || (n > 0 && (!cells || !dC || !sens)) || n < 0) fail(OTM_E_ARG, "null argument");
    check_where(where);
    hip_check(hipSetDevice(b->device), "hipSetDevice");
    std::vector<long long> offs(n + 1, 0);
    for (int i = 0; i < n; i++) {
      if (!cells[i] || !sens[i]) fail(OTM_E_ARG, "null context or output");
      if (cells[i]->device != b->device) fail(OTM_E_ARG, "context on another device than the batch");
      if (!cells[i]->has_solution) fail(OTM_E_STATE, "otm_batch_sensitivity before a solve");
      const int pt = pitch > 0 ? pitch : cells[i]->g0.nx;
      if (pt < cells[i]->g0.nx) fail(OTM_E_ARG, "pitch smaller than reso[0]");
      offs[i + 1] = offs[i] + (long long)pt * cells[i]->g0.ny * cells[i]->g0.nz;
      hip_check(hipStreamSynchronize(cells[i]->s), "hipStreamSynchronize");
    }
    if (n == 0) return;
    float* stage = nullptr;
    if (where == OTM_HOST) stage = dalloc<float>(offs[n]);
    try {
      ensure_cells(b, n);
      b->members.assign(n, Member());
      b->hcells.resize(n);
      std::vector<Member*> set;
      for (int i = 0; i < n; i++) {
        Member& m = b->members[i];
        m.c = cells[i];
        m.slot = i;
        m.nlev = (int)cells[i]->lv.size();
        CellDev d = describe(m);
        float w[6];
        const double* q = dC + 9 * i;
        // contraction weights for S in the order (00,01,02,11,12,22)
        w[0] = (float)q[0];
        w[1] = (float)(q[1] + q[3]);
        w[2] = (float)(q[2] + q[6]);
        w[3] = (float)q[4];
        w[4] = (float)(q[5] + q[7]);
        w[5] = (float)q[8];
        for (int k = 0; k < 6; k++) d.sw[k] = w[k];
        d.sens = where == OTM_HOST ? stage + offs[i] : sens[i];
        d.sens_pitch = pitch > 0 ? pitch : cells[i]->g0.nx;
        b->hcells[i] = d;
        set.push_back(&m);
        if (d.sens_pitch != cells[i]->g0.nx) {
          const long long tot = offs[i + 1] - offs[i];
          if (where == OTM_HOST)
            hip_check(hipMemsetAsync(d.sens, 0, sizeof(float) * tot, b->s), "clear sens");
          else
            hip_check(hipMemsetAsync(sens[i], 0, sizeof(float) * tot, b->s), "clear sens");
        }
      }
      hip_check(hipMemcpyAsync(b->dcells, b->hcells.data(), sizeof(CellDev) * n, hipMemcpyHostToDevice, b->s),
                "upload cells");
      add_op(b, set, OpSensitivity{}, [](const Member& m) { return (long long)nblk(m.c->g0.nn); });
      flush(b);
      if (where == OTM_HOST)
        for (int i = 0; i < n; i++)
          hip_check(hipMemcpyAsync(sens[i], stage + offs[i], sizeof(float) * (offs[i + 1] - offs[i]),
                                   hipMemcpyDeviceToHost, b->s),
                    "copy sens");
      hip_check(hipStreamSynchronize(b->s), "hipStreamSynchronize");
    } catch (...) {
      b->members.clear();
      dfree(stage);
      throw;
    }
    b->members.clear();
    dfree(stage);
  });
}

OTM_API int otm_batch_update(otm_batch* b, otm_ctx* const* cells, int n, const float* const* kappa, int pitch,
                             int where) {
  return guarded([&] {
    if (!b || (n > 0 && (!cells || !kappa)) || n < 0) fail(OTM_E_ARG, "null argument");
    check_where(where);
    hip_check(hipSetDevice(b->device), "hipSetDevice");
    for (int i = 0; i < n; i++) {
      if (!cells[i] || !kappa[i]) fail(OTM_E_ARG, "null context or conductivity");
      if (cells[i]->device != b->device) fail(OTM_E_ARG, "context on another device than the batch");
      const int pt = pitch > 0 ? pitch : cells[i]->g0.nx;
      if (pt < cells[i]->g0.nx) fail(OTM_E_ARG, "pitch smaller than reso[0]");
      if (where == OTM_HOST && (long long)pt * cells[i]->g0.ny * cells[i]->g0.nz > 3 * cells[i]->g0.nn)
        fail(OTM_E_ARG, "host pitch too wide for the staging buffer");
      for (int j = 0; j < i; j++)
        if (cells[j] == cells[i]) fail(OTM_E_ARG, "context listed twice");
      hip_check(hipStreamSynchronize(cells[i]->s), "hipStreamSynchronize");
    }
    if (n == 0) return;
    ensure_cells(b, n);
    std::vector<Member> tmp(n);
    std::vector<const float*> src(n);
    for (int i = 0; i < n; i++) {
      otm_ctx* c = cells[i];
      tmp[i].c = c;
      tmp[i].slot = i;
      tmp[i].nlev = (int)c->lv.size();
      const int pt = pitch > 0 ? pitch : c->g0.nx;
      if (where == OTM_HOST) {
        // staged through the context's q vector, which the PCG rewrites before use
        hip_check(hipMemcpyAsync(c->q, kappa[i], sizeof(float) * pt * c->g0.ny * c->g0.nz, hipMemcpyHostToDevice,
                                 b->s),
                  "upload kappa");
        src[i] = c->q;
      } else {
        src[i] = kappa[i];
      }
    }
    plan_setup(b, tmp, src, pitch, false);
    flush(b);
    hip_check(hipStreamSynchronize(b->s), "hipStreamSynchronize");
    for (int i = 0; i < n; i++) cells[i]->has_kappa = true;
  });
}

OTM_API int otm_batch_stats(const otm_batch* b, long long* launches, int* steps) {
  return guarded([&] {
    if (!b) fail(OTM_E_ARG, "null batch");
    if (launches) *launches = b->launches;
    if (steps) *steps = b->steps;
  });
}

}  // extern "C"

// ---- serving: continuous batching over resident hierarchies ------------------
// A worker thread owns one batch. Between two PCG steps it admits queued
// requests while the member count and the HBM budget allow (a request's
// context comes from a cache of idle contexts of its shape, or is created),
// uploads and prepares the request's conductivity, and adds it to the
// running batch; members that finish leave at step boundaries and return
// their context to the cache. Requests therefore join work already on the
// GPU instead of waiting for it to drain, and every result is computed with
// the same per-cell arithmetic as when it is solved alone.
namespace {

using Clock = std::chrono::steady_clock;

struct ServerReq {
  long long id = 0;
  otm_config cfg;
  std::vector<float> kappa;
  double dC[9];
  bool want_sens = false;
  // results
  double C[9];
  std::vector<float> sens;
  int status = OTM_OK;
  std::string err;
  int iters[3] = {0, 0, 0};
  double relres[3] = {0, 0, 0};
  Clock::time_point t_submit, t_admit, t_done;
  bool done = false;
  // while admitted
  otm_ctx* ctx = nullptr;
  float* dsens = nullptr;
  size_t bytes = 0;
  int slot = -1;
};

struct IdleCtx {
  otm_ctx* ctx;
  float* dsens;
  size_t bytes;
  long long last_used;
};

// Device bytes of a context of this configuration (hierarchy, PCG vectors,
// coarsest matrices) plus the sensitivity buffer, from the level sizes.
size_t ctx_bytes(const otm_config& cfg) {
  long long d[3] = {cfg.reso[0], cfg.reso[1], cfg.reso[2]};
  const long long n0 = d[0] * d[1] * d[2];
  size_t bytes = (size_t)n0 * (4 + 6 * 12 + 12 + 4);  // kappa, x r z z2 p q, residual, sensitivity
  if (cfg.fp64_refine) bytes += (size_t)n0 * 24;
  const int minc = cfg.min_coarse > 2 ? cfg.min_coarse : 4;
  long long nc = n0;
  while ((d[0] % 2 == 0) && (d[1] % 2 == 0) && (d[2] % 2 == 0) && d[0] / 2 >= minc && d[1] / 2 >= minc &&
         d[2] / 2 >= minc) {
    for (int a = 0; a < 3; a++) d[a] /= 2;
    nc = d[0] * d[1] * d[2];
    bytes += (size_t)nc * (4 * 12 + 108 + 144 + (cfg.fp16_stencils ? 54 : 0));
  }
  bytes += (size_t)nc * nc * 8 * 3 + 2048 * 9 * 8;
  return bytes;
}

std::string shape_key(const otm_config& c) {
  char buf[160];
  snprintf(buf, sizeof(buf), "%d,%d,%d|%g|%g,%g|%d,%d,%d,%d,%d,%d", c.reso[0], c.reso[1], c.reso[2], c.fem_rel_thres,
           c.heat_ratio[0], c.heat_ratio[1], c.pre_sweeps, c.coarse_sweeps, c.min_coarse, c.fp64_refine,
           c.fp16_stencils, c.max_cg_iter);
  return buf;
}

}  // namespace

struct otm_server {
  int device = 0;
  size_t budget = 0;
  int max_active = 0;
  otm_batch* batch = nullptr;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<ServerReq>> queue;
  std::unordered_map<long long, std::shared_ptr<ServerReq>> reqs;
  std::multimap<std::string, IdleCtx> idle;
  long long next_id = 1;
  long long use_clock = 0;
  size_t resident = 0;  // bytes of contexts held (active and idle)
  bool stop = false;
  std::thread worker;
  // statistics (completed, created and evicted change under mu)
  long long completed = 0, created = 0, evicted = 0;
  std::atomic<long long> steps{0};
  std::atomic<int> active_peak{0};
};

namespace {

// Returns a request's context to the idle cache (keep) or frees it. A
// request whose context was never created still holds the budget reserved
// for it at admission; that reservation is given back here.
void release_ctx(otm_server* sv, ServerReq& r, bool keep) {
  if (!r.ctx) {
    sv->resident -= r.bytes;
    r.bytes = 0;
    dfree(r.dsens);
    r.dsens = nullptr;
    return;
  }
  if (keep) {
    sv->idle.emplace(shape_key(r.cfg), IdleCtx{r.ctx, r.dsens, r.bytes, sv->use_clock++});
  } else {
    otm_destroy(r.ctx);
    dfree(r.dsens);
    sv->resident -= r.bytes;
  }
  r.ctx = nullptr;
  r.dsens = nullptr;
  r.bytes = 0;
}

bool evict_one(otm_server* sv) {
  if (sv->idle.empty()) return false;
  auto victim = sv->idle.begin();
  for (auto it = sv->idle.begin(); it != sv->idle.end(); ++it)
    if (it->second.last_used < victim->second.last_used) victim = it;
  otm_destroy(victim->second.ctx);
  dfree(victim->second.dsens);
  sv->resident -= victim->second.bytes;
  sv->idle.erase(victim);
  sv->evicted++;
  return true;
}

void finish_request(otm_server* sv, ServerReq& r, int status, const std::string& err) {
  r.status = status;
  r.err = err;
  r.t_done = Clock::now();
  r.done = true;
  sv->completed++;
}

void server_loop(otm_server* sv) {
  (void)hipSetDevice(sv->device);
  otm_batch* b = sv->batch;
  std::map<int, std::shared_ptr<ServerReq>> active;  // by slot
  std::vector<int> free_slots;
  int next_slot = 0;
  while (true) {
    // ---- admission (between steps)
    std::vector<std::shared_ptr<ServerReq>> admit;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      if (active.empty())
        sv->cv_work.wait(lk, [&] { return sv->stop || !sv->queue.empty(); });
      if (sv->stop && sv->queue.empty() && active.empty()) break;
      while (!sv->queue.empty() && (int)(active.size() + admit.size()) < sv->max_active) {
        std::shared_ptr<ServerReq> r = sv->queue.front();
        const std::string key = shape_key(r->cfg);
        auto it = sv->idle.find(key);
        if (it != sv->idle.end()) {
          r->ctx = it->second.ctx;
          r->dsens = it->second.dsens;
          r->bytes = it->second.bytes;
          sv->idle.erase(it);
        } else {
          const size_t need = ctx_bytes(r->cfg);
          while (sv->resident + need > sv->budget && evict_one(sv)) {
          }
          if (sv->resident + need > sv->budget) {
            if (active.empty() && admit.empty()) {
              sv->queue.pop_front();
              finish_request(sv, *r, OTM_E_OOM, "request exceeds the server's HBM budget");
              sv->cv_done.notify_all();
              continue;
            }
            break;  // wait for members to leave
          }
          r->bytes = need;
          sv->resident += need;
        }
        sv->queue.pop_front();
        admit.push_back(r);
      }
    }
    // contexts and slots for the admitted requests, then one batched setup
    // (conductivity upload, Galerkin hierarchy, coarsest inverse, cold
    // start) for all of them on the batch's stream
    std::vector<std::shared_ptr<ServerReq>> ready;
    for (auto& r : admit) {
      try {
        if (!r->ctx) {
          const int cst = otm_create(&r->cfg, &r->ctx);
          if (cst != OTM_OK) fail(cst, otm_last_error());
          r->dsens = dalloc<float>((size_t)r->cfg.reso[0] * r->cfg.reso[1] * r->cfg.reso[2]);
          std::lock_guard<std::mutex> lk(sv->mu);
          sv->created++;
        }
        int slot;
        if (!free_slots.empty()) {
          slot = free_slots.back();
          free_slots.pop_back();
        } else {
          slot = next_slot++;
          if (slot >= b->dcap) {
            ensure_cells(b, slot + 1);
            for (auto& kv : active)  // the new array holds no descriptors yet
              hip_check(hipMemcpyAsync(b->dcells + kv.first, &b->hcells[kv.first], sizeof(CellDev),
                                       hipMemcpyHostToDevice, b->s),
                        "upload cell");
          }
        }
        r->slot = slot;
        ready.push_back(r);
      } catch (const OtmError& e) {
        std::lock_guard<std::mutex> lk(sv->mu);
        release_ctx(sv, *r, false);
        finish_request(sv, *r, e.code, e.msg);
        sv->cv_done.notify_all();
      }
    }
    if (!ready.empty()) {
      try {
        std::vector<Member> tmp(ready.size());
        std::vector<const float*> src(ready.size());
        for (size_t i = 0; i < ready.size(); i++) {
          otm_ctx* c = ready[i]->ctx;
          tmp[i].c = c;
          tmp[i].slot = ready[i]->slot;
          tmp[i].nlev = (int)c->lv.size();
          hip_check(hipMemcpyAsync(c->q, ready[i]->kappa.data(), sizeof(float) * c->g0.nn, hipMemcpyHostToDevice,
                                   b->s),
                    "upload kappa");
          src[i] = c->q;
        }
        plan_setup(b, tmp, src, -1, true);
        flush(b);
        for (auto& r : ready) {
          r->ctx->has_kappa = true;
          r->t_admit = Clock::now();
          add_member(b, r->ctx, r->slot, r->slot);
          active[r->slot] = r;
        }
      } catch (const OtmError& e) {
        std::lock_guard<std::mutex> lk(sv->mu);
        for (auto& r : ready) {
          free_slots.push_back(r->slot);
          release_ctx(sv, *r, false);
          finish_request(sv, *r, e.code, e.msg);
        }
        sv->cv_done.notify_all();
      }
    }
    if ((int)active.size() > sv->active_peak.load()) sv->active_peak.store((int)active.size());
    if (active.empty()) continue;
    // ---- one step of every member, then the members that finished
    try {
      step(b);
      finish(b);
    } catch (const OtmError& e) {
      // a HIP failure poisons every member of this step
      std::lock_guard<std::mutex> lk(sv->mu);
      for (auto& kv : active) {
        release_ctx(sv, *kv.second, false);
        finish_request(sv, *kv.second, e.code, e.msg);
      }
      active.clear();
      b->members.clear();
      sv->cv_done.notify_all();
      continue;
    }
    sv->steps++;
    std::vector<Member> left;
    std::vector<Member> keep;
    for (Member& m : b->members) (live(m) ? keep : left).push_back(m);
    if (left.empty()) continue;
    b->members.swap(keep);
    // sensitivities of the members that finished, in one launch
    std::vector<Member*> sens_set;
    for (Member& m : left) {
      const int st = conclude(m);
      auto& r = active[m.slot];
      r->status = st;
      if (st != OTM_OK) r->err = m.fail_msg.empty() ? "solve failed" : m.fail_msg;
      memcpy(r->C, m.C, sizeof(r->C));
      for (int k = 0; k < 3; k++) {
        r->iters[k] = r->ctx->iters[k];
        r->relres[k] = m.relres[k];
      }
      const bool solved = st == OTM_OK || st == OTM_E_NOCONV;  // NOCONV: results published, status says so
      if (solved && r->want_sens) {
        CellDev d = describe(m);
        const double* q = r->dC;
        d.sw[0] = (float)q[0];
        d.sw[1] = (float)(q[1] + q[3]);
        d.sw[2] = (float)(q[2] + q[6]);
        d.sw[3] = (float)q[4];
        d.sw[4] = (float)(q[5] + q[7]);
        d.sw[5] = (float)q[8];
        d.sens = r->dsens;
        d.sens_pitch = r->cfg.reso[0];
        b->hcells[m.slot] = d;
        hip_check(hipMemcpyAsync(b->dcells + m.slot, &b->hcells[m.slot], sizeof(CellDev), hipMemcpyHostToDevice,
                                 b->s),
                  "upload cell");
        sens_set.push_back(&m);
      }
    }
    try {
      if (!sens_set.empty()) {
        add_op(b, sens_set, OpSensitivity{}, [](const Member& m) { return (long long)nblk(m.c->g0.nn); });
        flush(b);
        for (Member* m : sens_set) {
          auto& r = active[m->slot];
          r->sens.resize((size_t)m->c->g0.nn);
          hip_check(hipMemcpyAsync(r->sens.data(), r->dsens, sizeof(float) * m->c->g0.nn, hipMemcpyDeviceToHost,
                                   b->s),
                    "copy sens");
        }
        hip_check(hipStreamSynchronize(b->s), "hipStreamSynchronize");
      }
    } catch (const OtmError& e) {
      for (Member* m : sens_set) {
        active[m->slot]->status = e.code;
        active[m->slot]->err = e.msg;
      }
    }
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      for (Member& m : left) {
        auto r = active[m.slot];
        release_ctx(sv, *r, r->status == OTM_OK || r->status == OTM_E_NOCONV);
        finish_request(sv, *r, r->status, r->err);
        free_slots.push_back(m.slot);
        active.erase(m.slot);
      }
    }
    sv->cv_done.notify_all();
  }
}

}  // namespace

extern "C" {

OTM_API int otm_server_create(int device, double hbm_budget_bytes, int max_active, otm_server** out) {
  return guarded([&] {
    if (!out || !(hbm_budget_bytes > 0) || max_active <= 0) fail(OTM_E_ARG, "invalid argument");
    *out = nullptr;
    size_t free_b = 0, total_b = 0;
    hip_check(hipSetDevice(device), "hipSetDevice");
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    if (hbm_budget_bytes > (double)free_b) fail(OTM_E_ARG, "HBM budget exceeds the device's free memory");
    otm_server* sv = new otm_server();
    sv->device = device;
    sv->budget = (size_t)hbm_budget_bytes;
    sv->max_active = max_active;
    if (otm_batch_create(device, &sv->batch) != OTM_OK) {
      const std::string m = otm_last_error();
      delete sv;
      fail(OTM_E_HIP, m);
    }
    sv->worker = std::thread(server_loop, sv);
    *out = sv;
  });
}

OTM_API int otm_server_destroy(otm_server* sv) {
  return guarded([&] {
    if (!sv) return;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->stop = true;
    }
    sv->cv_work.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    (void)hipSetDevice(sv->device);
    for (auto& kv : sv->idle) {
      otm_destroy(kv.second.ctx);
      dfree(kv.second.dsens);
    }
    otm_batch_destroy(sv->batch);
    delete sv;
  });
}

OTM_API void otm_request_init(otm_request* r) {
  if (!r) return;
  memset(r, 0, sizeof(*r));
  otm_config_init(&r->cfg);
}

OTM_API int otm_server_submit(otm_server* sv, const otm_request* req, long long* id) {
  return guarded([&] {
    if (!sv || !req || !id || !req->kappa) fail(OTM_E_ARG, "null argument");
    const otm_config& c = req->cfg;
    (void)plan_levels(c);  // the shapes otm_create accepts, checked before anything is reserved
    if (c.device != sv->device) fail(OTM_E_ARG, "request for another device than the server");
    auto r = std::make_shared<ServerReq>();
    r->cfg = c;
    const size_t n = (size_t)c.reso[0] * c.reso[1] * c.reso[2];
    r->kappa.assign(req->kappa, req->kappa + n);
    memcpy(r->dC, req->dC, sizeof(r->dC));
    r->want_sens = req->want_sens != 0;
    r->t_submit = Clock::now();
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      if (sv->stop) fail(OTM_E_STATE, "server is shutting down");
      r->id = sv->next_id++;
      sv->reqs[r->id] = r;
      sv->queue.push_back(r);
    }
    sv->cv_work.notify_one();
    *id = r->id;
  });
}

OTM_API int otm_server_wait(otm_server* sv, long long id, otm_result* res, float* sens) {
  return guarded([&] {
    if (!sv || !res) fail(OTM_E_ARG, "null argument");
    std::shared_ptr<ServerReq> r;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      auto it = sv->reqs.find(id);
      if (it == sv->reqs.end()) fail(OTM_E_ARG, "unknown request id");
      r = it->second;
      sv->cv_done.wait(lk, [&] { return r->done; });
      sv->reqs.erase(id);
    }
    memset(res, 0, sizeof(*res));
    res->status = r->status;
    memcpy(res->C, r->C, sizeof(res->C));
    for (int k = 0; k < 3; k++) {
      res->iters[k] = r->iters[k];
      res->relres[k] = r->relres[k];
    }
    auto ms = [](Clock::time_point a, Clock::time_point b) {
      return std::chrono::duration<double, std::milli>(b - a).count();
    };
    res->latency_ms = ms(r->t_submit, r->t_done);
    res->queue_ms = r->t_admit.time_since_epoch().count() ? ms(r->t_submit, r->t_admit) : res->latency_ms;
    if (sens && (r->status == OTM_OK || r->status == OTM_E_NOCONV) && !r->sens.empty()) memcpy(sens, r->sens.data(), sizeof(float) * r->sens.size());
    if (r->status != OTM_OK) fail(r->status, r->err);
  });
}

OTM_API int otm_server_stats(otm_server* sv, otm_server_counters* st) {
  return guarded([&] {
    if (!sv || !st) fail(OTM_E_ARG, "null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    st->completed = sv->completed;
    st->steps = sv->steps.load();
    st->contexts_created = sv->created;
    st->contexts_evicted = sv->evicted;
    st->active_peak = sv->active_peak.load();
    st->queued = (int)sv->queue.size();
    st->resident_bytes = (double)sv->resident;
    st->budget_bytes = (double)sv->budget;
  });
}

}  // extern "C"
