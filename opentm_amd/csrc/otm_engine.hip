// otm_engine.hip -- host engine behind include/opentm.h.
//
// One context = one periodic cell with its multigrid hierarchy, three
// lock-step PCG recurrences (one per load case), a HIP stream and the
// design-density buffers of the optimisation loop. There is no global device
// state: every kernel receives its level geometry and constants by value,
// so any number of contexts can be resident at once (the reference keeps
// per-grid state in __constant__ symbols, grid.cu:19-56, and allows one
// hierarchy per process).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/opentm.h"
#include "otm_ctx.hpp"

using namespace otm;

// ---- error plumbing --------------------------------------------------------
static thread_local std::string g_last_error;
void otm::set_last_error(const std::string& msg) { g_last_error = msg; }

namespace {

// Element constants (closed forms of templateMatrix.cpp:21-87 after the node
// remap {1,3,2,0,5,7,6,4}): load case k is a unit gradient along grid axis
// kAxis[k] with sign kSign[k].
constexpr int kAxis[3] = {1, 0, 2};
constexpr double kSign[3] = {1.0, -1.0, 1.0};

inline double ke_pattern(int a, int b) {
  if (a == b) return 4.0;
  return __builtin_popcount(a ^ b) == 1 ? 0.0 : -1.0;
}

FineConsts make_fine_consts(double h) {
  FineConsts fc;
  fc.c = (float)(h / 12.0);
  for (int l = 0; l < 8; l++)
    for (int k = 0; k < 3; k++) {
      const int bit = (l >> kAxis[k]) & 1;
      fc.femu[l][k] = (float)(kSign[k] * (bit ? 1.0 : -1.0) * h * h / 4.0);
      fc.disp[l][k] = (float)(kSign[k] * bit * h);
    }
  return fc;
}

inline double w1d_h(int p, int M) {
  const int d = p - 2 * M;
  return d == 0 ? 1.0 : ((d == 1 || d == -1) ? 0.5 : 0.0);
}

GalerkinConsts make_galerkin_consts(double h) {
  GalerkinConsts G;
  const double c = h / 12.0;
  for (int ch = 0; ch < 8; ch++) {
    const int cx = ch & 1, cy = (ch >> 1) & 1, cz = ch >> 2;
    double Q[8][8];
    for (int m = 0; m < 8; m++)
      for (int M = 0; M < 8; M++)
        Q[m][M] = w1d_h(cx + (m & 1), M & 1) * w1d_h(cy + ((m >> 1) & 1), (M >> 1) & 1) *
                  w1d_h(cz + (m >> 2), M >> 2);
    for (int a = 0; a < 8; a++)
      for (int b = a; b < 8; b++) {
        double s = 0;
        for (int m = 0; m < 8; m++)
          for (int n = 0; n < 8; n++) s += Q[m][a] * c * ke_pattern(m, n) * Q[n][b];
        G.G[ch][tri8(a, b)] = (float)s;
      }
  }
  return G;
}

}  // namespace

// Radial taps (TensorExpression.h:1053-1093): Spline4 1-6q+8q^1.5-3q^2,
// Linear 1-sqrt(q), q = |d|^2/r^2 < 1, window half-width int(r+0.5), each
// weight divided by the window total.
Taps otm::make_taps(int kind, double r, int periodic) {
  if (!(r > 0.0)) fail(OTM_E_ARG, "filter radius must be positive");
  const int h = (int)(r + 0.5);
  if ((2 * h + 1) * (2 * h + 1) * (2 * h + 1) > kMaxTaps) fail(OTM_E_UNSUPPORTED, "filter radius too large (max 3.49)");
  if (kind != 0 && kind != 1) fail(OTM_E_ARG, "filter kind must be 0 (Spline4) or 1 (Linear)");
  Taps T;
  memset(&T, 0, sizeof(T));
  T.periodic = periodic ? 1 : 0;
  double tot = 0.0;
  std::vector<double> ws;
  std::vector<int> offs;
  for (int dz = -h; dz <= h; dz++)
    for (int dy = -h; dy <= h; dy++)
      for (int dx = -h; dx <= h; dx++) {
        const double q = (double)(dx * dx + dy * dy + dz * dz) / (r * r);
        double w = 0.0;
        if (q < 1.0) w = kind == 0 ? 1.0 - 6.0 * q + 8.0 * q * sqrt(q) - 3.0 * q * q : 1.0 - sqrt(q);
        tot += w;
        if (w != 0.0) {
          ws.push_back(w);
          offs.push_back(dx);
          offs.push_back(dy);
          offs.push_back(dz);
        }
      }
  T.n = (int)ws.size();
  for (int i = 0; i < T.n; i++) {
    T.w[i] = (float)(ws[i] / tot);
    T.dx[i] = (signed char)offs[3 * i];
    T.dy[i] = (signed char)offs[3 * i + 1];
    T.dz[i] = (signed char)offs[3 * i + 2];
  }
  return T;
}

struct otm_oc {
  int ne = 0;
  int device = 0;
  double min_rho = 0.001, step = 0.02, damp = 0.5;
  int max_steps = 20;
  double vol_tol = 1e-4;
  hipStream_t s = nullptr;
  float *newsens = nullptr, *a = nullptr, *b = nullptr;
  unsigned* maxbits = nullptr;
  OcState* st = nullptr;
  OcHost* hm = nullptr;      // coherent pinned block the apply kernel publishes to
  OcHost* hm_dev = nullptr;  // its device address
  unsigned tag = 0;
  double* partials = nullptr;
};

// ---- context lifetime --------------------------------------------------------
static void ctx_free(otm_ctx* c) {
  if (!c) return;
  if (c->s) (void)hipStreamSynchronize(c->s);
  for (auto& L : c->lv) {
    if (L.own_uf) {
      dfree(L.u);
      dfree(L.f);
    }
    dfree(L.res);
    dfree(L.st);
    dfree(L.E);
    dfree(L.st16);
    dfree(L.scale);
    if (L.own_uf) dfree(L.u2);
  }
  dfree(c->energy);
  dfree(c->kappa); dfree(c->x); dfree(c->x64); dfree(c->xlo); dfree(c->r); dfree(c->z); dfree(c->p); dfree(c->q); dfree(c->z2);
  dfree(c->cB); dfree(c->cBinv);
  dfree(c->partials); dfree(c->tile_partials); dfree(c->red_out); dfree(c->pcg);
  if (c->pcg_host) (void)hipHostFree(c->pcg_host);
  if (c->out_host) (void)hipHostFree(c->out_host);
  if (c->hm) (void)hipHostFree(c->hm);
  dfree(c->rho); dfree(c->rho_f); dfree(c->W); dfree(c->kraw); dfree(c->dkappa);
  dfree(c->tmp); dfree(c->drho);
  for (auto& p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  for (auto& v : c->vgraphs) (void)hipGraphExecDestroy(v.exec);
  if (c->s_side) (void)hipStreamSynchronize(c->s_side);
  if (c->ev_main) (void)hipEventDestroy(c->ev_main);
  if (c->ev_side) (void)hipEventDestroy(c->ev_side);
  if (c->s_side) (void)hipStreamDestroy(c->s_side);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
}

// Level sizes of the hierarchy a configuration gets (fine to coarse); throws
// OTM_E_UNSUPPORTED for shapes the engine does not solve. Shared by context
// creation and the server's admission check.
std::vector<Geom> otm::plan_levels(const otm_config& cfg) {
  for (int a = 0; a < 3; a++) {
    if (cfg.reso[a] < 8 || (cfg.reso[a] & 1)) fail(OTM_E_UNSUPPORTED, "each resolution must be even and >= 8");
    if (cfg.reso[a] > 1024) fail(OTM_E_UNSUPPORTED, "resolution exceeds 1024 (grid.cpp:23-25 limit)");
  }
  const int minc = cfg.min_coarse > 2 ? cfg.min_coarse : 4;
  int d[3] = {cfg.reso[0], cfg.reso[1], cfg.reso[2]};
  std::vector<Geom> geoms;
  while (true) {
    Geom g{d[0], d[1], d[2], (long long)d[0] * d[1] * d[2]};
    geoms.push_back(g);
    bool can = true;
    for (int a = 0; a < 3; a++)
      if ((d[a] & 1) || d[a] / 2 < minc) can = false;
    if (!can) break;
    for (int a = 0; a < 3; a++) d[a] /= 2;
  }
  if (geoms.size() < 2) fail(OTM_E_UNSUPPORTED, "resolution too small for a two-level hierarchy");
  const long long ncoarse = geoms.back().nn;
  if (ncoarse > 4096) fail(OTM_E_UNSUPPORTED, "coarsest level has " + std::to_string(ncoarse) +
                                                  " nodes (> 4096); use resolutions with more factors of two");
  if (ncoarse > kTailLdsNodes) fail(OTM_E_UNSUPPORTED, "coarsest level too large for the tail");
  return geoms;
}

static void build_hierarchy(otm_ctx* c) {
  const otm_config& cfg = c->cfg;
  const std::vector<Geom> geoms = plan_levels(cfg);
  const long long ncoarse = geoms.back().nn;
  c->g0 = geoms[0];
  const long long n0 = c->g0.nn;
  c->kappa = dalloc<float>(n0);
  c->x = dalloc<float>(3 * n0);
  c->r = dalloc<float>(3 * n0);
  c->z = dalloc<float>(3 * n0);
  c->p = dalloc<float>(3 * n0);
  c->q = dalloc<float>(3 * n0);
  for (size_t l = 0; l < geoms.size(); l++) {
    Level L;
    L.g = geoms[l];
    const long long n = L.g.nn;
    if (l == 0) {
      L.u = c->z;
      L.f = c->r;
      L.own_uf = false;
      c->fine_tile = cfg.fine_tile < 0 ? fine_tile_default(L.g) : cfg.fine_tile;
      if (c->fine_tile >= 0 && !fine_tile_ok(L.g, c->fine_tile))
        fail(OTM_E_UNSUPPORTED, "the cell's edges are not multiples of the requested fine tile shape");
      if (c->fine_tile >= 0 && cfg.pre_sweeps == cfg.post_sweeps) {
        L.brick = true;
        c->z2 = dalloc<float>(3 * n);
        L.u2 = c->z2;
      }
    } else {
      L.u = dalloc<float>(3 * n);
      L.f = dalloc<float>(3 * n);
      L.st = dalloc<float>(27 * n);
      L.E = dalloc<float>(36 * n);
      if (cfg.fp16_stencils && l + 1 < geoms.size()) {
        L.st16 = dalloc<_Float16>(27 * n);
        L.scale = dalloc<float>(1);
      }
      if (l + 1 < geoms.size() && brick_ok(L.g) && cfg.pre_sweeps == cfg.post_sweeps) {
        L.brick = true;
        L.u2 = dalloc<float>(3 * n);
      }
    }
    if (l + 1 < geoms.size()) L.res = dalloc<float>(3 * n);
    c->lv.push_back(L);
  }
  c->cB = dalloc<double>(coarsest_work_doubles(ncoarse));  // Gauss-Jordan work space
  // the tail V-cycle takes over from the first coarse level with at most
  // kTailStartNodes nodes whose levels fit the tail's LDS (the coarsest level
  // alone otherwise). A tail from 16^3 with stencils re-read from global
  // memory measured slower than the separate launches (131 against about
  // 100 us per V-cycle at 128^3); from 8^3 with the stencils in LDS it
  // replaces eleven launches.
  c->tail = (int)c->lv.size() - 1;
  for (size_t l = 1; l + 1 < c->lv.size(); l++) {
    long long st_nodes = 0, nodes = 0;
    for (size_t m = l; m < c->lv.size(); m++) {
      nodes += c->lv[m].g.nn;
      if (m + 1 < c->lv.size()) st_nodes += c->lv[m].g.nn;
    }
    if (c->lv[l].g.nn <= kTailStartNodes && (int)(c->lv.size() - l) <= kTailMaxLevels &&
        st_nodes <= kTailStNodes && nodes <= kTailLdsNodes) {
      c->tail = (int)l;
      break;
    }
  }
  c->cBinv = dalloc<double>(ncoarse * ncoarse);
  c->partials = dalloc<double>((size_t)kMaxReduceBlocks * kMaxReduce);
  // one row per fine tile of the smallest shape (32 x 4 x 4 = 512 nodes)
  c->tile_partials = dalloc<double>((size_t)std::min<long long>(kTileDotSlots, n0 / 512 + 1) * OTM_NRHS);
  c->red_out = dalloc<double>(16);
  c->pcg = dalloc<PcgScalars>(1);
  hip_check(hipMemsetAsync(c->pcg, 0, sizeof(PcgScalars), c->s), "hipMemset pcg");
  hip_check(hipHostMalloc((void**)&c->pcg_host, sizeof(PcgScalars), hipHostMallocDefault), "hipHostMalloc");
  hip_check(hipHostMalloc((void**)&c->out_host, 16 * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
  hip_check(hipHostMalloc((void**)&c->hm, sizeof(HostMirror), hipHostMallocCoherent), "hipHostMalloc");
  memset(c->hm, 0, sizeof(HostMirror));
  hip_check(hipHostGetDevicePointer((void**)&c->hm_dev, c->hm, 0), "hipHostGetDevicePointer");
  launch_fill(c->s, c->x, 3 * n0, 0.f);
  launch_fill(c->s, c->p, 3 * n0, 0.f);
  if (cfg.fp64_refine) {
    c->x64 = dalloc<double>(3 * n0);
    launch_fill_d(c->s, c->x64, 3 * n0, 0.0);
  } else {
    c->xlo = dalloc<float>(3 * n0);
    launch_fill(c->s, c->xlo, 3 * n0, 0.f);
  }
  hip_check(hipGetLastError(), "initial fill");
  const double h = 1.0 / cfg.reso[0];  // element side (homogenization.cpp:47 passes h/2 as half-length)
  c->fc = make_fine_consts(h);
  c->G = make_galerkin_consts(h);
}

// Galerkin coarse operators + coarsest factorisation (MG_H::updateStencils,
// MG.cpp:144-150).
// start_res: launch the next solve's FP64 start residual (it needs only
// kappa and the warm-start iterate) on the side stream while the coarsest
// inverse, one workgroup on one CU, is formed; the main stream waits for it
// afterwards. Overlapping it with the bandwidth-bound Galerkin kernels
// instead saved nothing.
static void build_operators(otm_ctx* c, bool start_res = false) {
  ProfScope ps(c, OTM_PROF_SETUP, 0.0);
  for (size_t l = 1; l < c->lv.size(); l++) {
    Level& L = c->lv[l];
    if (l == 1)
      launch_galerkin_first(c->s, c->g0, L.g, c->G, c->kappa, L.E);
    else
      launch_galerkin_next(c->s, c->lv[l - 1].g, L.g, c->lv[l - 1].E, L.E);
    launch_stencil_from_elements(c->s, L.g, L.E, L.st);
    if (L.st16) launch_stencil_pack(c->s, L.g, L.st, L.scale, L.st16);
  }
  const Level& Lc = c->lv.back();
  if (start_res) {
    if (!c->s_side) {
      hip_check(hipStreamCreateWithFlags(&c->s_side, hipStreamNonBlocking), "hipStreamCreate");
      hip_check(hipEventCreateWithFlags(&c->ev_main, hipEventDisableTiming), "hipEventCreate");
      hip_check(hipEventCreateWithFlags(&c->ev_side, hipEventDisableTiming), "hipEventCreate");
    }
    const double rtol2 = c->cfg.fem_rel_thres * c->cfg.fem_rel_thres;
    hip_check(hipEventRecord(c->ev_main, c->s), "hipEventRecord");
    hip_check(hipStreamWaitEvent(c->s_side, c->ev_main, 0), "hipStreamWaitEvent");
    launch_fine_init_residual(c->s_side, c->g0, c->fc, c->kappa, c->x, c->xlo, c->r,
                              c->reducer_host(FIN_BB_RR, rtol2, 1.0 / (double)c->g0.nn), red_blocks_for(c->g0.nn));
    hip_check(hipEventRecord(c->ev_side, c->s_side), "hipEventRecord");
  }
  launch_coarsest_factor(c->s, Lc.g, Lc.st, c->cB, c->cBinv);
  if (start_res) {
    hip_check(hipStreamWaitEvent(c->s, c->ev_side, 0), "hipStreamWaitEvent");
    c->start_res_ready = true;
  }
  hip_check(hipGetLastError(), "operator build");
}

// ---- multigrid V-cycle (MG_H::v_cycle, MG.cpp:46-71) ----------------------------
static void smooth(otm_ctx* c, size_t l, int sweeps, bool forward) {
  Level& L = c->lv[l];
  const int ncol = l == 0 ? 4 : 8;
  for (int s = 0; s < sweeps; s++)
    for (int i = 0; i < ncol; i++) {
      const int col = forward ? i : ncol - 1 - i;
      const double nn = (double)L.g.nn;
      if (l == 0) {
        // kappa 4 + other colours' rows 3*4*3/4 + f and u of the colour 2*3*4/4 bytes per node
        ProfScope ps(c, OTM_PROF_FINE_GS, 19.0 * nn);
        launch_fine_gs_color(c->s, L.g, c->fc, c->kappa, L.f, L.u, col);
      } else {
        // stencil rows of the colour 27*4/8 + other colours 3*4*7/8 + f and u 2*3*4/8
        ProfScope ps(c, OTM_PROF_COARSE_GS, 27.0 * nn);
        launch_st_gs_color(c->s, L.g, stref(L), L.f, L.u, col);
      }
    }
}

static TailLevels tail_levels(const otm_ctx* c, int nsw, float* uout) {
  TailLevels T;
  memset(&T, 0, sizeof(T));
  T.n = (int)c->lv.size() - c->tail;
  T.sweeps = nsw;
  for (int i = 0; i < T.n; i++) {
    const Level& L = c->lv[c->tail + i];
    T.g[i] = L.g;
    T.f[i] = L.f;
    T.res[i] = L.res;
    T.st[i] = stref(L);
  }
  T.f0 = c->lv[c->tail].f;
  T.Binv = c->cBinv;
  T.uout = uout;
  return T;
}

static void vcycle(otm_ctx* c, size_t l) {
  Level& L = c->lv[l];
  const double nn = (double)L.g.nn;
  const int nsw = (l > 0 && c->cfg.coarse_sweeps > 0) ? c->cfg.coarse_sweeps : c->cfg.pre_sweeps;
  if ((int)l == c->tail) {
    ProfScope ps(c, OTM_PROF_COARSE_GS, 0.0);
    launch_tail_vcycle(c->s, tail_levels(c, nsw, L.u));
    return;
  }
  if (l + 1 == c->lv.size()) {
    ProfScope ps(c, OTM_PROF_COARSEST, 8.0 * nn * nn + 24.0 * nn);
    launch_coarsest_solve(c->s, L.g, c->cBinv, L.f, L.u);
    return;
  }
  if (L.brick && l == 0) {
    // compulsory bytes per node: kappa 4 + f 12 + u written 12 (zero start)
    ProfScope ps(c, OTM_PROF_FINE_GS, 28.0 * nn);
    launch_fine_smooth_brick(c->s, L.g, c->fc, c->kappa, L.f, nullptr, L.u, nsw, true, true, c->fine_tile);
  } else if (L.brick) {
    // pre-smoothing from a zero guess straight into u
    ProfScope ps(c, OTM_PROF_COARSE_GS, 132.0 * nn);  // stencil 108 + f 12 + u 12 bytes per node
    launch_st_smooth_brick(c->s, L.g, stref(L), L.f, nullptr, L.u, nsw, true, true);
  } else {
    {
      ProfScope ps(c, OTM_PROF_VECTOR, 12.0 * nn);
      launch_fill(c->s, L.u, 3 * L.g.nn, 0.f);
    }
    smooth(c, l, nsw, true);
  }
  if (l == 0) {
    ProfScope ps(c, OTM_PROF_FINE_RESID, 40.0 * nn);
    launch_fine_residual(c->s, L.g, c->fc, c->kappa, L.f, L.u, L.res);
  } else {
    ProfScope ps(c, OTM_PROF_COARSE_RESID, 144.0 * nn);
    launch_st_residual(c->s, L.g, stref(L), L.f, L.u, L.res);
  }
  Level& C = c->lv[l + 1];
  {
    ProfScope ps(c, OTM_PROF_TRANSFER, 12.0 * nn + 12.0 * (double)C.g.nn);
    launch_restrict(c->s, L.g, C.g, L.res, C.f);
  }
  vcycle(c, l + 1);
  {
    ProfScope ps(c, OTM_PROF_TRANSFER, 24.0 * nn + 12.0 * (double)C.g.nn);
    launch_prolong_add(c->s, L.g, C.g, C.u, L.u);
  }
  if (L.brick && l == 0) {
    // r.z of the PCG from the post-smoother's own values when asked for
    const bool dot = c->post_dot_mode >= 0 && fine_brick_dot_ok(L.g, c->fine_tile);
    Reduce RD = c->reducer(dot ? c->post_dot_mode : FIN_STORE);
    RD.partials = c->tile_partials;
    {
      // compulsory bytes per node: kappa 4 + u read 12 + f 12 + u written 12
      ProfScope ps(c, OTM_PROF_FINE_GS, 40.0 * nn);
      launch_fine_smooth_brick(c->s, L.g, c->fc, c->kappa, L.f, L.u, L.u2, nsw, false, false, c->fine_tile,
                               dot ? &RD : nullptr);
    }
    if (dot) launch_fine_brick_dot_final(c->s, L.g, RD, c->fine_tile);
    c->post_dot_done = dot;
    std::swap(L.u, L.u2);
    std::swap(c->z, c->z2);
  } else if (L.brick) {
    // post-smoothing reads u (frozen halos) and writes u2; swap so u holds the result
    ProfScope ps(c, OTM_PROF_COARSE_GS, 155.0 * nn);  // plus the u halo read, 12 * 10^3/8^3
    launch_st_smooth_brick(c->s, L.g, stref(L), L.f, L.u, L.u2, nsw, false, false);
    std::swap(L.u, L.u2);
  } else {
    smooth(c, l, nsw, false);
  }
}

// ---- the V-cycle as a replayed HIP graph -------------------------------------
// A V-cycle's launches are fixed by a few pieces of host state: which buffer
// of each brick level holds the current correction (the brick smoothers
// write a second buffer and the host swaps the pointers), whether the fine
// post-smoother also forms the PCG's r.z (and for which finaliser), and
// whether the bricks are on. Each state's launch sequence is captured once
// into a graph and replayed; replaying applies the same pointer swaps on
// the host as the eager V-cycle. Timing with events (otm_profile_enable)
// runs the eager path.
static unsigned long long vcycle_key(const otm_ctx* c) {
  unsigned long long key = 0;
  for (size_t l = 0; l < c->lv.size() && l < 48; l++) {
    const Level& L = c->lv[l];
    if (L.u2 && (uintptr_t)L.u > (uintptr_t)L.u2) key |= 1ull << l;
  }
  key |= (unsigned long long)(c->post_dot_mode + 1) << 56;  // -1, FIN_RZ_FIRST, FIN_RZ
  key |= (unsigned long long)(c->bricks_on ? 1 : 0) << 62;
  key |= (unsigned long long)(c->fine_tile + 1) << 48;  // 0..5 in bits 48-50
  key |= (unsigned long long)(c->cfg.coarse_sweeps & 7) << 51;
  return key;
}

// The host-side effects of vcycle(c, 0) without its launches: the pointer
// swaps of the brick levels it smooths and the fused r.z flag.
static void vcycle_effects(otm_ctx* c) {
  const size_t stop = std::min<size_t>((size_t)c->tail, c->lv.size() - 1);
  c->post_dot_done = false;
  for (size_t l = 0; l < stop; l++) {
    Level& L = c->lv[l];
    if (!L.brick) continue;
    if (l == 0) {
      c->post_dot_done = c->post_dot_mode >= 0 && fine_brick_dot_ok(L.g, c->fine_tile);
      std::swap(c->z, c->z2);
    }
    std::swap(L.u, L.u2);
  }
}

static void run_vcycle(otm_ctx* c) {
  if (!c->cfg.graph_vcycle || c->graphs_failed || c->prof_on) {
    vcycle(c, 0);
    return;
  }
  const unsigned long long key = vcycle_key(c);
  for (const auto& v : c->vgraphs)
    if (v.key == key) {
      hip_check(hipGraphLaunch(v.exec, c->s), "hipGraphLaunch");
      vcycle_effects(c);
      return;
    }
  // capture this state's launches; vcycle() applies the host effects itself
  std::vector<std::pair<float*, float*>> saved;
  for (auto& L : c->lv) saved.push_back({L.u, L.u2});
  float *z = c->z, *z2 = c->z2;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool ok = hipStreamBeginCapture(c->s, hipStreamCaptureModeRelaxed) == hipSuccess;
  if (ok) {
    try {
      vcycle(c, 0);
    } catch (...) {
      ok = false;
    }
    const hipError_t e = hipStreamEndCapture(c->s, &graph);
    ok = ok && e == hipSuccess && graph != nullptr;
  }
  if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
  if (graph) (void)hipGraphDestroy(graph);
  if (!ok) {
    // no graph: restore the host state and run eagerly from now on
    (void)hipGetLastError();
    if (exec) (void)hipGraphExecDestroy(exec);
    for (size_t l = 0; l < c->lv.size(); l++) {
      c->lv[l].u = saved[l].first;
      c->lv[l].u2 = saved[l].second;
    }
    c->z = z;
    c->z2 = z2;
    c->graphs_failed = true;
    vcycle(c, 0);
    return;
  }
  c->vgraphs.push_back({key, exec});
  hip_check(hipGraphLaunch(exec, c->s), "hipGraphLaunch");
}

// ---- PCG over the three load cases -------------------------------------------
// Three preconditioned CG recurrences in lock step, one per load case, with
// the V-cycle as preconditioner (the reference iterates the V-cycle as a
// stationary method instead, MG.cpp:80-142). Each restart computes the true
// residual in FP64 and makes it mean-free:
// - plain solve: the iterate x is FP32 and the iteration restarts from the
//   recomputed residual while it still improves (residual correction);
// - fp64_refine: the iterate is kept in FP64 (x64) and each restart solves
//   for an FP32 correction (in x) down to a fixed reduction of its own start
//   residual, then adds it to x64 (iterative refinement), which reaches
//   relative residuals far below the FP32 resolution of the solution.
// Returns false when the preconditioner turned out not to be positive.
constexpr double kRefineInner = 1e-4;  // inner reduction per refinement step
// *fresh tells whether the last true residual (pcg_host's rr and bb) was
// evaluated at the final iterate, so the caller need not evaluate it again.
static bool pcg_iterate(otm_ctx* c, int* its, int* total, bool* fresh) {
  const long long n = c->g0.nn;
  const unsigned nb = red_blocks_for(n);
  const double rtol = c->cfg.fem_rel_thres;
  const double rtol2 = rtol * rtol;
  const double inv_n = 1.0 / (double)n;
  const bool refine = c->x64 != nullptr;
  const int max_restarts = refine ? 8 : 3;
  int it = 0;
  double prev_worst = 1e300;
  bool pending = false;  // a correction in x not yet added to x64
  *fresh = false;
  for (int restart = 0; restart < max_restarts; restart++) {
    {
      ProfScope ps(c, OTM_PROF_FINE_RESID, (refine ? 56.0 : 28.0) * (double)n);
      if (refine) {
        if (pending)
          launch_refine_add(c->s, 3 * n, c->x64, c->x);
        else
          launch_fill(c->s, c->x, 3 * n, 0.f);
        pending = false;
        Reduce R = c->reducer_host(FIN_BB_RR, rtol2, inv_n);
        R.inner2 = kRefineInner * kRefineInner;
        launch_fine_init_residual(c->s, c->g0, c->fc, c->kappa, c->x64, c->r, R, nb);
      } else if (c->start_res_ready) {
        c->start_res_ready = false;  // launched by do_update, same arguments
      } else {
        launch_fine_init_residual(c->s, c->g0, c->fc, c->kappa, c->x, c->xlo, c->r,
                                  c->reducer_host(FIN_BB_RR, rtol2, inv_n), nb);
      }
    }
    c->read_pcg();
    if (c->pcg_host->fault & 1) fail(OTM_E_NUMERIC, "non-finite residual (MG numerical explode)");
    *fresh = true;  // every exit below leaves x as just evaluated
    if (!c->pcg_host->active[0] && !c->pcg_host->active[1] && !c->pcg_host->active[2]) break;
    if (it >= c->cfg.max_cg_iter) break;
    // restart only while the true residual still improves: with the iterate
    // stored in FP32 it stalls near the FP32 resolution of the solution
    double worst = 0.0;
    for (int k = 0; k < 3; k++) {
      const double q = c->pcg_host->rr[k] / c->pcg_host->bb[k];
      worst = q > worst ? q : worst;
    }
    if (c->cfg.verbose)
      fprintf(stderr, "[otm] restart %d after %d iterations: true relres %.3e\n", restart, it, sqrt(worst));
    if (restart > 0 && worst > 0.25 * prev_worst) break;
    *fresh = false;
    prev_worst = worst;
    {
      ProfScope ps(c, OTM_PROF_VECTOR, 24.0 * (double)n);
      launch_sub_mean3(c->s, n, c->r, c->pcg);
    }
    c->post_dot_mode = FIN_RZ_FIRST;
    c->post_dot_done = false;
    run_vcycle(c);
    c->post_dot_mode = -1;
    {
      ProfScope ps(c, OTM_PROF_VECTOR, 24.0 * (double)n + 24.0 * (double)n);
      if (!c->post_dot_done) launch_dot3(c->s, n, c->r, c->z, c->reducer(FIN_RZ_FIRST), nb);
      launch_pcg_p(c->s, n, c->p, c->z, c->pcg, true);
    }
    pending = refine;
    while (it < c->cfg.max_cg_iter) {
      {
        ProfScope ps(c, OTM_PROF_FINE_APPLY, 28.0 * (double)n);
        launch_fine_apply_dot(c->s, c->g0, c->fc, c->kappa, c->p, c->q, c->reducer(FIN_PQ), nb);
      }
      {
        ProfScope ps(c, OTM_PROF_VECTOR, 72.0 * (double)n);
        launch_pcg_xr(c->s, n, c->x, c->xlo, c->r, c->p, c->q, c->pcg, c->reducer_host(FIN_RR, rtol2), nb);
      }
      it++;
      c->read_pcg();
      for (int k = 0; k < 3; k++)
        if (c->pcg_host->active[k]) its[k]++;
      if (c->pcg_host->fault & 1) fail(OTM_E_NUMERIC, "non-finite residual (MG numerical explode)");
      if (c->pcg_host->fault & 2) {
        *total = it;
        if (pending) launch_fill(c->s, c->x, 3 * n, 0.f);  // drop the partial correction
        return false;
      }
      if (!c->pcg_host->active[0] && !c->pcg_host->active[1] && !c->pcg_host->active[2]) break;
      c->post_dot_mode = FIN_RZ;
      c->post_dot_done = false;
      run_vcycle(c);
      c->post_dot_mode = -1;
      {
        ProfScope ps(c, OTM_PROF_VECTOR, 24.0 * (double)n + 36.0 * (double)n);
        if (!c->post_dot_done) launch_dot3(c->s, n, c->r, c->z, c->reducer(FIN_RZ), nb);
        launch_pcg_p(c->s, n, c->p, c->z, c->pcg);
      }
    }
  }
  if (pending) {
    launch_refine_add(c->s, 3 * n, c->x64, c->x);
    *fresh = false;
  }
  *total = it;
  return true;
}

static void set_bricks(otm_ctx* c, bool on) {
  for (size_t l = 0; l + 1 < c->lv.size(); l++) c->lv[l].brick = on && c->lv[l].u2 != nullptr;
  c->bricks_on = on;
}

// Removes the constant component of the iterate (translateForce,
// grid.cpp:654-664) and refreshes its FP32 copy; in stream order with what
// follows (no result the host needs).
static void remove_constant(otm_ctx* c) {
  const long long n = c->g0.nn;
  const unsigned nb = red_blocks_for(n);
  if (c->x64) {
    launch_sum3(c->s, n, c->x64, c->reducer(FIN_MEAN, 0.0, 1.0 / (double)n), nb);
    launch_sub_mean_cast(c->s, n, c->x64, c->x, c->pcg);
  } else {
    launch_sum3(c->s, n, c->x, c->reducer(FIN_MEAN, 0.0, 1.0 / (double)n), nb);
    launch_sub_mean3(c->s, n, c->x, c->pcg);
  }
}

// Returns false when the solve ended above the relative residual target
// (iteration cap, or the true residual stopped improving); the iterate and
// its statistics are kept either way.
static bool pcg_solve(otm_ctx* c) {
  c->energy_ok = false;
  const long long n = c->g0.nn;
  const unsigned nb = red_blocks_for(n);
  int its[3] = {0, 0, 0};
  int it = 0;
  c->post_dot_mode = -1;
  // the fault flags are cleared by the finaliser of each start residual
  bool fresh = false;
  if (!pcg_iterate(c, its, &it, &fresh)) {
    // The block-hybrid coarse smoother lost positivity on this operator:
    // fall back to exact multicolour Gauss-Seidel on every level (always a
    // positive preconditioner) for this context and solve again.
    set_bricks(c, false);
    c->fallbacks++;
    int it2 = 0;
    if (!pcg_iterate(c, its, &it2, &fresh)) fail(OTM_E_NUMERIC, "preconditioner not positive (MG numerical explode)");
    it += it2;
  }
  // true residual of the final iterate (already at hand when the restart
  // loop ended right after evaluating it)
  if (!fresh) {
    const Reduce Rf = c->reducer_host(FIN_BB_RR, c->cfg.fem_rel_thres * c->cfg.fem_rel_thres, 1.0 / (double)n);
    if (c->x64)
      launch_fine_init_residual(c->s, c->g0, c->fc, c->kappa, c->x64, c->r, Rf, nb);
    else
      launch_fine_init_residual(c->s, c->g0, c->fc, c->kappa, c->x, c->xlo, c->r, Rf, nb);
    c->read_pcg();
  }
  double relres[3];
  for (int k = 0; k < 3; k++) relres[k] = sqrt(c->pcg_host->rr[k] / c->pcg_host->bb[k]);
  remove_constant(c);
  for (int k = 0; k < 3; k++) {
    c->relres[k] = relres[k];
    c->iters[k] = its[k] + (its[k] < it ? 1 : 0);
  }
  if (c->pcg_host->fault & 1) fail(OTM_E_NUMERIC, "non-finite residual (MG numerical explode)");
  if (c->cfg.verbose)
    fprintf(stderr, "[otm] PCG its=%d relres=(%.3e, %.3e, %.3e)%s\n", it, c->relres[0], c->relres[1], c->relres[2],
            c->bricks_on ? "" : " (exact GS)");
  bool met = true;
  for (int k = 0; k < 3; k++) met = met && relres[k] <= c->cfg.fem_rel_thres;
  return met;
}

static std::string noconv_message(const otm_ctx* c) {
  char buf[200];
  snprintf(buf, sizeof(buf), "solve ended above the relative residual target %.3g: (%.3g, %.3g, %.3g) after %d PCG iterations",
           c->cfg.fem_rel_thres, c->relres[0], c->relres[1], c->relres[2],
           std::max(c->iters[0], std::max(c->iters[1], c->iters[2])));
  return buf;
}

static void kappa_h(otm_ctx* c, double C[9]) {
  const unsigned nb = red_blocks_for(c->g0.nn);
  if (!c->energy && hipMalloc((void**)&c->energy, sizeof(float) * 6 * (size_t)c->g0.nn) != hipSuccess) {
    (void)hipGetLastError();  // no room: sensitivities recompute the energies
    c->energy = nullptr;
  }
  {
    ProfScope ps(c, OTM_PROF_ENERGY, 16.0 * (double)c->g0.nn);
    launch_fine_kappa_h(c->s, c->g0, c->fc, c->kappa, c->chi_ref(), c->reducer_host(FIN_STORE), nb, c->energy);
  }
  c->energy_ok = c->energy != nullptr;
  c->wait_host();
  const double* s = c->hm->out;
  // order (00,01,02,11,12,22); |Y| = 1 (homogenization.cu:207)
  C[0] = s[0]; C[1] = s[1]; C[2] = s[2];
  C[3] = s[1]; C[4] = s[3]; C[5] = s[4];
  C[6] = s[2]; C[7] = s[4]; C[8] = s[5];
}

static void ensure_design_buffers(otm_ctx* c) {
  if (c->rho) return;
  const long long n = c->g0.nn;
  c->rho = dalloc<float>(n);
  c->rho_f = dalloc<float>(n);
  c->W = dalloc<float>(n);
  c->kraw = dalloc<float>(n);
  launch_fill(c->s, c->kraw, n, 0.f);
  c->dkappa = dalloc<float>(n);
  c->tmp = dalloc<float>(n);
  c->drho = dalloc<float>(n);
}

static void sens_weights(const double dC[9], float w[6]) {
  const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
  for (int t = 0; t < 6; t++) {
    const int i = ii[t], j = jj[t];
    w[t] = (float)(dC[i * 3 + j] + (i != j ? dC[j * 3 + i] : 0.0));
  }
}

// ---- C ABI -----------------------------------------------------------------------
extern "C" {

OTM_API void otm_config_init(otm_config* cfg) {
  if (!cfg) return;
  memset(cfg, 0, sizeof(*cfg));
  cfg->reso[0] = cfg->reso[1] = cfg->reso[2] = 128;  // cmdline.cpp:6-8
  cfg->heat_ratio[0] = 1.0;                           // cmdline.cpp:12
  cfg->heat_ratio[1] = 0.0;                           // cmdline.cpp:13
  cfg->fem_rel_thres = 1e-6;
  cfg->max_cg_iter = 200;  // the reference caps its V-cycle loop at 200 (MG.cpp:91)
  cfg->pre_sweeps = 2;
  cfg->post_sweeps = 2;
  cfg->min_coarse = 4;
  cfg->device = 0;
  cfg->verbose = 0;
  cfg->coarse_sweeps = 0;
  cfg->fp64_refine = 0;
  cfg->fp16_stencils = 0;
  cfg->exact_smoother = 0;
  cfg->graph_vcycle = 1;
  cfg->fine_tile = -1;
}

OTM_API const char* otm_last_error(void) { return g_last_error.c_str(); }
OTM_API int otm_version(void) { return 100; }

OTM_API int otm_create(const otm_config* cfg, otm_ctx** out) {
  return guarded([&] {
    if (!cfg || !out) fail(OTM_E_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
      (void)hipGetLastError();
      fail(OTM_E_NODEV, "no HIP device available");
    }
    if (cfg->device < 0 || cfg->device >= ndev) fail(OTM_E_ARG, "device ordinal out of range");
    if (cfg->pre_sweeps < 0 || cfg->post_sweeps < 0 || cfg->pre_sweeps + cfg->post_sweeps == 0)
      fail(OTM_E_ARG, "need at least one smoothing sweep");
    if (cfg->pre_sweeps != cfg->post_sweeps)
      fail(OTM_E_ARG, "pre and post sweeps must match (symmetric preconditioner for CG)");
    if (cfg->coarse_sweeps < 0) fail(OTM_E_ARG, "coarse_sweeps must be >= 0");
    if (cfg->fine_tile < -1 || cfg->fine_tile >= kFineTileVariants) fail(OTM_E_ARG, "fine_tile must be -1..4");
    if (!(cfg->fem_rel_thres > 0)) fail(OTM_E_ARG, "fem_rel_thres must be positive");
    otm_ctx* c = new otm_ctx();
    c->cfg = *cfg;
    c->device = cfg->device;
    try {
      hip_check(hipSetDevice(c->device), "hipSetDevice");
      // a blocking stream: work a caller queues on the default stream (a
      // hipMemset of an output, say) is ordered with the context's work, as
      // in the reference, which runs everything on the default stream
      hip_check(hipStreamCreateWithFlags(&c->s, hipStreamDefault), "hipStreamCreate");
      build_hierarchy(c);
      if (cfg->exact_smoother) set_bricks(c, false);
      c->sync();
    } catch (...) {
      ctx_free(c);
      throw;
    }
    *out = c;
  });
}

OTM_API int otm_destroy(otm_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    ctx_free(ctx);
  });
}

OTM_API int otm_levels(const otm_ctx* ctx, int* nlevels) {
  return guarded([&] {
    if (!ctx || !nlevels) fail(OTM_E_ARG, "null argument");
    *nlevels = (int)ctx->lv.size();
  });
}

OTM_API int otm_level_reso(const otm_ctx* ctx, int l, int reso[3]) {
  return guarded([&] {
    if (!ctx || !reso) fail(OTM_E_ARG, "null argument");
    if (l < 0 || l >= (int)ctx->lv.size()) fail(OTM_E_ARG, "level out of range");
    reso[0] = ctx->lv[l].g.nx;
    reso[1] = ctx->lv[l].g.ny;
    reso[2] = ctx->lv[l].g.nz;
  });
}

// overlap_start: launch the next solve's FP64 start residual (it needs only
// kappa and the warm-start iterate) on the side stream, concurrently with
// the coarse operator rebuild, whose coarsest inverse occupies one CU; the
// main stream waits for it after the rebuild. Only for a caller that solves
// next with nothing in between (otm_eval), on the plain path.
static void do_update(otm_ctx* c, const float* kappa, int pitch, int where, bool overlap_start = false) {
  if (kappa) {
    check_where(where);
    if (pitch < c->g0.nx) fail(OTM_E_ARG, "pitch smaller than reso[0]");
    const long long rows = (long long)c->g0.ny * c->g0.nz;
    const float* src = kappa;
    float* staging = nullptr;
    if (where == OTM_HOST) {
      staging = c->q;  // 3n floats of scratch; the PCG rewrites q before use
      if ((long long)pitch * rows > 3 * c->g0.nn) {
        // pitched host rows wider than the scratch: pack on the host first
        std::vector<float> packed((size_t)c->g0.nn);
        for (long long r = 0; r < rows; r++)
          memcpy(&packed[(size_t)(r * c->g0.nx)], kappa + r * pitch, sizeof(float) * c->g0.nx);
        hip_check(hipMemcpyAsync(staging, packed.data(), sizeof(float) * c->g0.nn, hipMemcpyHostToDevice, c->s),
                  "upload kappa");
        c->sync();
        pitch = c->g0.nx;
      } else {
        hip_check(hipMemcpyAsync(staging, kappa, sizeof(float) * pitch * rows, hipMemcpyHostToDevice, c->s),
                  "upload kappa");
      }
      src = staging;
    }
    // kappa_eff = rhoPenalMin + kappa (grid.h:429, grid.cu:405)
    launch_copy_pitched(c->s, c->g0, src, pitch, c->kappa, 1e-9f);
    c->has_kappa = true;
  }
  if (!c->has_kappa) fail(OTM_E_STATE, "otm_update without a conductivity field");
  c->start_res_ready = false;
  build_operators(c, overlap_start && c->x64 == nullptr);
}

OTM_API int otm_update(otm_ctx* ctx, const float* kappa, int pitch, int where) {
  return guarded([&] {
    if (!ctx) fail(OTM_E_ARG, "null context");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    do_update(ctx, kappa, pitch, where);
    ctx->sync();
  });
}

// Solve + kappa^H. Returns false when the solve ended above its relative
// residual target; C, the iterate and the statistics are published anyway
// (the reference prints "Failed to converge" and carries on, MG.cpp:101-118).
static bool do_heat_matrix(otm_ctx* c, double C[9]) {
  if (!c->has_kappa) fail(OTM_E_STATE, "otm_heat_matrix before otm_update");
  const bool met = pcg_solve(c);
  kappa_h(c, C);
  memcpy(c->C, C, sizeof(c->C));
  c->has_solution = true;
  return met;
}

OTM_API int otm_heat_matrix(otm_ctx* ctx, double C[9]) {
  return guarded([&] {
    if (!ctx || !C) fail(OTM_E_ARG, "null argument");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    if (!do_heat_matrix(ctx, C)) fail(OTM_E_NOCONV, noconv_message(ctx));
  });
}

OTM_API int otm_fine_tile_shape(int variant, int shape[4]) {
  return guarded([&] {
    if (!shape || variant < 0 || variant >= kFineTileVariants) fail(OTM_E_ARG, "unknown fine tile variant");
    const FineTileShape& t = kFineTileShapes[variant];
    shape[0] = t.bx;
    shape[1] = t.by;
    shape[2] = t.bz;
    shape[3] = t.bx * t.tr;
  });
}

OTM_API int otm_fine_tile(const otm_ctx* ctx, int* variant) {
  return guarded([&] {
    if (!ctx || !variant) fail(OTM_E_ARG, "null argument");
    *variant = ctx->lv[0].u2 ? ctx->fine_tile : -1;
  });
}

// ---- autotuner ------------------------------------------------------------------
// Times one full three-load-case solve of the current design for every
// candidate (each fine tile shape the cell admits, times coarse sweep counts
// 1 and 2), each from the same start (zero, or with warm the current
// iterate, e.g. the previous design's solution after otm_apply_design of the
// next one) and twice (the first run also
// captures the candidate's V-cycle graphs; the second is timed), and keeps
// the fastest in the context. A candidate whose solve misses the tolerance
// or falls back to exact sweeps is not chosen. The shapes change the block
// Gauss-Seidel and so the iteration counts: the tuner weighs both.
OTM_API int otm_autotune(otm_ctx* c, int warm, otm_tune_report* rep) {
  return guarded([&] {
    if (!c || !rep) fail(OTM_E_ARG, "null argument");
    if (!c->has_kappa) fail(OTM_E_STATE, "otm_autotune before otm_update");
    if (!c->lv[0].u2 || c->fine_tile < 0) fail(OTM_E_UNSUPPORTED, "the fine level has no brick smoother to tune");
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    memset(rep, 0, sizeof(*rep));
    rep->chosen = -1;
    const size_t n3 = 3 * (size_t)c->g0.nn;
    struct Saved {
      float *x = nullptr, *xlo = nullptr;
      double* x64 = nullptr;
      ~Saved() {
        dfree(x);
        dfree(xlo);
        dfree(x64);
      }
    } sv;
    sv.x = dalloc<float>(n3);
    hip_check(hipMemcpyAsync(sv.x, c->x, sizeof(float) * n3, hipMemcpyDeviceToDevice, c->s), "save start");
    if (c->xlo) {
      sv.xlo = dalloc<float>(n3);
      hip_check(hipMemcpyAsync(sv.xlo, c->xlo, sizeof(float) * n3, hipMemcpyDeviceToDevice, c->s), "save start");
    }
    if (c->x64) {
      sv.x64 = dalloc<double>(n3);
      hip_check(hipMemcpyAsync(sv.x64, c->x64, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->s), "save start");
    }
    // candidates are timed from a zero start (a warm start from the current
    // iterate may need no iteration at all); the iterate is restored after
    auto cold = [&] {
      launch_fill(c->s, c->x, (long long)n3, 0.f);
      if (c->xlo) launch_fill(c->s, c->xlo, (long long)n3, 0.f);
      if (c->x64) launch_fill_d(c->s, c->x64, (long long)n3, 0.0);
      c->start_res_ready = false;
      set_bricks(c, true);
      c->sync();
    };
    auto restore = [&] {
      hip_check(hipMemcpyAsync(c->x, sv.x, sizeof(float) * n3, hipMemcpyDeviceToDevice, c->s), "restore start");
      if (c->xlo)
        hip_check(hipMemcpyAsync(c->xlo, sv.xlo, sizeof(float) * n3, hipMemcpyDeviceToDevice, c->s), "restore start");
      if (c->x64)
        hip_check(hipMemcpyAsync(c->x64, sv.x64, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->s), "restore");
      c->start_res_ready = false;
      set_bricks(c, true);
      c->sync();
    };
    const int keep_tile = c->fine_tile, keep_cs = c->cfg.coarse_sweeps;
    double best = 1e300;
    for (int v = 0; v < kFineTileVariants; v++) {
      if (!fine_tile_ok(c->g0, v)) continue;
      for (int cs = 1; cs <= 2 && rep->n < OTM_TUNE_MAX; cs++) {
        const int i = rep->n++;
        rep->fine_tile[i] = v;
        rep->coarse_sweeps[i] = cs;
        c->fine_tile = v;
        c->cfg.coarse_sweeps = cs;
        bool ok = true;
        double ms = 0.0;
        for (int run = 0; run < 2 && ok; run++) {
          if (warm)
            restore();
          else
            cold();
          const int f0 = c->fallbacks;
          const auto t0 = std::chrono::steady_clock::now();
          double C[9];
          try {
            ok = do_heat_matrix(c, C);
          } catch (const OtmError&) {
            ok = false;
          }
          c->sync();
          ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
          ok = ok && c->fallbacks == f0;
        }
        int its = 0;
        for (int k = 0; k < 3; k++) its = std::max(its, c->iters[k]);
        rep->ms[i] = ms;
        rep->iters[i] = its;
        rep->ok[i] = ok ? 1 : 0;
        if (ok && ms < best) {
          best = ms;
          rep->chosen = i;
        }
      }
    }
    if (rep->chosen >= 0) {
      c->fine_tile = rep->fine_tile[rep->chosen];
      c->cfg.coarse_sweeps = rep->coarse_sweeps[rep->chosen];
    } else {
      c->fine_tile = keep_tile;
      c->cfg.coarse_sweeps = keep_cs;
    }
    restore();
    c->has_solution = false;
    c->energy_ok = false;
    if (rep->chosen < 0) fail(OTM_E_NOCONV, "no tuning candidate reached the tolerance");
  });
}

OTM_API int otm_solve_stats(const otm_ctx* ctx, int iters[3], double relres[3]) {
  return guarded([&] {
    if (!ctx) fail(OTM_E_ARG, "null context");
    for (int k = 0; k < 3; k++) {
      if (iters) iters[k] = ctx->iters[k];
      if (relres) relres[k] = ctx->relres[k];
    }
  });
}

// Contracted sensitivity of the current solve: from the element energies
// the kappa^H kernel stored when they are still those of this iterate, else
// recomputed from the fluctuation fields (the same bits either way).
static void launch_sensitivity(otm_ctx* c, const float w[6], float* sens, int pitch, const float* scale = nullptr) {
  if (c->energy_ok)
    launch_sens_from_energy(c->s, c->g0, c->energy, w, sens, pitch, scale);
  else
    launch_fine_sensitivity(c->s, c->g0, c->fc, c->chi_ref(), w, sens, pitch, scale);
}

OTM_API int otm_sensitivity(otm_ctx* ctx, const double dC[9], float* sens, int pitch, int where) {
  return guarded([&] {
    if (!ctx || !dC || !sens) fail(OTM_E_ARG, "null argument");
    check_where(where);
    if (!ctx->has_solution) fail(OTM_E_STATE, "otm_sensitivity before otm_heat_matrix");
    if (pitch < ctx->g0.nx) fail(OTM_E_ARG, "pitch smaller than reso[0]");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    float w[6];
    sens_weights(dC, w);
    const long long rows = (long long)ctx->g0.ny * ctx->g0.nz;
    if (where == OTM_DEVICE) {
      // the reference zero-fills the pitched destination first (homogenization.cu:427)
      if (pitch != ctx->g0.nx)
        hip_check(hipMemsetAsync(sens, 0, sizeof(float) * pitch * rows, ctx->s), "zero sens");
      launch_sensitivity(ctx, w, sens, pitch);
      ctx->sync();
    } else {
      launch_sensitivity(ctx, w, ctx->q, ctx->g0.nx);
      std::vector<float> h((size_t)ctx->g0.nn);
      hip_check(hipMemcpyAsync(h.data(), ctx->q, sizeof(float) * ctx->g0.nn, hipMemcpyDeviceToHost, ctx->s),
                "download sens");
      ctx->sync();
      for (long long r = 0; r < rows; r++) {
        memcpy(sens + r * pitch, &h[(size_t)(r * ctx->g0.nx)], sizeof(float) * ctx->g0.nx);
        for (int x = ctx->g0.nx; x < pitch; x++) sens[r * pitch + x] = 0.f;
      }
    }
  });
}

OTM_API int otm_element_energy(otm_ctx* ctx, float* S6, int where) {
  return guarded([&] {
    if (!ctx || !S6) fail(OTM_E_ARG, "null argument");
    check_where(where);
    if (!ctx->has_solution) fail(OTM_E_STATE, "otm_element_energy before otm_heat_matrix");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const long long n = ctx->g0.nn;
    if (ctx->energy_ok) {
      hip_check(hipMemcpyAsync(S6, ctx->energy, sizeof(float) * 6 * n, kind_of(where, OTM_DEVICE), ctx->s), "copy S6");
      ctx->sync();
    } else if (where == OTM_DEVICE) {
      launch_fine_energy6(ctx->s, ctx->g0, ctx->fc, ctx->chi_ref(), S6);
      ctx->sync();
    } else {
      float* d = dalloc<float>(6 * n);
      launch_fine_energy6(ctx->s, ctx->g0, ctx->fc, ctx->chi_ref(), d);
      hipError_t e = hipMemcpyAsync(S6, d, sizeof(float) * 6 * n, hipMemcpyDeviceToHost, ctx->s);
      hipError_t e2 = hipStreamSynchronize(ctx->s);
      dfree(d);
      hip_check(e, "download S6");
      hip_check(e2, "sync");
    }
  });
}

OTM_API int otm_get_chi(otm_ctx* ctx, float* chi, int where) {
  return guarded([&] {
    if (!ctx || !chi) fail(OTM_E_ARG, "null argument");
    check_where(where);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    hip_check(hipMemcpyAsync(chi, ctx->x, sizeof(float) * 3 * ctx->g0.nn, kind_of(where, OTM_DEVICE), ctx->s),
              "copy chi");
    ctx->sync();
  });
}

OTM_API int otm_set_chi(otm_ctx* ctx, const float* chi, int where) {
  return guarded([&] {
    if (!ctx || !chi) fail(OTM_E_ARG, "null argument");
    check_where(where);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    ctx->energy_ok = false;
    hip_check(hipMemcpyAsync(ctx->x, chi, sizeof(float) * 3 * ctx->g0.nn, kind_of(OTM_DEVICE, where), ctx->s),
              "copy chi");
    if (ctx->x64) launch_f2d(ctx->s, 3 * ctx->g0.nn, ctx->x, ctx->x64);
    if (ctx->xlo) launch_fill(ctx->s, ctx->xlo, 3 * ctx->g0.nn, 0.f);
    ctx->sync();
  });
}

OTM_API int otm_reset_displacement(otm_ctx* ctx) {
  return guarded([&] {
    if (!ctx) fail(OTM_E_ARG, "null context");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    ctx->energy_ok = false;
    launch_fill(ctx->s, ctx->x, 3 * ctx->g0.nn, 0.f);
    if (ctx->x64) launch_fill_d(ctx->s, ctx->x64, 3 * ctx->g0.nn, 0.0);
    if (ctx->xlo) launch_fill(ctx->s, ctx->xlo, 3 * ctx->g0.nn, 0.f);
    ctx->sync();
  });
}

OTM_API int otm_apply_level(otm_ctx* ctx, int level, const float* u, float* q, int where) {
  return guarded([&] {
    if (!ctx || !u || !q) fail(OTM_E_ARG, "null argument");
    check_where(where);
    if (level < 0 || level >= (int)ctx->lv.size()) fail(OTM_E_ARG, "level out of range");
    if (!ctx->has_kappa) fail(OTM_E_STATE, "otm_apply_level before otm_update");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const Level& L = ctx->lv[level];
    const long long n = L.g.nn;
    float* du = dalloc<float>(3 * n);
    float* dq = dalloc<float>(3 * n);
    float* dz = dalloc<float>(3 * n);
    try {
      hip_check(hipMemcpyAsync(du, u, sizeof(float) * 3 * n, kind_of(OTM_DEVICE, where), ctx->s), "copy u");
      launch_fill(ctx->s, dz, 3 * n, 0.f);
      // q = 0 - (-A u): residual with f = 0 gives -A u
      if (level == 0)
        launch_fine_residual(ctx->s, L.g, ctx->fc, ctx->kappa, dz, du, dq);
      else
        launch_st_residual(ctx->s, L.g, stref(L), dz, du, dq);
      std::vector<float> h((size_t)(3 * n));
      hip_check(hipMemcpyAsync(h.data(), dq, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, ctx->s), "copy q");
      ctx->sync();
      for (auto& v : h) v = -v;
      hip_check(hipMemcpy(q, h.data(), sizeof(float) * 3 * n, kind_of(where, OTM_HOST)), "copy q out");
    } catch (...) {
      dfree(du); dfree(dq); dfree(dz);
      throw;
    }
    dfree(du); dfree(dq); dfree(dz);
  });
}

OTM_API int otm_apply_fine_pcg(otm_ctx* ctx, const float* u, float* q, double pq[3], int where) {
  return guarded([&] {
    if (!ctx || !u || !q) fail(OTM_E_ARG, "null argument");
    check_where(where);
    if (!ctx->has_kappa) fail(OTM_E_STATE, "otm_apply_fine_pcg before otm_update");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const long long n = ctx->g0.nn;
    float* du = dalloc<float>(3 * n);
    float* dq = dalloc<float>(3 * n);
    try {
      hip_check(hipMemcpyAsync(du, u, sizeof(float) * 3 * n, kind_of(OTM_DEVICE, where), ctx->s), "copy u");
      launch_fine_apply_dot(ctx->s, ctx->g0, ctx->fc, ctx->kappa, du, dq, ctx->reducer(FIN_STORE), red_blocks_for(n));
      hip_check(hipMemcpyAsync(q, dq, sizeof(float) * 3 * n, kind_of(where, OTM_DEVICE), ctx->s), "copy q");
      double out[OTM_NRHS];
      hip_check(hipMemcpyAsync(out, ctx->red_out, sizeof(out), hipMemcpyDeviceToHost, ctx->s), "copy u.q");
      ctx->sync();
      if (pq)
        for (int k = 0; k < OTM_NRHS; k++) pq[k] = out[k];
    } catch (...) {
      dfree(du); dfree(dq);
      throw;
    }
    dfree(du); dfree(dq);
  });
}

OTM_API int otm_vcycle(otm_ctx* ctx, const float* r, float* z, int where) {
  return guarded([&] {
    if (!ctx || !r || !z) fail(OTM_E_ARG, "null argument");
    check_where(where);
    if (!ctx->has_kappa) fail(OTM_E_STATE, "otm_vcycle before otm_update");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const long long n = ctx->g0.nn;
    hip_check(hipMemcpyAsync(ctx->r, r, sizeof(float) * 3 * n, kind_of(OTM_DEVICE, where), ctx->s), "copy r");
    ctx->post_dot_mode = -1;
    run_vcycle(ctx);
    hip_check(hipMemcpyAsync(z, ctx->z, sizeof(float) * 3 * n, kind_of(where, OTM_DEVICE), ctx->s), "copy z");
    ctx->sync();
  });
}

// ---- MG_H entry points and level operators ---------------------------------------
static void stationary_cycle(otm_ctx* c, double relres[3]) {
  c->energy_ok = false;
  const long long n = c->g0.nn;
  const unsigned nb = red_blocks_for(n);
  const Reduce R = c->reducer_host(FIN_BB_RR, 0.0, 1.0 / (double)n);
  if (c->x64)
    launch_fine_init_residual(c->s, c->g0, c->fc, c->kappa, c->x64, c->r, R, nb);
  else
    launch_fine_init_residual(c->s, c->g0, c->fc, c->kappa, c->x, c->xlo, c->r, R, nb);
  c->read_pcg();
  if (c->pcg_host->fault & 1) fail(OTM_E_NUMERIC, "non-finite residual (MG numerical explode)");
  for (int k = 0; k < 3; k++) relres[k] = sqrt(c->pcg_host->rr[k] / c->pcg_host->bb[k]);
  launch_sub_mean3(c->s, n, c->r, c->pcg);
  c->post_dot_mode = -1;
  run_vcycle(c);
  if (c->x64)
    launch_refine_add(c->s, 3 * n, c->x64, c->z);
  else
    launch_add_pair(c->s, 3 * n, c->x, c->xlo, c->z);
}

OTM_API int otm_solve(otm_ctx* ctx, double rtol, int warm, double relres[3]) {
  return guarded([&] {
    if (!ctx) fail(OTM_E_ARG, "null context");
    if (!ctx->has_kappa) fail(OTM_E_STATE, "otm_solve before otm_update");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const long long n = ctx->g0.nn;
    if (!warm) {
      launch_fill(ctx->s, ctx->x, 3 * n, 0.f);
      if (ctx->x64) launch_fill_d(ctx->s, ctx->x64, 3 * n, 0.0);
      if (ctx->xlo) launch_fill(ctx->s, ctx->xlo, 3 * n, 0.f);
    }
    const double keep = ctx->cfg.fem_rel_thres;
    if (rtol > 0) ctx->cfg.fem_rel_thres = rtol;
    bool met = false;
    std::string msg;
    try {
      met = pcg_solve(ctx);
      if (!met) msg = noconv_message(ctx);
    } catch (...) {
      ctx->cfg.fem_rel_thres = keep;
      throw;
    }
    ctx->cfg.fem_rel_thres = keep;
    ctx->sync();
    if (relres)
      for (int k = 0; k < 3; k++) relres[k] = ctx->relres[k];
    if (!met) fail(OTM_E_NOCONV, msg);
  });
}

OTM_API int otm_mg_cycles(otm_ctx* ctx, int cycles, double relres[3]) {
  return guarded([&] {
    if (!ctx || cycles < 0) fail(OTM_E_ARG, "invalid argument");
    if (!ctx->has_kappa) fail(OTM_E_STATE, "otm_mg_cycles before otm_update");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    double rr[3] = {0, 0, 0};
    for (int i = 0; i < cycles; i++) stationary_cycle(ctx, rr);
    remove_constant(ctx);
    ctx->sync();
    if (relres)
      for (int k = 0; k < 3; k++) relres[k] = rr[k];
  });
}

// Host or device arrays of `count` floats staged into a fresh device buffer.
struct Staged {
  float* d = nullptr;
  ~Staged() { dfree(d); }
};
static float* stage_in(otm_ctx* c, Staged& st, const float* src, long long count, int where) {
  st.d = dalloc<float>(count);
  hip_check(hipMemcpyAsync(st.d, src, sizeof(float) * count, kind_of(OTM_DEVICE, where), c->s), "copy in");
  return st.d;
}
static void stage_out(otm_ctx* c, float* dst, const float* src, long long count, int where) {
  hip_check(hipMemcpyAsync(dst, src, sizeof(float) * count, kind_of(where, OTM_DEVICE), c->s), "copy out");
  c->sync();
}

OTM_API int otm_transfer(otm_ctx* ctx, int level, int direction, const float* in, float* out, int where) {
  return guarded([&] {
    if (!ctx || !in || !out) fail(OTM_E_ARG, "null argument");
    check_where(where);
    if (level < 1 || level >= (int)ctx->lv.size()) fail(OTM_E_ARG, "level must be in 1..levels-1");
    if (direction != 0 && direction != 1) fail(OTM_E_ARG, "direction must be 0 (restrict) or 1 (prolong)");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const Geom& F = ctx->lv[level - 1].g;
    const Geom& C = ctx->lv[level].g;
    Staged a, b;
    if (direction == 0) {
      const float* r = stage_in(ctx, a, in, 3 * F.nn, where);
      b.d = dalloc<float>(3 * C.nn);
      launch_restrict(ctx->s, F, C, r, b.d);
      stage_out(ctx, out, b.d, 3 * C.nn, where);
    } else {
      const float* uc = stage_in(ctx, a, in, 3 * C.nn, where);
      float* uf = stage_in(ctx, b, out, 3 * F.nn, where);
      launch_prolong_add(ctx->s, F, C, uc, uf);
      stage_out(ctx, out, uf, 3 * F.nn, where);
    }
  });
}

OTM_API int otm_smooth_level(otm_ctx* ctx, int level, const float* f, float* u, int sweeps, int mode, int where) {
  return guarded([&] {
    if (!ctx || !f || !u || sweeps < 1) fail(OTM_E_ARG, "invalid argument");
    check_where(where);
    if (level < 0 || level + 1 >= (int)ctx->lv.size()) fail(OTM_E_ARG, "level must be in 0..levels-2");
    if (mode < 0 || mode > 3) fail(OTM_E_ARG, "mode must be 0..3");
    if (!ctx->has_kappa) fail(OTM_E_STATE, "otm_smooth_level before otm_update");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const Level& L = ctx->lv[level];
    const long long n = L.g.nn;
    const bool brick = mode >= 2, forward = (mode & 1) == 0;
    if (brick && !L.u2) fail(OTM_E_UNSUPPORTED, "this level has no brick smoother");
    Staged a, b, c2;
    const float* df = stage_in(ctx, a, f, 3 * n, where);
    float* du = stage_in(ctx, b, u, 3 * n, where);
    if (brick) {
      c2.d = dalloc<float>(3 * n);
      if (level == 0)
        launch_fine_smooth_brick(ctx->s, L.g, ctx->fc, ctx->kappa, df, du, c2.d, sweeps, forward, false,
                                 ctx->fine_tile);
      else
        launch_st_smooth_brick(ctx->s, L.g, stref(L), df, du, c2.d, sweeps, forward, false);
      stage_out(ctx, u, c2.d, 3 * n, where);
    } else {
      const int ncol = level == 0 ? 4 : 8;
      for (int s = 0; s < sweeps; s++)
        for (int i = 0; i < ncol; i++) {
          const int col = forward ? i : ncol - 1 - i;
          if (level == 0)
            launch_fine_gs_color(ctx->s, L.g, ctx->fc, ctx->kappa, df, du, col);
          else
            launch_st_gs_color(ctx->s, L.g, stref(L), df, du, col);
        }
      stage_out(ctx, u, du, 3 * n, where);
    }
  });
}

OTM_API int otm_coarsest_solve(otm_ctx* ctx, const float* f, float* u, int where) {
  return guarded([&] {
    if (!ctx || !f || !u) fail(OTM_E_ARG, "null argument");
    check_where(where);
    if (!ctx->has_kappa) fail(OTM_E_STATE, "otm_coarsest_solve before otm_update");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const Level& L = ctx->lv.back();
    const long long n = L.g.nn;
    Staged a, b;
    const float* df = stage_in(ctx, a, f, 3 * n, where);
    b.d = dalloc<float>(3 * n);
    launch_coarsest_solve(ctx->s, L.g, ctx->cBinv, df, b.d);
    stage_out(ctx, u, b.d, 3 * n, where);
  });
}

OTM_API int otm_get_kappa(otm_ctx* ctx, float* kappa, int where) {
  return guarded([&] {
    if (!ctx || !kappa) fail(OTM_E_ARG, "null argument");
    check_where(where);
    if (!ctx->has_kappa) fail(OTM_E_STATE, "otm_get_kappa before otm_update");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const long long n = ctx->g0.nn;
    Staged a;
    float* dst = kappa;
    if (where == OTM_HOST) dst = a.d = dalloc<float>(n);
    // the context holds kappa + 1e-9 (the floor of grid.cu:405)
    launch_copy_pitched(ctx->s, ctx->g0, ctx->kappa, ctx->g0.nx, dst, -1e-9f);
    if (where == OTM_HOST)
      stage_out(ctx, kappa, dst, n, OTM_HOST);
    else
      ctx->sync();
  });
}

OTM_API int otm_get_stencil(otm_ctx* ctx, int level, float* st, int where) {
  return guarded([&] {
    if (!ctx || !st) fail(OTM_E_ARG, "null argument");
    check_where(where);
    if (level < 1 || level >= (int)ctx->lv.size()) fail(OTM_E_ARG, "stencils exist on levels 1..L-1");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const Level& L = ctx->lv[level];
    hip_check(hipMemcpyAsync(st, L.st, sizeof(float) * 27 * L.g.nn, kind_of(where, OTM_DEVICE), ctx->s),
              "copy stencil");
    ctx->sync();
  });
}

OTM_API int otm_conv_filter(const int reso[3], int kind, double radius, int periodic, int adjoint, const float* in,
                            float* out, int where) {
  return guarded([&] {
    if (!reso || !in || !out) fail(OTM_E_ARG, "null argument");
    check_where(where);
    for (int a = 0; a < 3; a++)
      if (reso[a] <= 0) fail(OTM_E_ARG, "reso must be positive");
    Taps T = make_taps(kind, radius, periodic);
    Geom g{reso[0], reso[1], reso[2], (long long)reso[0] * reso[1] * reso[2]};
    float* din = dalloc<float>(g.nn);
    float* dout = dalloc<float>(g.nn);
    float* dW = dalloc<float>(g.nn);
    float* dtmp = dalloc<float>(g.nn);
    hipStream_t s = nullptr;
    hipError_t e = hipMemcpy(din, in, sizeof(float) * g.nn, kind_of(OTM_DEVICE, where));
    if (e == hipSuccess) {
      if (adjoint) {
        launch_filter_fwd(s, g, T, din, dtmp, dW);  // weight sums W
        launch_filter_adj(s, g, T, din, dW, dout);
      } else {
        launch_filter_fwd(s, g, T, din, dout, dW);
      }
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(float) * g.nn, kind_of(where, OTM_DEVICE));
    dfree(din); dfree(dout); dfree(dW); dfree(dtmp);
    hip_check(e, "conv filter");
  });
}

// ---- the filter without allocation --------------------------------------------
// Scratch of otm_conv_filter_noalloc: the per-node weight sums W of one
// cell, at most kNoAllocNodes floats (32 MiB on every device the library
// runs on, held until the process ends). In host code `scratch` names a
// host-side stand-in: its device address comes from hipGetSymbolAddress,
// once per device with that device current, and only that address is
// handed to kernels.
constexpr long long kNoAllocNodes = 1ll << 23;
static __device__ __align__(256) float scratch[kNoAllocNodes];
constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_dev[kNoAllocMaxDevices];

static float* scratch_of_current_device() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kNoAllocMaxDevices) fail(OTM_E_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) fail(OTM_E_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_dev[dev] = (float*)p;
  }
  return g_scratch_dev[dev];
}

OTM_API int otm_conv_filter_noalloc(const int reso[3], int kind, double radius, int periodic, int adjoint,
                                    const float* in, float* out, void* stream) {
  return guarded([&] {
    if (!reso || !in || !out) fail(OTM_E_ARG, "null argument");
    for (int a = 0; a < 3; a++)
      if (reso[a] <= 0) fail(OTM_E_ARG, "reso must be positive");
    const long long n = (long long)reso[0] * reso[1] * reso[2];
    if (n > kNoAllocNodes)
      fail(OTM_E_UNSUPPORTED, "cell exceeds the no-allocation filter's scratch of 8388608 nodes");
    // the adjoint writes the forward output to out before it reads in again
    if (in < out + n && out < in + n) fail(OTM_E_ARG, "in and out overlap");
    float* W = scratch_of_current_device();
    const Taps T = make_taps(kind, radius, periodic);
    const Geom g{reso[0], reso[1], reso[2], n};
    hipStream_t s = (hipStream_t)stream;
    // the forward pass writes every node's weight sum before any read of W
    launch_filter_fwd(s, g, T, in, out, W);
    if (adjoint) launch_filter_adj(s, g, T, in, W, out);
    hip_check(hipGetLastError(), "conv filter");
  });
}

static_assert(kReduceSum == OTM_REDUCE_SUM && kReducePnorm == OTM_REDUCE_PNORM &&
                  kReduceLogSumExp == OTM_REDUCE_LOGSUMEXP,
              "reduction kinds");

OTM_API int otm_field_reduce(const float* x, long long n, int kind, double p, int where, double* value, float* grad,
                             double upstream) {
  return guarded([&] {
    if (!x || !value || n <= 0) fail(OTM_E_ARG, "invalid argument");
    check_where(where);
    if (kind < OTM_REDUCE_SUM || kind > OTM_REDUCE_LOGSUMEXP) fail(OTM_E_ARG, "unknown reduction");
    if (kind == OTM_REDUCE_PNORM && !(p > 0)) fail(OTM_E_ARG, "p must be positive");
    hipStream_t s = nullptr;
    unsigned nb = (unsigned)std::min<long long>((n + kBlock - 1) / kBlock, 1024);
    struct Bufs {
      float *x = nullptr, *g = nullptr;
      double* part = nullptr;
      ~Bufs() {
        dfree(x);
        dfree(g);
        dfree(part);
      }
    } b;
    const float* dx = x;
    if (where == OTM_HOST) {
      b.x = dalloc<float>(n);
      hip_check(hipMemcpy(b.x, x, sizeof(float) * n, hipMemcpyHostToDevice), "copy in");
      dx = b.x;
    }
    b.part = dalloc<double>(nb);
    std::vector<double> part(nb);
    double shift = 0.0;
    if (kind == OTM_REDUCE_LOGSUMEXP) {  // exp(x - max): the same value, no overflow
      launch_field_max(s, n, dx, b.part, nb);
      hip_check(hipMemcpy(part.data(), b.part, sizeof(double) * nb, hipMemcpyDeviceToHost), "max");
      shift = -INFINITY;
      for (double m : part) shift = std::max(shift, m);
    }
    launch_field_sum(s, kind, n, dx, p, shift, b.part, nb);
    hip_check(hipMemcpy(part.data(), b.part, sizeof(double) * nb, hipMemcpyDeviceToHost), "sum");
    double sum = 0.0;
    for (double v : part) sum += v;
    const double val = kind == OTM_REDUCE_SUM ? sum : (kind == OTM_REDUCE_PNORM ? pow(sum, 1.0 / p) : shift + log(sum));
    *value = val;
    if (grad) {
      float* dg = grad;
      if (where == OTM_HOST) dg = b.g = dalloc<float>(n);
      launch_field_grad(s, kind, n, dx, p, val, upstream, dg);
      hip_check(hipGetLastError(), "field gradient");
      if (where == OTM_HOST)
        hip_check(hipMemcpy(grad, dg, sizeof(float) * n, hipMemcpyDeviceToHost), "copy out");
      else
        hip_check(hipStreamSynchronize(s), "sync");
    }
  });
}

// Symmetry groups of symmetrizeField (Framework.cuh:221-235) on the cell's
// centre: reflect3 = the 8 axis reflections, reflect6 = reflections x axis
// permutations (48), rotate3 = the 24 rotations of the cube (permutation
// parity = parity of the number of reflected axes), rotate2 = the point
// reflection (2).
static SymGroup make_sym_group(int sym, const Geom& g) {
  SymGroup G;
  memset(&G, 0, sizeof(G));
  const int perms[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {2, 1, 0}, {1, 0, 2}};  // 3 even, 3 odd
  auto add = [&](int p, int f) {
    for (int j = 0; j < 3; j++) G.perm[G.n][j] = (signed char)perms[p][j];
    G.flip[G.n] = (unsigned char)f;
    G.n++;
  };
  const bool cubic = g.nx == g.ny && g.ny == g.nz;
  switch (sym) {
    case OTM_SYM_REFLECT3:
      for (int f = 0; f < 8; f++) add(0, f);
      break;
    case OTM_SYM_REFLECT6:
      if (!cubic) fail(OTM_E_ARG, "reflect6 symmetry needs a cubic cell");
      for (int p = 0; p < 6; p++)
        for (int f = 0; f < 8; f++) add(p, f);
      break;
    case OTM_SYM_ROTATE3:
      if (!cubic) fail(OTM_E_ARG, "rotate3 symmetry needs a cubic cell");
      for (int p = 0; p < 6; p++)
        for (int f = 0; f < 8; f++)
          if ((__builtin_popcount(f) & 1) == (p >= 3 ? 1 : 0)) add(p, f);
      break;
    case OTM_SYM_ROTATE2:
      add(0, 0);
      add(0, 7);
      break;
    default:
      fail(OTM_E_ARG, "unknown symmetry");
  }
  return G;
}

OTM_API int otm_symmetrize(const int reso[3], int sym, float* field, int where) {
  return guarded([&] {
    if (!reso || !field) fail(OTM_E_ARG, "null argument");
    check_where(where);
    for (int a = 0; a < 3; a++)
      if (reso[a] <= 0) fail(OTM_E_ARG, "reso must be positive");
    if (sym == OTM_SYM_NONE) return;
    Geom g{reso[0], reso[1], reso[2], (long long)reso[0] * reso[1] * reso[2]};
    const SymGroup G = make_sym_group(sym, g);
    float* din = dalloc<float>(g.nn);
    float* dout = dalloc<float>(g.nn);
    hipError_t e = hipMemcpy(din, field, sizeof(float) * g.nn, kind_of(OTM_DEVICE, where));
    if (e == hipSuccess) {
      launch_symmetrize(nullptr, g, G, din, dout);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(field, dout, sizeof(float) * g.nn, kind_of(where, OTM_DEVICE));
    dfree(din);
    dfree(dout);
    hip_check(e, "symmetrize");
  });
}

OTM_API int otm_symmetrize_design(otm_ctx* ctx, int sym) {
  return guarded([&] {
    if (!ctx) fail(OTM_E_ARG, "null context");
    if (!ctx->rho) fail(OTM_E_STATE, "no design density set");
    if (sym == OTM_SYM_NONE) return;
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const SymGroup G = make_sym_group(sym, ctx->g0);
    const long long n = ctx->g0.nn;
    launch_symmetrize(ctx->s, ctx->g0, G, ctx->rho, ctx->tmp);
    hip_check(hipMemcpyAsync(ctx->rho, ctx->tmp, sizeof(float) * n, hipMemcpyDeviceToDevice, ctx->s), "copy");
    if (ctx->has_drho) {
      launch_symmetrize(ctx->s, ctx->g0, G, ctx->drho, ctx->tmp);
      hip_check(hipMemcpyAsync(ctx->drho, ctx->tmp, sizeof(float) * n, hipMemcpyDeviceToDevice, ctx->s), "copy");
    }
    ctx->sync();
  });
}

OTM_API void otm_density_params_init(otm_density_params* p) {
  if (!p) return;
  p->filter_kind = 0;       // Spline4 (Framework.cu:163)
  p->filter_radius = 1.5;
  p->filter_periodic = 0;   // TensorExpression.h:1083
  p->penal = 3.0;
}

OTM_API int otm_set_design(otm_ctx* ctx, const float* rho, int where) {
  return guarded([&] {
    if (!ctx || !rho) fail(OTM_E_ARG, "null argument");
    check_where(where);
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    ensure_design_buffers(ctx);
    hip_check(hipMemcpyAsync(ctx->rho, rho, sizeof(float) * ctx->g0.nn, kind_of(OTM_DEVICE, where), ctx->s),
              "copy rho");
    ctx->sync();
    ctx->has_drho = false;
  });
}

OTM_API int otm_get_design(otm_ctx* ctx, float* rho, int where) {
  return guarded([&] {
    if (!ctx || !rho) fail(OTM_E_ARG, "null argument");
    check_where(where);
    if (!ctx->rho) fail(OTM_E_STATE, "no design density set");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    hip_check(hipMemcpyAsync(rho, ctx->rho, sizeof(float) * ctx->g0.nn, kind_of(where, OTM_DEVICE), ctx->s),
              "copy rho");
    ctx->sync();
  });
}

// symmetrizeField (Framework.cuh:221-235) of the context's design in place.
static void symmetrize_rho(otm_ctx* c, int sym) {
  if (sym == OTM_SYM_NONE) return;
  const SymGroup G = make_sym_group(sym, c->g0);
  launch_symmetrize(c->s, c->g0, G, c->rho, c->tmp);
  hip_check(hipMemcpyAsync(c->rho, c->tmp, sizeof(float) * c->g0.nn, hipMemcpyDeviceToDevice, c->s), "copy");
}

static double mean_on_device(otm_ctx* c, const float* v) {
  const long long n = c->g0.nn;
  launch_sum2(c->s, n, v, v, c->reducer(FIN_STORE), red_blocks_for(n));
  hip_check(hipMemcpyAsync(c->out_host, c->red_out, sizeof(double), hipMemcpyDeviceToHost, c->s), "read sum");
  c->sync();
  return c->out_host[0] / (double)n;
}

// randTri (Framework.cuh:238-310): coefficients uniform in [-1, 1] (the
// reference draws them with cuRAND), the basis sum per node, symmetrisation,
// then the level-set value t of sigmoid(., 15, t) bisected (at most 60
// steps, relative volume error 1e-3) so that the mapped field has mean
// volRatio, and the map itself (values in [0.001, min(1.5 volRatio, 1)]).
static void rand_tri(otm_ctx* c, const otm_init_params& p, std::vector<float>& coeffs) {
  const int P = p.n_period;
  if (P < 1 || P > rand_tri_max_period()) fail(OTM_E_UNSUPPORTED, "randTri period must be 1..10");
  const bool centered = p.initway == 1 || p.initway == 5;
  const int nb = 6 * P + 36 * P * P + (centered ? 4 : 0);
  if (p.coeffs) {
    if (p.ncoeffs != nb) fail(OTM_E_ARG, "randTri needs " + std::to_string(nb) + " coefficients");
    coeffs.assign(p.coeffs, p.coeffs + nb);
  } else {
    if (p.initway == 5) fail(OTM_E_ARG, "rep_randcenter needs the coefficients of an earlier randcenter start");
    std::mt19937_64 gen((uint64_t)p.seed * 0x9E3779B97F4A7C15ull + 0x5851F42D4C957F2Dull);
    std::uniform_real_distribution<float> ud(-1.f, 1.f);
    coeffs.resize(nb);
    for (auto& v : coeffs) v = ud(gen);
  }
  RandTriRot R;
  memset(&R, 0, sizeof(R));
  R.m[0] = R.m[4] = R.m[8] = 1.f;
  if (centered) {
    // glm::toMat3(normalize(quat(w, x, y, z))) from the last four coefficients
    const int q0 = 6 * P + 36 * P * P;
    double w = coeffs[q0], x = coeffs[q0 + 1], y = coeffs[q0 + 2], z = coeffs[q0 + 3];
    const double nrm = sqrt(w * w + x * x + y * y + z * z);
    if (!(nrm > 0)) fail(OTM_E_ARG, "randcenter rotation quaternion is zero");
    w /= nrm; x /= nrm; y /= nrm; z /= nrm;
    const double m[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                         2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                         2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)};
    for (int i = 0; i < 9; i++) R.m[i] = (float)m[i];
  }
  const long long n = c->g0.nn;
  float* dco = dalloc<float>(nb);
  double* part = nullptr;
  try {
    hip_check(hipMemcpyAsync(dco, coeffs.data(), sizeof(float) * nb, hipMemcpyHostToDevice, c->s), "upload coeffs");
    launch_rand_tri(c->s, c->g0, P, centered ? 1 : 0, R, dco, c->rho);
    hip_check(hipGetLastError(), "randTri");
    symmetrize_rho(c, p.sym);
    const float V = (float)p.volratio;
    const float hi = fminf(V * 1.5f, 1.f);
    const unsigned nbk = level_sum_blocks(n);
    part = dalloc<double>(nbk);
    std::vector<double> hpart(nbk);
    float tu = (float)P * (float)P * 37.f, tl = -tu, t = 0.5f * (tu + tl);
    for (int it = 0; it < 60; it++) {
      t = (tu + tl) / 2;
      launch_level_sum(c->s, n, c->rho, t, hi, part);
      hip_check(hipMemcpyAsync(hpart.data(), part, sizeof(double) * nbk, hipMemcpyDeviceToHost, c->s), "sums");
      c->sync();
      double sum = 0.0;
      for (double v : hpart) sum += v;
      const float sm = (float)(sum / (double)n);
      const float rel = fabsf(sm - V) / V;
      if (rel < 1e-3f) break;
      if (sm > V) tl = t;
      if (sm < V) tu = t;
    }
    launch_level_apply(c->s, n, c->rho, t, hi);
    hip_check(hipGetLastError(), "randTri map");
    c->sync();
  } catch (...) {
    dfree(dco);
    dfree(part);
    throw;
  }
  dfree(dco);
  dfree(part);
}

OTM_API void otm_init_params_init(otm_init_params* p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->initway = 9;      // IWP (cmdline.cpp:10)
  p->volratio = 0.3;   // cmdline.cpp:11
  p->n_period = 10;    // HomoConfig::initperiod (cmdline.cpp)
  p->sym = OTM_SYM_NONE;
}

OTM_API int otm_init_design(otm_ctx* ctx, const otm_init_params* p, float* coeffs_out) {
  return guarded([&] {
    if (!ctx || !p) fail(OTM_E_ARG, "null argument");
    const int iw = p->initway;
    if (iw == 3 || iw == 4) fail(OTM_E_UNSUPPORTED, "OpenVDB initial densities are not supported");
    if (iw < 0 || iw > 10) fail(OTM_E_ARG, "unknown initialisation");
    if (!(p->volratio > 0 && p->volratio <= 1)) fail(OTM_E_ARG, "volratio must be in (0, 1]");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    ensure_design_buffers(ctx);
    const long long n = ctx->g0.nn;
    if (iw == 0 || iw == 1 || iw == 5) {
      std::vector<float> coeffs;
      rand_tri(ctx, *p, coeffs);
      if (coeffs_out) memcpy(coeffs_out, coeffs.data(), sizeof(float) * coeffs.size());
    } else if (iw == 2) {
      // noise (Framework.cu:75-88): uniform values, symmetrised, projected to
      // {1e-9, 1} (proj(20, 0.5)), scaled to the volume ratio, clamped
      launch_noise(ctx->s, n, (unsigned long long)(unsigned)p->seed, ctx->rho);
      symmetrize_rho(ctx, p->sym);
      launch_project_binary(ctx->s, n, ctx->rho);
      const double m = mean_on_device(ctx, ctx->rho);
      const float sc = (float)p->volratio / (float)m;
      launch_scale_clamp(ctx->s, n, ctx->rho, sc, 0.001f, 1.f);
    } else {
      InitModes M;
      memset(&M, 0, sizeof(M));
      launch_init_density(ctx->s, ctx->g0, iw, M, ctx->rho);
    }
    hip_check(hipGetLastError(), "init density");
    // initDensity's last steps (Framework.cu:140-144): symmetrise, clamp to [0.1, 1]
    symmetrize_rho(ctx, p->sym);
    launch_scale_clamp(ctx->s, n, ctx->rho, 1.f, 0.1f, 1.f);
    ctx->sync();
    ctx->has_drho = false;
  });
}

OTM_API int otm_init_density(otm_ctx* ctx, int initway, int seed, double volratio) {
  otm_init_params p;
  otm_init_params_init(&p);
  p.initway = initway;
  p.seed = seed;
  p.volratio = volratio;
  return otm_init_design(ctx, &p, nullptr);
}

OTM_API int otm_design_stats(otm_ctx* ctx, double* mean_rho, double* mean_kappa) {
  return guarded([&] {
    if (!ctx) fail(OTM_E_ARG, "null context");
    if (!ctx->rho) fail(OTM_E_STATE, "no design density set");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const long long n = ctx->g0.nn;
    const unsigned nb = red_blocks_for(n);
    launch_sum2(ctx->s, n, ctx->rho, ctx->kraw, ctx->reducer(FIN_STORE), nb);
    hip_check(hipMemcpyAsync(ctx->out_host, ctx->red_out, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->s),
              "read stats");
    ctx->sync();
    if (mean_rho) *mean_rho = ctx->out_host[0] / (double)n;
    if (mean_kappa) *mean_kappa = ctx->out_host[1] / (double)n;
  });
}

// The density pipeline of otm_eval: filter, SIMP map, operator rebuild.
static void apply_design(otm_ctx* ctx, const otm_density_params* p) {
    if (!ctx->rho) fail(OTM_E_STATE, "otm_eval before otm_set_design");
    if (p->penal != 3.0) fail(OTM_E_UNSUPPORTED, "only the cubic SIMP penalty is supported");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const long long n = ctx->g0.nn;
    const float* src = ctx->rho;
    if (p->filter_kind >= 0) {
      if (ctx->taps_kind != p->filter_kind || ctx->taps_r != p->filter_radius ||
          ctx->taps_periodic != p->filter_periodic) {
        ctx->taps = make_taps(p->filter_kind, p->filter_radius, p->filter_periodic);
        ctx->taps_kind = p->filter_kind;
        ctx->taps_r = p->filter_radius;
        ctx->taps_periodic = p->filter_periodic;
        ctx->w_valid = false;
      }
      launch_filter_fwd(ctx->s, ctx->g0, ctx->taps, ctx->rho, ctx->rho_f, ctx->W, ctx->w_valid);
      ctx->w_valid = true;
      src = ctx->rho_f;
    }
    launch_simp_fwd(ctx->s, n, src, (float)ctx->cfg.heat_ratio[0], (float)ctx->cfg.heat_ratio[1], 0.f, ctx->kraw,
                    ctx->dkappa);
    do_update(ctx, ctx->kraw, ctx->g0.nx, OTM_DEVICE, true);
    ctx->has_drho = false;
    ctx->taps_kind = p->filter_kind >= 0 ? ctx->taps_kind : -1;
}

OTM_API int otm_apply_design(otm_ctx* ctx, const otm_density_params* p) {
  return guarded([&] {
    if (!ctx || !p) fail(OTM_E_ARG, "null argument");
    apply_design(ctx, p);
    ctx->has_solution = false;
    ctx->sync();
  });
}

OTM_API int otm_eval(otm_ctx* ctx, const otm_density_params* p, double C[9]) {
  return guarded([&] {
    if (!ctx || !p || !C) fail(OTM_E_ARG, "null argument");
    apply_design(ctx, p);
    if (!do_heat_matrix(ctx, C)) fail(OTM_E_NOCONV, noconv_message(ctx));
  });
}

OTM_API int otm_backward(otm_ctx* ctx, const double dC[9], float* drho, int where) {
  return guarded([&] {
    if (!ctx || !dC) fail(OTM_E_ARG, "null argument");
    if (!ctx->rho || !ctx->has_solution) fail(OTM_E_STATE, "otm_backward before otm_eval");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const long long n = ctx->g0.nn;
    float w[6];
    sens_weights(dC, w);
    launch_sensitivity(ctx, w, ctx->tmp, ctx->g0.nx, ctx->dkappa);  // times d kappa / d rho
    if (ctx->taps_kind >= 0)
      launch_filter_adj(ctx->s, ctx->g0, ctx->taps, ctx->tmp, ctx->W, ctx->drho);
    else
      hip_check(hipMemcpyAsync(ctx->drho, ctx->tmp, sizeof(float) * n, hipMemcpyDeviceToDevice, ctx->s), "copy");
    if (drho) {
      check_where(where);
      hip_check(hipMemcpyAsync(drho, ctx->drho, sizeof(float) * n, kind_of(where, OTM_DEVICE), ctx->s), "copy drho");
      ctx->sync();
    }
    ctx->has_drho = true;  // on the device, in stream order
  });
}

// ---- OC optimizer ------------------------------------------------------------------
static void oc_free(otm_oc* o) {
  if (!o) return;
  if (o->s) (void)hipStreamSynchronize(o->s);
  dfree(o->newsens); dfree(o->a); dfree(o->b); dfree(o->maxbits);
  dfree(o->st); dfree(o->partials);
  if (o->hm) (void)hipHostFree(o->hm);
  if (o->s) (void)hipStreamDestroy(o->s);
  delete o;
}

OTM_API int otm_oc_create(int ne, double min_rho, double step_limit, double damp, int device, otm_oc** out) {
  return guarded([&] {
    if (!out || ne <= 0) fail(OTM_E_ARG, "invalid argument");
    *out = nullptr;
    otm_oc* o = new otm_oc();
    o->ne = ne;
    o->device = device;
    o->min_rho = min_rho;
    o->step = step_limit;
    o->damp = damp;
    try {
      hip_check(hipSetDevice(device), "hipSetDevice");
      hip_check(hipStreamCreateWithFlags(&o->s, hipStreamDefault), "hipStreamCreate");  // ordered with the default stream
      o->newsens = dalloc<float>(ne);
      o->a = dalloc<float>(ne);
      o->b = dalloc<float>(ne);
      o->maxbits = dalloc<unsigned>(1);
      hip_check(hipMemsetAsync(o->maxbits, 0, sizeof(unsigned), o->s), "hipMemsetAsync");
      o->st = dalloc<OcState>(1);
      o->partials = dalloc<double>((size_t)1024 * kOcCand);
      hip_check(hipHostMalloc((void**)&o->hm, sizeof(OcHost), hipHostMallocCoherent), "hipHostMalloc");
      memset(o->hm, 0, sizeof(OcHost));
      hip_check(hipHostGetDevicePointer((void**)&o->hm_dev, o->hm, 0), "hipHostGetDevicePointer");
      // otm_oc_step runs the optimiser's kernels on a context's stream: the
      // cleared word must be in memory before this call returns
      hip_check(hipStreamSynchronize(o->s), "hipStreamSynchronize");
    } catch (...) {
      oc_free(o);
      throw;
    }
    *out = o;
  });
}

OTM_API int otm_oc_destroy(otm_oc* oc) {
  return guarded([&] { oc_free(oc); });
}

OTM_API int otm_oc_set_bisection(otm_oc* oc, int max_steps, double vol_tol) {
  return guarded([&] {
    if (!oc || max_steps <= 0 || !(vol_tol > 0)) fail(OTM_E_ARG, "invalid argument");
    oc->max_steps = max_steps;
    oc->vol_tol = vol_tol;
  });
}

// Filtered sensitivities into o->newsens (on stream s); copied back over
// sens when copy_back is set (filterSens works in place).
static void oc_filter_dev(otm_oc* o, hipStream_t s, float* sens, const float* rho, const Geom& g, double radius,
                          bool copy_back) {
  Taps T = make_taps(1, radius, 1);  // OCOptimizer.cu:143: Linear, periodic
  launch_oc_filter_sens(s, g, T, sens, rho, o->newsens);
  if (copy_back)
    hip_check(hipMemcpyAsync(sens, o->newsens, sizeof(float) * g.nn, hipMemcpyDeviceToDevice, s), "copy sens");
}

OTM_API int otm_oc_filter_sens(otm_oc* oc, float* sens, const float* rho, int pitch, const int reso[3],
                               double radius, int where) {
  return guarded([&] {
    if (!oc || !sens || !rho || !reso) fail(OTM_E_ARG, "null argument");
    check_where(where);
    Geom g{reso[0], reso[1], reso[2], (long long)reso[0] * reso[1] * reso[2]};
    if (g.nn != oc->ne) fail(OTM_E_ARG, "reso does not match ne");
    if (pitch < reso[0]) fail(OTM_E_ARG, "pitch smaller than reso[0]");
    hip_check(hipSetDevice(oc->device), "hipSetDevice");
    const long long rows = (long long)reso[1] * reso[2];
    // gather to contiguous device copies
    hip_check(hipMemcpy2DAsync(oc->a, sizeof(float) * reso[0], sens, sizeof(float) * pitch, sizeof(float) * reso[0],
                               rows, kind_of(OTM_DEVICE, where), oc->s),
              "gather sens");
    hip_check(hipMemcpy2DAsync(oc->b, sizeof(float) * reso[0], rho, sizeof(float) * pitch, sizeof(float) * reso[0],
                               rows, kind_of(OTM_DEVICE, where), oc->s),
              "gather rho");
    oc_filter_dev(oc, oc->s, oc->a, oc->b, g, radius, true);
    hip_check(hipMemcpy2DAsync(sens, sizeof(float) * pitch, oc->a, sizeof(float) * reso[0], sizeof(float) * reso[0],
                               rows, kind_of(where, OTM_DEVICE), oc->s),
              "scatter sens");
    hip_check(hipStreamSynchronize(oc->s), "sync");
  });
}

// The OC update on stream s. With final_sync the call returns once rho is
// updated; without, rho is updated in stream order on s.
static void oc_update_dev(otm_oc* o, hipStream_t s, const float* sens, float* rho, double volratio, int* steps,
                          double* volume, bool final_sync = true) {
  const long long n = o->ne;
  const unsigned nb = std::min(512u, red_blocks_for(n));  // fewer partials for the 31-way final sum
  Reduce R;
  R.partials = o->partials;
  R.out = nullptr;
  R.pcg = nullptr;
  R.rtol2 = 0;
  R.inv_nn = 0;
  R.mode = FIN_STORE;
  // passes of kOcTree bisection levels are queued without host round trips
  // (a pass after the tolerance is met returns at once), each group followed
  // by an apply that updates rho only once the bisection has ended and
  // publishes the state: one host round trip per group, normally one
  launch_oc_begin(s, n, sens, o->maxbits, o->st);
  const int passes = (o->max_steps + kOcTree - 1) / kOcTree;
  const int per_check = 4;
  for (int done = 0; done < passes;) {
    const int m = std::min(per_check, passes - done);
    for (int i = 0; i < m; i++)
      launch_oc_tree_pass(s, n, sens, rho, o->st, (float)o->min_rho, (float)o->step, (float)o->damp, volratio,
                          o->vol_tol, o->max_steps, R, nb);
    done += m;
    const unsigned tag = ++o->tag;
    launch_oc_apply(s, n, sens, rho, o->st, (float)o->min_rho, (float)o->step, (float)o->damp, o->max_steps,
                    done >= passes ? 1 : 0, o->hm_dev, tag);
    wait_tag(s, &o->hm->seq, tag);
    if (o->hm->applied || o->hm->st.steps == 0) break;
  }
  if (final_sync) hip_check(hipStreamSynchronize(s), "sync");  // rho updated
  if (steps) *steps = o->hm->st.steps;
  if (volume) *volume = o->hm->st.vol;
}

OTM_API int otm_oc_update(otm_oc* oc, const float* sens, float* rho, double volratio, int where, int* steps,
                          double* volume) {
  return guarded([&] {
    if (!oc || !sens || !rho) fail(OTM_E_ARG, "null argument");
    check_where(where);
    hip_check(hipSetDevice(oc->device), "hipSetDevice");
    const long long n = oc->ne;
    if (where == OTM_DEVICE) {
      oc_update_dev(oc, oc->s, sens, rho, volratio, steps, volume);
    } else {
      hip_check(hipMemcpyAsync(oc->a, sens, sizeof(float) * n, hipMemcpyHostToDevice, oc->s), "upload sens");
      hip_check(hipMemcpyAsync(oc->b, rho, sizeof(float) * n, hipMemcpyHostToDevice, oc->s), "upload rho");
      oc_update_dev(oc, oc->s, oc->a, oc->b, volratio, steps, volume);
      hip_check(hipMemcpyAsync(rho, oc->b, sizeof(float) * n, hipMemcpyDeviceToHost, oc->s), "download rho");
      hip_check(hipStreamSynchronize(oc->s), "sync");
    }
  });
}

OTM_API int otm_oc_step(otm_ctx* ctx, otm_oc* oc, double filter_radius, double volratio, int* steps,
                        double* volume) {
  return guarded([&] {
    if (!ctx || !oc) fail(OTM_E_ARG, "null argument");
    if (!ctx->has_drho) fail(OTM_E_STATE, "otm_oc_step before otm_backward");
    if (oc->ne != ctx->g0.nn) fail(OTM_E_ARG, "optimizer size does not match the cell");
    if (oc->device != ctx->device) fail(OTM_E_ARG, "optimizer and context on different devices");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    // on the context's stream, in order with the gradient before and the
    // next evaluation after: no synchronisation but the bisection's own
    const float* g = ctx->drho;
    if (filter_radius > 0) {
      oc_filter_dev(oc, ctx->s, ctx->drho, ctx->rho, ctx->g0, filter_radius, false);
      g = oc->newsens;
    }
    oc_update_dev(oc, ctx->s, g, ctx->rho, volratio, steps, volume, false);
    ctx->has_drho = false;
  });
}

// ---- MMA optimizer ------------------------------------------------------------------
struct otm_mma {
  int m = 0, device = 0;
  long long n = 0;
  double a0 = 1, a = 0, c = 1e3, d = 1;
  hipStream_t s = nullptr;
  MmaArrays A;
  double* gval = nullptr;
  void* small = nullptr;
  MmaHost* mh = nullptr;       // coherent pinned block the finalisers publish to
  MmaHost* mh_dev = nullptr;   // its device address
  float* fx = nullptr;         // design, previous designs, gradients (FP32): (5 + m) n floats
  float* fbounds = nullptr;    // per-element bounds (2n floats), once set
  float* xout = nullptr;       // FP32 result staging for host callers
  MmaProf prof;
  std::vector<double*> owned;
};

static void mma_free(otm_mma* o) {
  if (!o) return;
  if (o->s) (void)hipStreamSynchronize(o->s);
  for (double* p : o->owned) dfree(p);
  dfree(o->gval);
  dfree(o->small);
  dfree(o->fx);
  dfree(o->fbounds);
  dfree(o->xout);
  mma_prof_release(o->prof);
  if (o->mh) (void)hipHostFree(o->mh);
  if (o->s) (void)hipStreamDestroy(o->s);
  delete o;
}

OTM_API int otm_mma_create(int m, int n, double a0, double a, double c, double d, int device, otm_mma** out) {
  return guarded([&] {
    if (!out || n <= 0) fail(OTM_E_ARG, "invalid argument");
    if (m < 1 || m > kMmaMaxM) fail(OTM_E_UNSUPPORTED, "MMA supports 1..4 constraints");
    *out = nullptr;
    otm_mma* o = new otm_mma();
    o->m = m;
    o->n = n;
    o->a0 = a0;
    o->a = a;
    o->c = c;
    o->d = d;
    o->device = device;
    try {
      hip_check(hipSetDevice(device), "hipSetDevice");
      hip_check(hipStreamCreateWithFlags(&o->s, hipStreamDefault), "hipStreamCreate");  // ordered with the default stream
      auto vec = [&](long long cnt) {
        double* p = dalloc<double>(cnt);
        o->owned.push_back(p);
        return p;
      };
      MmaArrays& A = o->A;
      A.n = n;
      const size_t nf = (size_t)n;
      o->fx = dalloc<float>((size_t)(4 + m) * nf);
      A.xvar = o->fx;
      A.xold1 = o->fx + nf;
      A.xold2 = o->fx + 2 * nf;
      A.df0dx = o->fx + 3 * nf;
      A.dgdx = o->fx + 4 * nf;
      A.xmin = A.xmax = nullptr;  // uniform [0, 1] until set (mmaOptimizer.h:21-22)
      A.xmin_u = 0.0;
      A.xmax_u = 1.0;
      A.move = 1.0;
      A.low = vec(n); A.upp = vec(n);
      A.xs = vec(n); A.xsi = vec(n); A.eta = vec(n);
      A.xs_t = vec(n); A.xsi_t = vec(n); A.eta_t = vec(n);
      A.partials = vec((long long)kMmaMaxBlocks * 32);
      o->gval = dalloc<double>(kMmaMaxM);
      o->small = dalloc<char>(mma_small_bytes());
      o->xout = dalloc<float>(n);
      hip_check(hipHostMalloc((void**)&o->mh, sizeof(MmaHost), hipHostMallocCoherent), "hipHostMalloc");
      memset(o->mh, 0, sizeof(MmaHost));
      hip_check(hipHostGetDevicePointer((void**)&o->mh_dev, o->mh, 0), "hipHostGetDevicePointer");
      launch_fill(o->s, A.xold1, 2 * n, 0.f);
      hip_check(hipStreamSynchronize(o->s), "sync");
    } catch (...) {
      mma_free(o);
      throw;
    }
    *out = o;
  });
}

OTM_API int otm_mma_destroy(otm_mma* mma) {
  return guarded([&] { mma_free(mma); });
}

OTM_API int otm_mma_set_bounds(otm_mma* o, double xmin, double xmax) {
  return guarded([&] {
    if (!o || !(xmax > xmin)) fail(OTM_E_ARG, "invalid bounds");
    o->A.xmin = o->A.xmax = nullptr;
    o->A.xmin_u = xmin;
    o->A.xmax_u = xmax;
  });
}

// copies n floats from the caller's host or device array into dst (device)
static void mma_copy_in(otm_mma* o, float* dst, const float* src, int where) {
  hip_check(hipMemcpyAsync(dst, src, sizeof(float) * o->n,
                           where == OTM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, o->s),
            "MMA input copy");
}

OTM_API int otm_mma_set_bounds_array(otm_mma* o, const float* xmin, const float* xmax, int where) {
  return guarded([&] {
    if (!o || !xmin || !xmax) fail(OTM_E_ARG, "null argument");
    check_where(where);
    hip_check(hipSetDevice(o->device), "hipSetDevice");
    if (!o->fbounds) o->fbounds = dalloc<float>(2 * (size_t)o->n);
    mma_copy_in(o, o->fbounds, xmin, where);
    mma_copy_in(o, o->fbounds + o->n, xmax, where);
    hip_check(hipStreamSynchronize(o->s), "sync");
    o->A.xmin = o->fbounds;
    o->A.xmax = o->fbounds + o->n;
  });
}

OTM_API int otm_mma_set_move(otm_mma* o, double move) {
  return guarded([&] {
    if (!o || !(move > 0)) fail(OTM_E_ARG, "invalid move");
    o->A.move = move;
  });
}

OTM_API int otm_mma_update(otm_mma* o, int itn, float* x, const float* dfdx, const double* gval,
                           const float* const* dgdx, int where, int* newton) {
  return guarded([&] {
    if (!o || !x || !dfdx || !gval || !dgdx) fail(OTM_E_ARG, "null argument");
    check_where(where);
    for (int i = 0; i < o->m; i++)
      if (!dgdx[i]) fail(OTM_E_ARG, "null constraint gradient");
    hip_check(hipSetDevice(o->device), "hipSetDevice");
    const long long n = o->n;
    // the FP32 design and gradients are kept as they are; every pass widens
    // them to FP64 (exact), as the reference's type_cast does (mmaOptimizer.cu:63-68)
    mma_copy_in(o, o->A.xvar, x, where);
    mma_copy_in(o, (float*)o->A.df0dx, dfdx, where);
    for (int i = 0; i < o->m; i++) mma_copy_in(o, (float*)o->A.dgdx + (long long)i * n, dgdx[i], where);
    hip_check(hipMemcpyAsync(o->gval, gval, sizeof(double) * o->m, hipMemcpyHostToDevice, o->s), "gval");
    float* xout = where == OTM_DEVICE ? x : o->xout;
    const int rc =
        mma_update(o->s, o->m, o->A, o->small, o->mh, o->mh_dev, o->gval, itn, o->a0, o->a, o->c, o->d, xout,
                   &o->prof);
    if (rc == -2) fail(OTM_E_NUMERIC, "MMA subsolver produced non-finite values");
    if (rc < 0) fail(OTM_E_HIP, "MMA update failed");
    if (where == OTM_HOST)
      hip_check(hipMemcpyAsync(x, o->xout, sizeof(float) * n, hipMemcpyDeviceToHost, o->s), "download x");
    hip_check(hipStreamSynchronize(o->s), "sync");
    if (newton) *newton = rc;
  });
}

OTM_API int otm_mma_profile_enable(otm_mma* o, int enable) {
  return guarded([&] {
    if (!o) fail(OTM_E_ARG, "null optimiser");
    hip_check(hipStreamSynchronize(o->s), "sync");
    for (int k = 0; k < 3; k++) {
      o->prof.launches[k] = 0;
      o->prof.ms[k] = 0.0;
      o->prof.bytes[k] = 0.0;
    }
    o->prof.on = enable != 0;
  });
}

OTM_API int otm_mma_profile_read(otm_mma* o, int kclass, long long* launches, double* ms, double* bytes) {
  return guarded([&] {
    if (!o || kclass < 0 || kclass > 2) fail(OTM_E_ARG, "invalid argument");
    if (launches) *launches = o->prof.launches[kclass];
    if (ms) *ms = o->prof.ms[kclass];
    if (bytes) *bytes = o->prof.bytes[kclass];
  });
}

OTM_API int otm_profile_enable(otm_ctx* ctx, int enable) {
  return guarded([&] {
    if (!ctx) fail(OTM_E_ARG, "null context");
    ctx->sync();
    ctx->prof_on = enable != 0;
    ctx->prof_mask = enable;
    if (ctx->prof_on) {
      for (int k = 0; k < OTM_PROF_NCLASSES; k++) {
        ctx->prof_launches[k] = 0;
        ctx->prof_ms[k] = 0.0;
        ctx->prof_bytes[k] = 0.0;
      }
    }
  });
}

OTM_API int otm_profile_read(otm_ctx* ctx, int kclass, long long* launches, double* ms, double* bytes) {
  return guarded([&] {
    if (!ctx) fail(OTM_E_ARG, "null context");
    if (kclass < 0 || kclass >= OTM_PROF_NCLASSES) fail(OTM_E_ARG, "kernel class out of range");
    ctx->sync();
    if (launches) *launches = ctx->prof_launches[kclass];
    if (ms) *ms = ctx->prof_ms[kclass];
    if (bytes) *bytes = ctx->prof_bytes[kclass];
  });
}

OTM_API int otm_device_alloc(int device, size_t bytes, void** ptr) {
  return guarded([&] {
    if (!ptr) fail(OTM_E_ARG, "null argument");
    hip_check(hipSetDevice(device), "hipSetDevice");
    *ptr = dalloc<char>(bytes);
  });
}

OTM_API int otm_device_free(void* ptr) {
  return guarded([&] { dfree(ptr); });
}

OTM_API int otm_memcpy(void* dst, const void* src, size_t bytes, int dst_where, int src_where) {
  return guarded([&] {
    if (dst_where == OTM_ANY || src_where == OTM_ANY) {
      if (dst_where != OTM_ANY) check_where(dst_where);
      if (src_where != OTM_ANY) check_where(src_where);
      hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDefault), "hipMemcpy");
      return;
    }
    check_where(dst_where);
    check_where(src_where);
    hip_check(hipMemcpy(dst, src, bytes, kind_of(dst_where, src_where)), "hipMemcpy");
  });
}

OTM_API int otm_synchronize(otm_ctx* ctx) {
  return guarded([&] {
    if (!ctx) fail(OTM_E_ARG, "null context");
    ctx->sync();
  });
}

}  // extern "C"

// Solve with exact multicolour Gauss-Seidel on every level (the batched
// runtime's fallback for a cell whose brick smoother lost positivity).
bool otm::ctx_solve_exact(otm_ctx* c, double C[9]) {
  set_bricks(c, false);
  return do_heat_matrix(c, C);
}
