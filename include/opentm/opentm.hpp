// opentm.hpp -- C++ classes with the reference's names over the C ABI of
// opentm.h, so that code written against quanyuchen2000/OPENTM's
// homo::Homogenization_H / MG_H / Grid_H (subset) / OCOptimizer /
// MMAOptimizer, cfg::HomoConfig,
// runCustom, runInstance and wrapper_homo compiles against this engine.
// Errors throw std::runtime_error with the engine's message (the reference
// throws std::runtime_error too, e.g. "MG numerical explode", MG.cpp:116).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../opentm.h"

namespace cfg {
// cmdline.h:8-11
enum class Objective : uint8_t { bulk, shear, npr, custom };
enum class Symmetry : uint8_t { reflect3, reflect6, rotate3, rotate2, NONE };
enum class InitWay : uint8_t { random, randcenter, noise, manual, interp, rep_randcenter, P, G, D, IWP, example };
enum class Model : uint8_t { mma, oc };

// cmdline.h:12-33; defaults of HomoConfig::init (cmdline.cpp:4-35). The
// solver knobs after `initperiod` are this engine's.
struct HomoConfig {
  Symmetry sym;
  InitWay winit;
  Model model;
  int reso[3];
  double volRatio;
  double heatRatio[2];
  double finthres;
  double filterRadius;
  double designStep;
  double dampRatio;
  double femRelThres;
  double target_tensor[6];
  std::string outprefix;
  std::string testname;
  std::string inputrho;
  bool useManagedMemory = true;
  int logrho, logc, logsens, logobj;
  int max_iter = 100;
  int initperiod;
  int seed = 0;          // seed of the random initialisers (the reference seeds from the clock)
  int coarse_sweeps = 0;  // Gauss-Seidel sweeps on coarse levels (0: as on the fine level)
  int fp64_refine = 0;    // FP64 iterate with FP32 inner solves (iterative refinement)
  int fp16_stencils = 0;  // coarse stencils stored in FP16 with a per-level scale
  int device = 0;
  void init() {
    reso[0] = reso[1] = reso[2] = 128;
    sym = Symmetry::NONE;
    winit = InitWay::IWP;
    volRatio = 0.3;
    heatRatio[0] = 1;
    heatRatio[1] = 0;
    outprefix = "";
    logrho = logc = logsens = logobj = 0;
    testname = "none";
    useManagedMemory = true;
    inputrho = "init rand";
    max_iter = 500;
    initperiod = 10;
    finthres = 5e-4;
    filterRadius = 2.0;
    designStep = 0.05;
    dampRatio = 0.5;
    femRelThres = 1e-6;  // the reference's 1e-2 targets its stationary MG loop
    target_tensor[0] = 0.3;
    target_tensor[1] = 0.2;
    target_tensor[2] = 0.1;
    target_tensor[3] = target_tensor[4] = target_tensor[5] = 0.;
    seed = 0;
    coarse_sweeps = 0;
    fp64_refine = 0;
    fp16_stencils = 0;
    device = 0;
  }
  otm_config engine() const {
    otm_config c;
    otm_config_init(&c);
    for (int a = 0; a < 3; a++) c.reso[a] = reso[a];
    c.heat_ratio[0] = heatRatio[0];
    c.heat_ratio[1] = heatRatio[1];
    c.fem_rel_thres = femRelThres;
    c.coarse_sweeps = coarse_sweeps;
    c.fp64_refine = fp64_refine;
    c.fp16_stencils = fp16_stencils;
    c.device = device;
    return c;
  }
};
}  // namespace cfg

namespace homo {

inline void check(int status) {
  if (status != OTM_OK) throw std::runtime_error(otm_last_error());
}

// grid.h:89
enum VoxelIOFormat { binary, openVDB };

// Grid_H (grid.h:104-421), the subset a caller of Homogenization_H / MG_H
// touches, as a view of one level of an engine context. The reference's
// per-grid device buffers (padded Gauss-Seidel order) have no counterpart:
// the engine keeps every level lexicographic and unpadded.
struct Grid_H {
  otm_ctx* ctx = nullptr;
  int level = 0;
  std::array<int, 3> cellReso{{0, 0, 0}};
  Grid_H(otm_ctx* c, int l) : ctx(c), level(l) { check(otm_level_reso(c, l, cellReso.data())); }
  int n_cells() const { return cellReso[0] * cellReso[1] * cellReso[2]; }
  // nodes of the level (the engine stores no periodic images or padding)
  int n_gsvertices() const { return n_cells(); }
  int n_gscells() const { return n_cells(); }
  // Grid_H::update (root level): SIMP-mapped conductivity, device pointer
  void update(float* rho, int pitchT = -1, bool lexiOrder = true) {
    if (level != 0) throw std::runtime_error("Grid_H::update applies to the root grid");
    if (!lexiOrder) throw std::runtime_error("only lexicographic conductivities are accepted");
    check(otm_update(ctx, rho, pitchT < 0 ? cellReso[0] : pitchT, OTM_DEVICE));
  }
  void reset_displacement() {
    if (level != 0) throw std::runtime_error("only the root grid's displacement is held between solves");
    check(otm_reset_displacement(ctx));
  }
  // Grid_H::getDensity (root level): lexicographic conductivity on the host
  void getDensity(std::vector<float>& rho, bool lexiOrder = true) {
    if (level != 0) throw std::runtime_error("Grid_H::getDensity applies to the root grid");
    (void)lexiOrder;  // the engine's order is lexicographic
    rho.resize((size_t)n_cells());
    check(otm_get_kappa(ctx, rho.data(), OTM_HOST));
  }
  // Grid_H::writeDensity (grid.cpp:330-345); OpenVDB is not available here
  void writeDensity(const std::string& fname, VoxelIOFormat frmat) {
    if (frmat != binary) throw std::runtime_error("OpenVDB output is not supported; use VoxelIOFormat::binary");
    std::vector<float> rho;
    getDensity(rho);
    check(otm_write_density(fname.c_str(), cellReso.data(), rho.data(), OTM_HOST));
  }
};

// MGConfig / MG_H (MG.h:8-32) over an engine context. MG_H::build makes
// a context of its own; the MG_H of a Homogenization_H shares that one's.
struct MGConfig {
  int reso[3];
  bool enableManagedMem = true;
  std::string namePrefix;
};

struct MG_H {
  std::vector<std::shared_ptr<Grid_H>> grids;  // fine to coarse
  MGConfig mgConfig;
  otm_ctx* ctx = nullptr;
  bool owns = false;
  MG_H() = default;
  MG_H(const MG_H&) = delete;
  MG_H& operator=(const MG_H&) = delete;
  ~MG_H() { release(); }
  void release() {
    grids.clear();
    if (owns) otm_destroy(ctx);
    ctx = nullptr;
    owns = false;
  }
  // view of a context owned elsewhere (Homogenization_H::mg_)
  void attach(otm_ctx* c) {
    release();
    ctx = c;
    int nl = 0;
    check(otm_levels(c, &nl));
    for (int l = 0; l < nl; l++) grids.push_back(std::make_shared<Grid_H>(c, l));
    for (int a = 0; a < 3; a++) mgConfig.reso[a] = grids[0]->cellReso[a];
  }
  void build(MGConfig config) {
    otm_config e;
    otm_config_init(&e);
    for (int a = 0; a < 3; a++) e.reso[a] = config.reso[a];
    otm_ctx* c = nullptr;
    check(otm_create(&e, &c));
    attach(c);
    owns = true;
    mgConfig = config;
  }
  std::shared_ptr<Grid_H> getRootGrid() { return grids.at(0); }
  // one stationary V-cycle of the three load cases (pre/post as in the
  // reference: ignored; the engine's smoothing schedule is fixed per context)
  void v_cycle(float w_SOR = 1.f, int pre = 1, int post = 1) {
    (void)w_SOR;
    (void)pre;
    (void)post;
    check(otm_mg_cycles(ctx, 1, nullptr));
  }
  void reset_displacement() { check(otm_reset_displacement(ctx)); }
  // MG-preconditioned CG of the three load cases to `tol`; returns the
  // largest relative residual reached (the reference returns its residual)
  double solveEquation(double tol = 1e-2, bool with_guess = true) {
    double rr[3] = {0, 0, 0};
    const int st = otm_solve(ctx, tol, with_guess ? 1 : 0, rr);
    if (st != OTM_OK && st != OTM_E_NOCONV) check(st);  // the reference reports and carries on
    return std::max(rr[0], std::max(rr[1], rr[2]));
  }
  void updateStencils() { check(otm_update(ctx, nullptr, mgConfig.reso[0], OTM_DEVICE)); }
};

// homogenization.h:15-32. Pointers are device pointers, as in the reference.
struct Homogenization_H {
  cfg::HomoConfig config;
  otm_ctx* ctx = nullptr;
  std::shared_ptr<Grid_H> grid;  // root grid (homogenization.h:21)
  std::unique_ptr<MG_H> mg_;     // the hierarchy (homogenization.h:22)
  float power_penal = 1;
  float diag_strength = 1e6;
  explicit Homogenization_H(cfg::HomoConfig c) { build(c); }
  Homogenization_H(const Homogenization_H&) = delete;
  Homogenization_H& operator=(const Homogenization_H&) = delete;
  ~Homogenization_H() {
    mg_.reset();
    grid.reset();
    otm_destroy(ctx);
  }
  void build(cfg::HomoConfig c) {
    mg_.reset();
    grid.reset();
    if (ctx) otm_destroy(ctx);
    ctx = nullptr;
    config = c;
    otm_config e = c.engine();
    check(otm_create(&e, &ctx));
    mg_.reset(new MG_H());
    mg_->attach(ctx);
    grid = mg_->getRootGrid();
  }
  void update(float* rho = nullptr, int pitchT = -1) {
    check(otm_update(ctx, rho, pitchT < 0 ? config.reso[0] : pitchT, OTM_DEVICE));
  }
  void heatMatrix(double C[3][3]) {
    double c[9];
    const int st = otm_heat_matrix(ctx, c);
    if (st != OTM_OK && st != OTM_E_NOCONV) check(st);  // NOCONV: C is written, as the reference carries on
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) C[i][j] = c[3 * i + j];
  }
  void heatMatrix(float C[3][3]) {
    double c[3][3];
    heatMatrix(c);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) C[i][j] = (float)c[i][j];
  }
  void Sensitivity(float dC[3][3], float* sens, int pitchT, bool lexiOrder = false) {
    if (!lexiOrder) throw std::runtime_error("only lexicographic sensitivity output is provided");
    double d[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) d[3 * i + j] = dC[i][j];
    check(otm_sensitivity(ctx, d, sens, pitchT, OTM_DEVICE));
  }
  // no effect on the solve in the reference either (homogenization.cpp:83-87)
  void ConfigDiagPrecondition(float strength) { diag_strength = strength; }
};

// OCOptimizer.h:7-20
struct OCOptimizer {
  int ne;
  float step_limit = 0.02f, damp = 0.5f, minRho = 0.001f;
  otm_oc* h = nullptr;
  OCOptimizer(int Ne, float min_density, float stepLimit, float dampExponent)
      : ne(Ne), step_limit(stepLimit), damp(dampExponent), minRho(min_density) {
    check(otm_oc_create(ne, minRho, step_limit, damp, 0, &h));
  }
  OCOptimizer(const OCOptimizer&) = delete;
  OCOptimizer& operator=(const OCOptimizer&) = delete;
  ~OCOptimizer() { otm_oc_destroy(h); }
  void filterSens(float* sens, const float* rho, size_t pitchT, int reso[3], float radius = 2) {
    check(otm_oc_filter_sens(h, sens, rho, (int)pitchT, reso, radius, OTM_DEVICE));
  }
  void update(const float* sens, float* rho, float volratio) {
    check(otm_oc_update(h, sens, rho, volratio, OTM_DEVICE, nullptr, nullptr));
  }
};

// mmaOptimizer.h:12-35
struct MMAOptimizer {
  int n, m;
  double a0, a, c, d;
  otm_mma* h = nullptr;
  MMAOptimizer(int nconstrain, int nvar, double a0_, double a_, double c_, double d_)
      : n(nvar), m(nconstrain), a0(a0_), a(a_), c(c_), d(d_) {
    check(otm_mma_create(m, n, a0, a, c, d, 0, &h));
  }
  MMAOptimizer(const MMAOptimizer&) = delete;
  MMAOptimizer& operator=(const MMAOptimizer&) = delete;
  ~MMAOptimizer() { otm_mma_destroy(h); }
  void setBound(float xmin, float xmax) { check(otm_mma_set_bounds(h, xmin, xmax)); }
  void setBound(float* xmin, float* xmax) { check(otm_mma_set_bounds_array(h, xmin, xmax, OTM_DEVICE)); }
  // x, dfdx, gval and dgdx[i] are device (or managed) float arrays, as the
  // reference's callers pass them (mmaOptimizer.cu:61-69 casts gval on the
  // device); the m constraint values are read back before the update
  void update(int itn, float* x, float* dfdx, float* gval, float** dgdx) {
    std::vector<float> gf((size_t)m);
    check(otm_memcpy(gf.data(), gval, sizeof(float) * (size_t)m, OTM_HOST, OTM_ANY));
    std::vector<double> g(gf.begin(), gf.end());
    check(otm_mma_update(h, itn, x, dfdx, g.data(), dgdx, OTM_DEVICE, nullptr));
  }
};

// Convergence test of runCustom (Framework.cuh:312-343). Behaviour: every
// iteration's objective is recorded; up to and including iteration 11 the
// answer is "not converged"; afterwards the loop has converged when each of
// the last three steps (obj[it-1] -> obj[it], obj[it-2] -> obj[it-1],
// obj[it-3] -> obj[it-2]) changed the objective by at most `thres` relative
// to the earlier value. A comparison that cannot be made (NaN) counts as
// within the threshold, as in the reference.
struct ConvergeChecker {
  explicit ConvergeChecker(double thres_ = 1e-3) : tol_(thres_) {}
  bool is_converge(int iter, double obj) {
    // the last four objectives, newest first
    for (int i = 3; i > 0; i--) recent_[i] = recent_[i - 1];
    recent_[0] = obj;
    if (iter <= 11) return false;
    for (int step = 0; step < 3; step++) {
      const double now = recent_[step], before = recent_[step + 1];
      const double change = std::fabs(now - before) / std::fabs(before);
      if (change > tol_) return false;
    }
    return true;
  }

 private:
  double tol_;
  double recent_[4] = {0.0, 0.0, 0.0, 0.0};
};

// Adaptive volume bound of runCustom's OC branch (Framework.cuh:345-411),
// as a small state machine over: the bound B (starts at 1), the lower
// limit L given by the caller, the last cut D and the cut factor F (starts
// at 1), the previous objective and its previous change, an oscillation
// counter and a stall counter. Per check(value v, lower, volfrac, itn):
//  1. L = lower. If v < 0 within 1e-4 (v + 1e-4 < 1e-4) or itn == 100, cut:
//     D = B - L, B -= F D, and F shrinks by 0.8 while above 0.05. If F is
//     then at most 0.05, report 1 (stop); otherwise reset the stall counter.
//  2. "reached" means |B - volfrac| < 1e-2.
//  3. With c = previous objective - v: a sign change of c against the
//     previous change counts one oscillation, anything else resets the count.
//     Eight oscillations in a row: widen (B += 0.3 D F, count reset) when
//     reached, otherwise cut as in 1 (count kept).
//  4. A step with c < 1e-4, v > 1e-4 and reached counts one stall (else the
//     stall count resets); five stalls widen the bound by 0.3 D F and reset.
//  5. Remember v and c; report 0.
// The state is FP32 and the constants are double literals, evaluated in the
// reference's precision (float state, double arithmetic where it mixes).
class VolumeGovernor {
 public:
  float get_volume_bound() const { return bound_; }
  float get_current_decrease() const { return factor_ * cut_; }
  int volume_check(float value, float lowBound_, float volfrac, int itn) {
    lower_ = lowBound_;
    if ((double)value + 0.0001 < 1e-4 || itn == 100) {
      cut();
      if ((double)factor_ <= 0.05) return 1;
      stalls_ = 0;
    }
    const bool reached = std::fabs(bound_ - volfrac) < 1e-2;
    const float change = last_value_ - value;
    oscillations_ = (last_change_ * change < 0) ? oscillations_ + 1 : 0;
    if (oscillations_ >= 8) {
      if (reached) {
        widen();
        oscillations_ = 0;
      } else {
        cut();
      }
    }
    last_change_ = change;
    const bool stalled = (double)change < 1e-4 && (double)value > 1e-4 && reached;
    stalls_ = stalled ? stalls_ + 1 : 0;
    if (stalls_ >= 5) {
      widen();
      stalls_ = 0;
    }
    last_value_ = value;
    return 0;
  }

 private:
  void cut() {
    cut_ = bound_ - lower_;
    bound_ = bound_ - factor_ * cut_;
    if ((double)factor_ > 0.05) factor_ = (float)((double)factor_ * 0.8);
  }
  void widen() { bound_ = (float)((double)bound_ + 0.3 * (double)cut_ * (double)factor_); }
  float bound_ = 1.0f, lower_ = 0.0f, cut_ = 0.0f, factor_ = 1.0f;
  float last_value_ = 1.0f, last_change_ = 0.0f;
  int oscillations_ = 0, stalls_ = 0;
};

}  // namespace homo

// runCustom (Framework.cu:147-244): the design loop on this engine. rho0
// (host, reso^3 floats) replaces the initialiser when given.
inline std::vector<float> runCustom(cfg::HomoConfig config, std::vector<float>* rho0 = nullptr) {
  otm_custom_params p;
  otm_custom_params_init(&p);
  p.base = config.engine();
  p.model = config.model == cfg::Model::mma ? OTM_MODEL_MMA : OTM_MODEL_OC;
  p.initway = (int)config.winit;
  p.seed = config.seed;
  p.max_iter = config.max_iter;
  p.finthres = config.finthres;
  p.volratio = config.volRatio;
  p.sym = (int)config.sym;
  for (int i = 0; i < 6; i++) p.target[i] = config.target_tensor[i];
  const size_t n = (size_t)config.reso[0] * config.reso[1] * config.reso[2];
  std::vector<float> rho(n);
  const float* r0 = nullptr;
  if (rho0) {
    if (rho0->size() != n) throw std::runtime_error("initial density has the wrong size");
    r0 = rho0->data();
  }
  homo::check(otm_run_custom(&p, r0, rho.data(), nullptr));
  return rho;
}

// main.cpp:45-62
inline std::vector<float> runInstance(int reso, std::vector<double> heat_ratios, std::vector<double> target_ratio,
                                      cfg::InitWay initway, cfg::Model model) {
  cfg::HomoConfig config;
  config.init();
  config.reso[0] = config.reso[1] = config.reso[2] = reso;
  if (heat_ratios.size() != 2) throw std::runtime_error("Input wrong size of heat_ratios");
  if (target_ratio.size() != 6) throw std::runtime_error("target tensor needs 6 entries");
  for (int i = 0; i < 2; i++) config.heatRatio[i] = heat_ratios[i];
  for (int i = 0; i < 6; i++) config.target_tensor[i] = target_ratio[i];
  config.winit = initway;
  config.model = model;
  return runCustom(config);
}

// main.cpp:17-44
class wrapper_homo {
  cfg::HomoConfig config;
  std::vector<float> rho;

 public:
  wrapper_homo() { config.init(); }
  void setDensity(std::vector<float> rho_) { rho = std::move(rho_); }
  void setConfig(int reso, std::vector<double> heat_ratios, std::vector<double> target_ratio, cfg::Model model) {
    config.init();
    config.reso[0] = config.reso[1] = config.reso[2] = reso;
    if (heat_ratios.size() != 2) throw std::runtime_error("Input wrong size of heat_ratios");
    if (target_ratio.size() != 6) throw std::runtime_error("target tensor needs 6 entries");
    for (int i = 0; i < 2; i++) config.heatRatio[i] = heat_ratios[i];
    for (int i = 0; i < 6; i++) config.target_tensor[i] = target_ratio[i];
    config.model = model;
  }
  cfg::HomoConfig& configuration() { return config; }
  std::vector<float> optimize() { return runCustom(config, rho.empty() ? nullptr : &rho); }
};
