// opentm.hpp -- C++ classes with the reference's names over the C ABI of
// opentm.h, so that code written against quanyuchen2000/OPENTM's
// homo::Homogenization_H / OCOptimizer / MMAOptimizer, cfg::HomoConfig,
// runCustom, runInstance and wrapper_homo compiles against this engine.
// Errors throw std::runtime_error with the engine's message (the reference
// throws std::runtime_error too, e.g. "MG numerical explode", MG.cpp:116).
#pragma once

#include <cmath>
#include <cstdint>
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

// homogenization.h:15-32. Pointers are device pointers, as in the reference.
struct Homogenization_H {
  cfg::HomoConfig config;
  otm_ctx* ctx = nullptr;
  float power_penal = 1;
  float diag_strength = 1e6;
  explicit Homogenization_H(cfg::HomoConfig c) { build(c); }
  Homogenization_H(const Homogenization_H&) = delete;
  Homogenization_H& operator=(const Homogenization_H&) = delete;
  ~Homogenization_H() { otm_destroy(ctx); }
  void build(cfg::HomoConfig c) {
    if (ctx) otm_destroy(ctx);
    ctx = nullptr;
    config = c;
    otm_config e = c.engine();
    check(otm_create(&e, &ctx));
  }
  void update(float* rho = nullptr, int pitchT = -1) {
    check(otm_update(ctx, rho, pitchT < 0 ? config.reso[0] : pitchT, OTM_DEVICE));
  }
  void heatMatrix(double C[3][3]) {
    double c[9];
    check(otm_heat_matrix(ctx, c));
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
  void update(int itn, float* x, float* dfdx, float* gval, float** dgdx) {
    std::vector<double> g(m);
    for (int i = 0; i < m; i++) g[i] = gval[i];  // gval is host memory here
    check(otm_mma_update(h, itn, x, dfdx, g.data(), dgdx, OTM_DEVICE, nullptr));
  }
};

// Framework.cuh:312-343: converged when the relative objective change stays
// below the threshold for three consecutive iterations after iteration 11.
struct ConvergeChecker {
 private:
  static constexpr int circleLen = 10;
  double thres = 1e-3;
  double Obj[circleLen] = {};
  int p_cir = 0;
  int cirId(int id) const {
    while (id < 0) id += circleLen;
    return id % circleLen;
  }
  double get(int id) const { return Obj[cirId(id)]; }
  bool checkSeq() const {
    bool out = false;
    for (int i = 0; i < 3; i++)
      out = out || std::fabs(get(p_cir - i) - get(p_cir - i - 1)) / std::fabs(get(p_cir - i - 1)) > thres;
    return !out;
  }

 public:
  explicit ConvergeChecker(double thres_ = 1e-3) : thres(thres_) {}
  bool is_converge(int iter, double obj) {
    const int id = cirId(iter);
    p_cir = id;
    Obj[id] = obj;
    if (iter <= 11) return false;
    return checkSeq();
  }
};

// Framework.cuh:345-411: adaptive volume bound for the OC tensor-matching loop.
struct VolumeGovernor {
 private:
  float lowBound = 0;
  float volume_bound = 1.0f;
  float decrease = 0;
  float decrease_factor = 1.f;
  float val_last = 1;
  float vibrate = 0;
  int vibrate_count = 0;
  int count = 0;
  void shrink() {
    decrease = volume_bound - lowBound;
    volume_bound = volume_bound - decrease_factor * decrease;
    if (decrease_factor > 0.05f) decrease_factor *= 0.8f;
  }
  void expand() { volume_bound += 0.3f * decrease * decrease_factor; }

 public:
  float get_volume_bound() const { return volume_bound; }
  float get_current_decrease() const { return decrease_factor * decrease; }
  int volume_check(float value, float lowBound_, float volfrac, int itn) {
    lowBound = lowBound_;
    if (value + 0.0001f < 1e-4f || itn == 100) {
      shrink();
      if (decrease_factor <= 0.05f) return 1;
      count = 0;
    }
    const bool reach = std::fabs(volume_bound - volfrac) < 1e-2f;
    if (vibrate * (val_last - value) < 0)
      vibrate_count++;
    else
      vibrate_count = 0;
    if (vibrate_count >= 8) {
      if (reach) {
        expand();
        vibrate_count = 0;
      } else {
        shrink();
      }
    }
    vibrate = val_last - value;
    if (val_last - value < 1e-4f && value > 1e-4f && reach)
      count++;
    else
      count = 0;
    if (count >= 5) {
      volume_bound += 0.3f * decrease * decrease_factor;
      count = 0;
    }
    val_last = value;
    return 0;
  }
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
