// otm_framework.cpp -- runCustom (Framework.cu:147-244) over the engine's
// public C ABI: the tensor-matching design loop with OC (VolumeGovernor +
// ConvergeChecker) or MMA. Everything stays on the device between
// iterations; only the 3x3 tensor, the volume statistics and the MMA
// constraint value come back to the host.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/opentm.h"
#include "../../include/opentm/opentm.hpp"

namespace {

struct Fail {
  int code;
};
inline void ck(int status) {
  if (status != OTM_OK) throw Fail{status};
}

// otm_eval inside the loop: a solve that ends above its tolerance is
// reported and the loop carries on with its result, as the reference's
// MG_H::solveEquation prints "Failed to converge" and returns (MG.cpp:101-118).
inline void eval_counted(otm_ctx* c, const otm_density_params* dp, double C[9], int* unconverged) {
  const int st = otm_eval(c, dp, C);
  if (st == OTM_E_NOCONV) {
    (*unconverged)++;
    return;
  }
  ck(st);
}

// objective of runCustom (Framework.cu:168-170) and its gradient wrt C
double tensor_objective(const double C[9], const double t[6], double dC[9]) {
  static const int ii[6] = {0, 1, 2, 0, 2, 0}, jj[6] = {0, 1, 2, 1, 1, 2};
  double f = 0.0;
  for (int q = 0; q < 9; q++) dC[q] = 0.0;
  for (int q = 0; q < 6; q++) {
    const double e = C[ii[q] * 3 + jj[q]] - t[q];
    f += e * e;
    dC[ii[q] * 3 + jj[q]] += 2.0 * e;
  }
  return f - 1e-4;
}

struct DevBuf {
  void* p = nullptr;
  DevBuf(int device, size_t bytes) { ck(otm_device_alloc(device, bytes, &p)); }
  ~DevBuf() { otm_device_free(p); }
  float* f() { return static_cast<float*>(p); }
};

struct Ctx {
  otm_ctx* c = nullptr;
  ~Ctx() { otm_destroy(c); }
};

}  // namespace

extern "C" {

OTM_API void otm_custom_params_init(otm_custom_params* p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  otm_config_init(&p->base);
  p->model = OTM_MODEL_OC;
  p->initway = 9;  // IWP (cmdline.cpp:10)
  p->seed = 0;
  p->max_iter = 500;         // cmdline.cpp:22
  p->finthres = 5e-4;        // cmdline.cpp:24
  p->volratio = 0.3;         // cmdline.cpp:11
  const double t[6] = {0.3, 0.2, 0.1, 0.0, 0.0, 0.0};  // cmdline.cpp:29-34
  for (int i = 0; i < 6; i++) p->target[i] = t[i];
  p->sym = OTM_SYM_NONE;  // cmdline.cpp:5
  otm_log_params_init(&p->log);
}

OTM_API void otm_log_params_init(otm_log_params* p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));  // no prefix, every interval off
}

// logIter (Framework.cu:7-59). Text lines are appended (the file is
// started afresh at iteration 0); density and gradient files are written
// whole in the binary density format.
OTM_API int otm_log_iteration(otm_ctx* ctx, const otm_log_params* p, int iter, const double C[9], double obj,
                              const float* sens, int where) {
  try {
    if (!ctx || !p || !C || iter < 0) return OTM_E_ARG;
    if (!p->prefix) return OTM_OK;
    int reso[3];
    ck(otm_level_reso(ctx, 0, reso));
    const size_t ne = (size_t)reso[0] * reso[1] * reso[2];
    const std::string pre(p->prefix);
    auto write_design = [&](const std::string& name) {
      std::vector<float> rho(ne);
      ck(otm_get_design(ctx, rho.data(), OTM_HOST));
      ck(otm_write_density((pre + name).c_str(), reso, rho.data(), OTM_HOST));
    };
    auto append = [&](const std::string& name, const std::string& line) {
      FILE* f = fopen((pre + name).c_str(), iter == 0 ? "w" : "a");
      if (!f) throw Fail{OTM_E_ARG};
      fputs(line.c_str(), f);
      fclose(f);
    };
    char buf[64];
    if (iter % 5 == 0) {
      write_design("rho");
      FILE* f = fopen((pre + "C").c_str(), "w");
      if (!f) throw Fail{OTM_E_ARG};
      for (int i = 0; i < 3; i++) fprintf(f, "%.9e %.9e %.9e\n", C[3 * i], C[3 * i + 1], C[3 * i + 2]);
      fclose(f);
    }
    if (p->logrho > 0 && iter % p->logrho == 0) {
      snprintf(buf, sizeof(buf), "rho_%04d", iter);
      write_design(buf);
    }
    if (p->logc > 0 && iter % p->logc == 0) {
      std::string line = "iter " + std::to_string(iter) + " ";
      for (int q = 0; q < 9; q++) {
        snprintf(buf, sizeof(buf), "%.9e ", C[q]);
        line += buf;
      }
      append("Clog", line + "\n");
    }
    if (p->logsens > 0 && iter % p->logsens == 0) {
      if (!sens) return OTM_E_ARG;
      snprintf(buf, sizeof(buf), "sens_%04d", iter);
      ck(otm_write_density((pre + buf).c_str(), reso, sens, where));
    }
    if (p->logobj > 0 && iter % p->logobj == 0) {
      snprintf(buf, sizeof(buf), "obj = %.9e", obj);
      append("objlog", "iter " + std::to_string(iter) + " " + buf + "\n");
    }
    return OTM_OK;
  } catch (const Fail& f) {
    return f.code;
  }
}

// Binary density format of Grid_H::writeDensity (grid.cpp:330-345): three
// int32 cell resolutions, then one float32 per element in lexicographic
// order (x fastest).
OTM_API int otm_write_density(const char* path, const int reso[3], const float* rho, int where) {
  if (!path || !reso || !rho || (where != OTM_HOST && where != OTM_DEVICE)) return OTM_E_ARG;
  const size_t n = (size_t)reso[0] * reso[1] * reso[2];
  if (reso[0] <= 0 || reso[1] <= 0 || reso[2] <= 0) return OTM_E_ARG;
  std::vector<float> host;
  const float* src = rho;
  if (where == OTM_DEVICE) {
    host.resize(n);
    const int st = otm_memcpy(host.data(), rho, sizeof(float) * n, OTM_HOST, OTM_DEVICE);
    if (st != OTM_OK) return st;
    src = host.data();
  }
  FILE* f = fopen(path, "wb");
  if (!f) return OTM_E_ARG;
  const int32_t hdr[3] = {reso[0], reso[1], reso[2]};
  const bool ok = fwrite(hdr, sizeof(int32_t), 3, f) == 3 && fwrite(src, sizeof(float), n, f) == n;
  return (fclose(f) == 0 && ok) ? OTM_OK : OTM_E_ARG;
}

OTM_API int otm_read_density(const char* path, int reso[3], float* rho, int where) {
  if (!path || !reso || (where != OTM_HOST && where != OTM_DEVICE)) return OTM_E_ARG;
  FILE* f = fopen(path, "rb");
  if (!f) return OTM_E_ARG;
  int32_t hdr[3];
  if (fread(hdr, sizeof(int32_t), 3, f) != 3 || hdr[0] <= 0 || hdr[1] <= 0 || hdr[2] <= 0) {
    fclose(f);
    return OTM_E_ARG;
  }
  for (int a = 0; a < 3; a++) reso[a] = hdr[a];
  const size_t n = (size_t)hdr[0] * hdr[1] * hdr[2];
  int st = OTM_OK;
  if (rho) {
    std::vector<float> host(n);
    if (fread(host.data(), sizeof(float), n, f) != n) st = OTM_E_ARG;
    if (st == OTM_OK) {
      if (where == OTM_DEVICE)
        st = otm_memcpy(rho, host.data(), sizeof(float) * n, OTM_DEVICE, OTM_HOST);
      else
        memcpy(rho, host.data(), sizeof(float) * n);
    }
  }
  fclose(f);
  return st;
}

OTM_API int otm_run_custom(const otm_custom_params* p, const float* rho0, float* rho_out, otm_custom_report* report) {
  try {
    if (!p || !rho_out) return OTM_E_ARG;
    const int* reso = p->base.reso;
    const long long ne = (long long)reso[0] * reso[1] * reso[2];
    const int dev = p->base.device;
    Ctx hom;
    ck(otm_create(&p->base, &hom.c));
    if (rho0) {
      ck(otm_set_design(hom.c, rho0, OTM_HOST));
    } else {
      otm_init_params ip;  // initDensity (Framework.cu:61-145), symmetry included
      otm_init_params_init(&ip);
      ip.initway = p->initway;
      ip.seed = p->seed;
      ip.volratio = p->volratio;
      ip.sym = p->sym;
      ck(otm_init_design(hom.c, &ip, nullptr));
    }
    otm_density_params dp;
    otm_density_params_init(&dp);  // Spline4 r = 1.5, non-periodic (Framework.cu:163)
    homo::ConvergeChecker criteria(p->finthres);
    double C[9] = {0}, dC[9], obj = 0.0;
    int itn = 0;
    int unconverged = 0;
    bool converged = false;
    if (p->model == OTM_MODEL_MMA) {
      // MMAOptimizer mma(1, ne, 1, 0, 1e7, 1); setBound(1e-4, 1) (Framework.cu:172-173):
      // the volume is the objective and the tensor error the constraint
      otm_mma* mma = nullptr;
      ck(otm_mma_create(1, (int)ne, 1.0, 0.0, 1e7, 1.0, dev, &mma));
      struct MmaGuard {
        otm_mma* m;
        ~MmaGuard() { otm_mma_destroy(m); }
      } guard{mma};
      ck(otm_mma_set_bounds(mma, 1e-4, 1.0));
      DevBuf rho(dev, sizeof(float) * ne), dfdx(dev, sizeof(float) * ne), dvdx(dev, sizeof(float) * ne);
      std::vector<float> vol((size_t)ne, (float)(1.0 / (double)ne));
      ck(otm_memcpy(dvdx.p, vol.data(), sizeof(float) * ne, OTM_DEVICE, OTM_HOST));
      for (itn = 0; itn < p->max_iter; itn++) {
        eval_counted(hom.c, &dp, C, &unconverged);
        obj = tensor_objective(C, p->target, dC);
        const float f0val = (float)obj;  // the reference's objective value is a float (Framework.cu:181)
        ck(otm_backward(hom.c, dC, dfdx.f(), OTM_DEVICE));
        ck(otm_log_iteration(hom.c, &p->log, itn, C, obj, dfdx.f(), OTM_DEVICE));
        if (criteria.is_converge(itn, f0val)) {
          converged = true;
          break;
        }
        ck(otm_get_design(hom.c, rho.f(), OTM_DEVICE));
        const float* dg[1] = {dfdx.f()};
        const double gval[1] = {f0val};
        ck(otm_mma_update(mma, itn, rho.f(), dvdx.f(), gval, dg, OTM_DEVICE, nullptr));
        ck(otm_set_design(hom.c, rho.f(), OTM_DEVICE));
      }
    } else {
      // OCOptimizer oc(ne, 0.001, 0.02, 0.5) with the adaptive volume bound
      // (Framework.cu:206-231)
      otm_oc* oc = nullptr;
      ck(otm_oc_create((int)ne, 0.001, 0.02, 0.5, dev, &oc));
      struct OcGuard {
        otm_oc* o;
        ~OcGuard() { otm_oc_destroy(o); }
      } guard{oc};
      homo::VolumeGovernor governor;
      DevBuf sens_log(dev, p->log.prefix && p->log.logsens > 0 ? sizeof(float) * ne : sizeof(float));
      for (itn = 0; itn < p->max_iter; itn++) {
        eval_counted(hom.c, &dp, C, &unconverged);
        obj = tensor_objective(C, p->target, dC);
        const float val = (float)obj;  // a float in the reference (Framework.cu:211)
        double vf = 0.0, lower = 0.0;
        ck(otm_design_stats(hom.c, &vf, &lower));
        governor.volume_check(val, (float)lower, (float)vf, itn);
        if (p->log.prefix && p->log.logsens > 0 && itn % p->log.logsens == 0) {
          ck(otm_backward(hom.c, dC, sens_log.f(), OTM_DEVICE));
          ck(otm_log_iteration(hom.c, &p->log, itn, C, obj, sens_log.f(), OTM_DEVICE));
        } else {
          ck(otm_backward(hom.c, dC, nullptr, OTM_DEVICE));
          ck(otm_log_iteration(hom.c, &p->log, itn, C, obj, nullptr, OTM_DEVICE));
        }
        if (criteria.is_converge(itn, val) && governor.get_current_decrease() < 1e-4) {
          converged = true;
          break;
        }
        ck(otm_symmetrize_design(hom.c, p->sym));  // design and gradient (Framework.cu:223-224)
        ck(otm_oc_step(hom.c, oc, 2.0, governor.get_volume_bound(), nullptr, nullptr));
      }
    }
    ck(otm_get_design(hom.c, rho_out, OTM_HOST));
    if (report) {
      report->iterations = itn;
      report->converged = converged ? 1 : 0;
      report->objective = obj;
      double vf = 0.0;
      ck(otm_design_stats(hom.c, &vf, nullptr));
      report->volume = vf;
      for (int q = 0; q < 9; q++) report->C[q] = C[q];
      report->unconverged_solves = unconverged;
    }
    return OTM_OK;
  } catch (const Fail& f) {
    return f.code;  // the failing engine call has set otm_last_error
  } catch (const std::exception&) {
    return OTM_E_ARG;
  }
}

}  // extern "C"
