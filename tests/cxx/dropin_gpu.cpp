// Drives the C++ drop-in classes (include/opentm/opentm.hpp) on a GPU the
// way the reference's callers do, with device memory from hipMalloc:
//  - homo::Homogenization_H on a 32^3 conductivity held with a row pitch of
//    128 floats (the TensorBuffer pitch of a 32^3 cell, TensorExpression.h:54):
//    update(rho, pitchT), heatMatrix, Sensitivity(dC, sens, pitchT, true);
//  - its mg_ (MG_H::solveEquation) and grid (Grid_H::writeDensity);
//  - homo::MMAOptimizer::update with x, dfdx, gval and dgdx in device memory
//    (Framework.cu:186-196).
// Usage: dropin_gpu <dir>. Reads <dir>/kappa.bin (binary density format)
// and <dir>/mma_in.bin (x, dfdx, dgdx: n floats each, then gval: 1 float);
// writes C.txt, sens.bin (pitched), mg.txt, kappa_out.bin, mma_x.bin.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "opentm/opentm.hpp"

#define HIP_OK(x)                                                           \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
      std::exit(2);                                                         \
    }                                                                       \
  } while (0)

static std::vector<float> read_floats(const std::string& path, size_t n) {
  std::vector<float> v(n);
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f || std::fread(v.data(), sizeof(float), n, f) != n) {
    std::fprintf(stderr, "cannot read %s\n", path.c_str());
    std::exit(2);
  }
  std::fclose(f);
  return v;
}

static void write_floats(const std::string& path, const float* p, size_t n) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(p, sizeof(float), n, f) != n) std::exit(3);
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 2) return 1;
  const std::string dir = argv[1];
  try {
    int reso[3];
    if (otm_read_density((dir + "/kappa.bin").c_str(), reso, nullptr, OTM_HOST) != OTM_OK) return 2;
    const int n = reso[0];
    const int pitch = 128;
    std::vector<float> k((size_t)n * n * n);
    homo::check(otm_read_density((dir + "/kappa.bin").c_str(), reso, k.data(), OTM_HOST));
    // conductivity in a pitched device buffer, as a TensorBuffer holds it
    float* dk = nullptr;
    HIP_OK(hipMalloc(&dk, sizeof(float) * (size_t)pitch * n * n));
    HIP_OK(hipMemset(dk, 0, sizeof(float) * (size_t)pitch * n * n));
    HIP_OK(hipMemcpy2D(dk, sizeof(float) * pitch, k.data(), sizeof(float) * n, sizeof(float) * n, (size_t)n * n,
                       hipMemcpyHostToDevice));

    cfg::HomoConfig config;
    config.init();
    config.reso[0] = config.reso[1] = config.reso[2] = n;
    config.heatRatio[0] = 1.0;
    config.heatRatio[1] = 1e-3;
    config.femRelThres = 1e-7;
    homo::Homogenization_H hom(config);
    hom.update(dk, pitch);
    double C[3][3];
    hom.heatMatrix(C);
    FILE* fc = std::fopen((dir + "/C.txt").c_str(), "w");
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) std::fprintf(fc, "%.17g\n", C[i][j]);
    std::fclose(fc);

    float dC[3][3] = {{-1.f, 0.f, 0.f}, {0.f, -1.f, 0.f}, {0.f, 0.f, -1.f}};
    float* ds = nullptr;
    HIP_OK(hipMalloc(&ds, sizeof(float) * (size_t)pitch * n * n));
    HIP_OK(hipMemset(ds, 0x7f, sizeof(float) * (size_t)pitch * n * n));  // overwritten with zeros in the padding
    hom.Sensitivity(dC, ds, pitch, true);
    std::vector<float> sens((size_t)pitch * n * n);
    HIP_OK(hipMemcpy(sens.data(), ds, sizeof(float) * sens.size(), hipMemcpyDeviceToHost));
    write_floats(dir + "/sens.bin", sens.data(), sens.size());

    // the hierarchy through MG_H: a warm solve to a tighter tolerance
    const double rr = hom.mg_->solveEquation(1e-9, true);
    FILE* fm = std::fopen((dir + "/mg.txt").c_str(), "w");
    std::fprintf(fm, "%d %.6e\n", (int)hom.mg_->grids.size(), rr);
    for (auto& g : hom.mg_->grids) std::fprintf(fm, "%d %d %d\n", g->cellReso[0], g->cellReso[1], g->cellReso[2]);
    std::fclose(fm);
    hom.grid->writeDensity(dir + "/kappa_out.bin", homo::binary);

    // MMA with every argument in device memory
    const size_t ne = (size_t)n * n * n;
    std::vector<float> in = read_floats(dir + "/mma_in.bin", 3 * ne + 1);
    float *dx = nullptr, *df = nullptr, *dg = nullptr, *gv = nullptr;
    HIP_OK(hipMalloc(&dx, sizeof(float) * ne));
    HIP_OK(hipMalloc(&df, sizeof(float) * ne));
    HIP_OK(hipMalloc(&dg, sizeof(float) * ne));
    HIP_OK(hipMalloc(&gv, sizeof(float)));
    HIP_OK(hipMemcpy(dx, in.data(), sizeof(float) * ne, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(df, in.data() + ne, sizeof(float) * ne, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dg, in.data() + 2 * ne, sizeof(float) * ne, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(gv, in.data() + 3 * ne, sizeof(float), hipMemcpyHostToDevice));
    homo::MMAOptimizer mma(1, (int)ne, 1, 0, 1e7, 1);
    mma.setBound(1e-4f, 1.f);
    float* dgdx[1] = {dg};
    mma.update(0, dx, df, gv, dgdx);
    std::vector<float> x(ne);
    HIP_OK(hipMemcpy(x.data(), dx, sizeof(float) * ne, hipMemcpyDeviceToHost));
    write_floats(dir + "/mma_x.bin", x.data(), ne);
    HIP_OK(hipFree(dx));
    HIP_OK(hipFree(df));
    HIP_OK(hipFree(dg));
    HIP_OK(hipFree(gv));
    HIP_OK(hipFree(ds));
    HIP_OK(hipFree(dk));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 4;
  }
  std::printf("dropin ok\n");
  return 0;
}
