// Compiles code written against the reference's C++ API (cfg::HomoConfig,
// homo::Homogenization_H / OCOptimizer / MMAOptimizer, ConvergeChecker,
// VolumeGovernor, runInstance, wrapper_homo) against the engine's headers.
// Without a GPU the constructors must throw std::runtime_error; on a
// machine with a GPU they construct, and the program says so (the GPU
// behaviour itself is tested by tests/cxx/dropin_gpu.cpp).
#include <cstdio>
#include <stdexcept>

#include "opentm/opentm.hpp"

int main(int argc, char** argv) {
  cfg::HomoConfig config;
  config.init();
  config.reso[0] = config.reso[1] = config.reso[2] = 16;
  homo::ConvergeChecker criteria(config.finthres);
  homo::VolumeGovernor governor;
  for (int itn = 0; itn < 3; itn++) {
    criteria.is_converge(itn, 1.0 - 0.1 * itn);
    governor.volume_check(0.5f, 0.2f, 0.3f, itn);
  }
  (void)argc;
  (void)argv;
  try {
    homo::Homogenization_H hom(config);
    double C[3][3];
    (void)C;
    homo::OCOptimizer oc(16 * 16 * 16, 0.001f, 0.02f, 0.5f);
    homo::MMAOptimizer mma(1, 16 * 16 * 16, 1, 0, 1e7, 1);
    mma.setBound(1e-4f, 1.f);
    wrapper_homo w;
    w.setConfig(16, {1.0, 1e-4}, {0.3, 0.2, 0.1, 0, 0, 0}, cfg::Model::oc);
    std::printf("constructed with a GPU\n");
    return 0;
  } catch (const std::runtime_error& e) {
    std::printf("runtime_error: %s\n", e.what());
    return 0;
  }
}
