// Prints the decisions of the drop-in ConvergeChecker and VolumeGovernor
// (include/opentm/opentm.hpp) on fixed scenarios; tests/test_abi.py checks
// them against hand-computed values. Host code only.
#include <cstdio>

#include "opentm/opentm.hpp"

int main() {
  {  // constant objective: converged from iteration 12
    homo::ConvergeChecker c(1e-3);
    std::printf("A");
    for (int it = 0; it < 14; it++) std::printf(" %d", c.is_converge(it, 2.0) ? 1 : 0);
    std::printf("\n");
  }
  {  // one jump at iteration 11: three calm steps needed after it
    homo::ConvergeChecker c(1e-3);
    std::printf("B");
    for (int it = 0; it < 16; it++) std::printf(" %d", c.is_converge(it, it == 11 ? 1.5 : 1.0) ? 1 : 0);
    std::printf("\n");
  }
  {  // slow drift below the threshold
    homo::ConvergeChecker c(1e-3);
    std::printf("D");
    for (int it = 0; it < 13; it++) std::printf(" %d", c.is_converge(it, 1.0 + 1e-4 * it) ? 1 : 0);
    std::printf("\n");
  }
  {  // negative objective cuts the bound, then five stalled steps widen it
    homo::VolumeGovernor g;
    const int r0 = g.volume_check(-0.5f, 0.2f, 0.9f, 0);
    std::printf("G1 %d %.7g %.7g\n", r0, g.get_volume_bound(), g.get_current_decrease());
    for (int it = 1; it <= 5; it++) {
      const int r = g.volume_check(0.3f, 0.2f, 0.2f, it);
      std::printf("G1 %d %.7g %.7g\n", r, g.get_volume_bound(), g.get_current_decrease());
    }
  }
  {  // oscillating objective away from the bound: cut after 8 sign changes
    homo::VolumeGovernor g;
    for (int it = 0; it < 10; it++) {
      const int r = g.volume_check(it % 2 ? 0.6f : 0.5f, 0.1f, 0.9f, it);
      std::printf("G2 %d %.7g %.7g\n", r, g.get_volume_bound(), g.get_current_decrease());
    }
  }
  {  // negative objective every step: the cut factor runs out on the 14th check
    homo::VolumeGovernor g;
    for (int it = 0; it < 14; it++) std::printf("G3 %d\n", g.volume_check(-1.0f, 0.0f, 0.5f, it));
  }
  {  // iteration 100 forces a cut
    homo::VolumeGovernor g;
    for (int it = 98; it <= 100; it++) g.volume_check(0.5f, 0.25f, 0.9f, it);
    std::printf("G4 %.7g\n", g.get_volume_bound());
  }
  return 0;
}
