// Prints value and gradient of a composite objective built with
// include/opentm/objective.hpp at the point given on the command line
// (9 tensor entries); the test compares them with central differences.
#include <stdio.h>
#include <stdlib.h>

#include "opentm/objective.hpp"

using homo::ad::Var;

static Var objective(const homo::ad::Tensor3& c) {
  Var f = (c[0][0] - 0.3).pow(2) + (c[1][1] - 0.2).pow(2) + (c[2][2] - 0.1).pow(2);
  f = f.log() + 0.1 * c[0][1].abs() + c[1][2].sgm(4.0, 0.05) * c[2][0].exp() / (c[0][0] + 2.0);
  return f - 1e-4;
}

int main(int argc, char** argv) {
  if (argc != 10) return 2;
  double C[9];
  for (int q = 0; q < 9; q++) C[q] = atof(argv[1 + q]);
  auto c = homo::ad::heat_tensor(C);
  const double v = objective(c).backward();
  double dC[9];
  homo::ad::tensor_grad(c, dC);
  printf("%.17g", v);
  for (int q = 0; q < 9; q++) printf(" %.17g", dC[q]);
  printf("\n");
  return 0;
}
