/* FP32 build of the oracle's multigrid pieces (mg_impl.h), the CPU port of
 * the GPU's plain FP32 MG-PCG timed as the CPU baseline: orc32_gs_ell,
 * orc32_brick_gs_ell, orc32_galerkin_ell, orc32_mg_homogenize (FP32
 * vectors and operators, FP64 inner products and coarsest solve). */
#include "oracle.h"
#define REAL float
#define FN(name) orc32_##name
#include "mg_impl.h"
