/* FP64 build of the oracle's multigrid pieces (mg_impl.h): orc_gs_ell,
 * orc_brick_gs_ell, orc_galerkin_ell, orc_mg_homogenize. Test
 * infrastructure only (see oracle.h). */
#include "oracle.h"
#define REAL double
#define FN(name) orc_##name
#include "mg_impl.h"
