/*
 * opentm.h -- C ABI of the MI355X (gfx950) thermal inverse-homogenization
 * engine. Plain pointers and sizes only; every call returns an otm_status.
 *
 * The reference (quanyuchen2000/OPENTM) has no C ABI: its boundary is the
 * C++ classes homo::Homogenization_H, homo::MG_H, homo::OCOptimizer and the
 * (commented-out) pybind module openTM. Each entry point below names the
 * reference interface it replaces. C++ classes with the reference's names
 * sit on top of this ABI in include/opentm/opentm.hpp.
 *
 * Conventions
 *  - Scalar fields are lexicographic with x fastest; a row pitch (in
 *    floats, >= reso[0]) may be given where the reference takes pitchT.
 *  - `where` says whether a pointer is host (OTM_HOST) or device
 *    (OTM_DEVICE) memory. Device pointers must be allocated on the
 *    context's device.
 *  - The homogenized tensor uses the reference's load-case order: load case
 *    0 is a unit temperature gradient along +y of the grid, 1 along -x,
 *    2 along +z (the element node remap of templateMatrix.cpp:74-82).
 *  - A context owns one MG hierarchy and one HIP stream; calls on
 *    different contexts are independent and may run from different host
 *    threads. Calls on one context must not overlap.
 */
#ifndef OPENTM_H
#define OPENTM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OTM_API __attribute__((visibility("default")))

typedef struct otm_ctx otm_ctx;
typedef struct otm_oc otm_oc;

typedef enum otm_status {
  OTM_OK = 0,
  OTM_E_ARG = -1,          /* invalid argument (sizes, pointers, flags) */
  OTM_E_HIP = -2,          /* HIP runtime error */
  OTM_E_NOCONV = -3,       /* solver did not reach the tolerance */
  OTM_E_UNSUPPORTED = -4,  /* resolution or option not supported */
  OTM_E_NUMERIC = -5,      /* NaN/Inf in the solve (reference: "MG numerical explode", MG.cpp:116) */
  OTM_E_NODEV = -6,        /* no HIP device */
  OTM_E_STATE = -7,        /* call order violated (e.g. solve before update) */
  OTM_E_OOM = -8           /* device allocation failed */
} otm_status;

enum { OTM_HOST = 0, OTM_DEVICE = 1, OTM_ANY = 2 /* otm_memcpy only: host, device or managed, found from the pointer */ };

/* Solver configuration. Mirrors cfg::HomoConfig (cmdline.h:12-33) for the
 * fields the solve uses, plus the multigrid knobs the reference hard-codes. */
typedef struct otm_config {
  int reso[3];              /* HomoConfig.reso: cells per axis (even, >= 8) */
  double heat_ratio[2];     /* HomoConfig.heatRatio: kappa at rho=1 and rho=0 */
  double fem_rel_thres;     /* HomoConfig.femRelThres: relative residual target */
  int max_cg_iter;          /* PCG iteration cap per solve */
  int pre_sweeps;           /* Gauss-Seidel sweeps before restriction */
  int post_sweeps;          /* Gauss-Seidel sweeps after prolongation */
  int min_coarse;           /* stop coarsening once a level edge would fall below this */
  int device;               /* HIP device ordinal */
  int verbose;              /* print per-solve statistics to stderr */
  int coarse_sweeps;        /* sweeps on levels >= 1 (0: same as the fine level) */
  int fp64_refine;          /* 1: FP64 iterate with FP32 inner PCG solves (iterative refinement) */
  int fp16_stencils;        /* 1: coarse stencils stored in FP16 with a per-level scale */
  int exact_smoother;       /* 1: exact multicolour Gauss-Seidel on every level instead of the
                               LDS brick smoother (the fallback a solve takes by itself when
                               the brick preconditioner loses positivity) */
  int graph_vcycle;         /* 1 (default): each V-cycle's fixed launch sequence is captured
                               once as a HIP graph and replayed (0: launched one by one) */
  int fine_tile;            /* fine brick smoother tile shape: -1 (default) 32x4x8, or 16x4x8
                               where x is not a multiple of 32; 0..4 a variant of
                               otm_fine_tile_shape (each is its own block Gauss-Seidel) */
} otm_config;

/* cfg::HomoConfig::init (cmdline.cpp:4-35) defaults where they apply;
 * fem_rel_thres defaults to 1e-6 (the reference's 1e-2 stationary-MG
 * threshold is not accurate enough for kappa^H parity). */
OTM_API void otm_config_init(otm_config* cfg);

/* Message of the last failing call on this host thread. */
OTM_API const char* otm_last_error(void);
OTM_API int otm_version(void);

/* Homogenization_H(cfg::HomoConfig) / Homogenization_H::build
 * (homogenization.h:15-32, homogenization.cpp:31-63): allocates the grid
 * hierarchy and element constants. */
OTM_API int otm_create(const otm_config* cfg, otm_ctx** out);
OTM_API int otm_destroy(otm_ctx* ctx);

/* Number of multigrid levels and the edge lengths of level `l`
 * (MG_H::grids, MG.h:14-32). */
OTM_API int otm_levels(const otm_ctx* ctx, int* nlevels);
OTM_API int otm_level_reso(const otm_ctx* ctx, int l, int reso[3]);

/* Homogenization_H::update(float* rho, int pitchT) (homogenization.cpp:76-81):
 * takes the SIMP-mapped conductivity per element (pitch >= reso[0] floats
 * per row) and rebuilds the coarse Galerkin operators. kappa == NULL only
 * rebuilds the operators. */
OTM_API int otm_update(otm_ctx* ctx, const float* kappa, int pitch, int where);

/* Homogenization_H::heatMatrix(double C[3][3]) (homogenization.cu:180-252):
 * solves the three load cases (warm-started from the previous solution)
 * and returns the symmetric 3x3 homogenized conductivity, row-major.
 * OTM_E_NOCONV: the solve ended above fem_rel_thres (max_cg_iter reached
 * or the true residual stopped improving); C, the fluctuation fields and
 * otm_solve_stats are still written (the reference reports "Failed to
 * converge" and carries on, MG.cpp:101-118). OTM_E_NUMERIC: NaN/Inf. */
OTM_API int otm_heat_matrix(otm_ctx* ctx, double C[9]);

/* Fine smoother tile shapes: variant 0..4 -> shape = {bx, by, bz, threads};
 * otm_fine_tile gives the context's variant (-1: exact colour sweeps). */
OTM_API int otm_fine_tile_shape(int variant, int shape[4]);
OTM_API int otm_fine_tile(const otm_ctx* ctx, int* variant);

/* Autotuner: no counterpart in the reference, whose launch shapes are fixed
 * (culib/lib.cuh:71-74). Times one full solve of the current design (after
 * otm_update) for each fine tile shape the cell admits and coarse sweep
 * counts 1 and 2, each from the same start (warm = 0: zero; warm = 1: the
 * current iterate, which after otm_apply_design of a design loop's next
 * design is the warm start its solve would take), keeps the fastest that
 * reaches the tolerance without falling back to exact sweeps in the context
 * (the iterate is restored, otm_heat_matrix must run again), and reports
 * every candidate. The choice can be pinned with otm_config.fine_tile and
 * coarse_sweeps. OTM_E_UNSUPPORTED when the fine level has no brick smoother. */
#define OTM_TUNE_MAX 16
typedef struct otm_tune_report {
  int n;                          /* candidates timed */
  int fine_tile[OTM_TUNE_MAX];    /* fine tile variant of candidate i */
  int coarse_sweeps[OTM_TUNE_MAX];
  double ms[OTM_TUNE_MAX];        /* solve time (the second of two runs) */
  int iters[OTM_TUNE_MAX];        /* PCG iterations (max over the load cases) */
  int ok[OTM_TUNE_MAX];           /* reached the tolerance with the brick preconditioner */
  int chosen;                     /* index of the kept candidate */
} otm_tune_report;
OTM_API int otm_autotune(otm_ctx* ctx, int warm, otm_tune_report* report);

/* PCG iterations and achieved relative residual of the last solve. */
OTM_API int otm_solve_stats(const otm_ctx* ctx, int iters[3], double relres[3]);

/* Homogenization_H::Sensitivity(float dC[3][3], float* sens, int pitchT,
 * bool lexiOrder=true) (homogenization.cu:415-479): per-element
 * d(sum_ij dC_ij kappa^H_ij)/d(kappa_e), written lexicographically with
 * row pitch `pitch`. Requires a prior otm_heat_matrix. */
OTM_API int otm_sensitivity(otm_ctx* ctx, const double dC[9], float* sens, int pitch, int where);

/* The six element energies S_e(i,j) = d kappa^H_ij / d kappa_e in the order
 * (00,01,02,11,12,22), each an array of reso[0]*reso[1]*reso[2] floats. */
OTM_API int otm_element_energy(otm_ctx* ctx, float* S6, int where);

/* The three fluctuation fields chi_k (3 * nodes floats, load case major),
 * mean-zero. Setting them warm-starts the next solve (Grid_H::useUchar,
 * grid.cpp:318-321). */
OTM_API int otm_get_chi(otm_ctx* ctx, float* chi, int where);
OTM_API int otm_set_chi(otm_ctx* ctx, const float* chi, int where);

/* MG_H::reset_displacement (MG.h:14-32): zero the warm start. */
OTM_API int otm_reset_displacement(otm_ctx* ctx);

/* Level operators, for tests and custom drivers: q = A_l u (three vectors
 * of level l, 3*nodes_l floats) and z = M^{-1} r (one V-cycle of the
 * preconditioner on the fine level; MG_H::v_cycle, MG.cpp:46-71). */
OTM_API int otm_apply_level(otm_ctx* ctx, int level, const float* u, float* q, int where);
/* The fine operator as the PCG applies it (MG_H::solveEquation's K p,
 * MG.cpp:80-142): q = K u for three vectors of the fine level in the
 * difference form, and the three dot products u.q (pq, may be NULL). */
OTM_API int otm_apply_fine_pcg(otm_ctx* ctx, const float* u, float* q, double pq[3], int where);
OTM_API int otm_vcycle(otm_ctx* ctx, const float* r, float* z, int where);
/* Grid_H::getDensity (grid.cpp:330-345 writes it): the element conductivity
 * last given to otm_update (to one FP32 rounding), lexicographic, x fastest. */
OTM_API int otm_get_kappa(otm_ctx* ctx, float* kappa, int where);
/* Level-l 27-point stencil, 27*nodes_l floats (offset major), l >= 1. */
OTM_API int otm_get_stencil(otm_ctx* ctx, int level, float* st, int where);

/* MG_H::solveEquation(double tol, bool with_guess) (MG.cpp:80-142) for the
 * three load cases together: MG-preconditioned CG to relative residual
 * rtol (<= 0: the context's fem_rel_thres) from the current fluctuation
 * fields (warm != 0) or from zero, without kappa^H. relres (may be NULL)
 * receives the final true relative residuals. OTM_E_NOCONV as for
 * otm_heat_matrix. */
OTM_API int otm_solve(otm_ctx* ctx, double rtol, int warm, double relres[3]);

/* MG_H::v_cycle(w, pre, post) (MG.cpp:46-71) iterated the way the
 * reference iterates it (MG_H::solveEquation's loop): `cycles` stationary
 * steps chi += M^{-1}(b - K chi) of the three load cases, with this
 * engine's V-cycle as M^{-1}. relres (may be NULL) receives the relative
 * residual measured at the start of the last cycle (the reference's check
 * lags one cycle the same way, MG.cpp:100). */
OTM_API int otm_mg_cycles(otm_ctx* ctx, int cycles, double relres[3]);

/* Grid_H::restrict_residual / prolongate_correction (grid.cu:811-940):
 * direction 0 restricts three level-(level-1) vectors `in` onto level
 * `level` (out = P^T in); direction 1 adds the prolongation of three
 * level-`level` vectors `in` to the level-(level-1) vectors `out`
 * (out += P in). Trilinear P, 1 <= level < levels. */
OTM_API int otm_transfer(otm_ctx* ctx, int level, int direction, const float* in, float* out, int where);

/* Grid_H::gs_relaxation (grid.cu:943-987) on level `level` (< levels-1)
 * for three right-hand sides f, in place on u (3*nodes_l floats each).
 * mode 0: exact multicolour Gauss-Seidel, colours in this engine's forward
 * order (the reverse of the reference's colour order 0..7); mode 1: the
 * reference's order; mode 2 / 3: the LDS brick smoother in those orders
 * (exact inside each brick, neighbours outside it frozen at the entry u),
 * the one the preconditioner uses. */
OTM_API int otm_smooth_level(otm_ctx* ctx, int level, const float* f, float* u, int sweeps, int mode, int where);

/* Grid_H::solveHostEquation (grid.cpp:272-286): u = B^{-1} f on the
 * coarsest level for three right-hand sides, with B the coarsest Galerkin
 * operator plus the rank-one term that fixes its constant null space. */
OTM_API int otm_coarsest_solve(otm_ctx* ctx, const float* f, float* u, int where);

/* Radial convolution filter (conv_tsexp_t, TensorExpression.h:1095-1185;
 * kernels TensorExpression.cuh:303-384). kind 0 = Spline4, 1 = Linear;
 * periodic 0 renormalises at the cell boundary. adjoint != 0 applies the
 * transpose. Arrays are contiguous reso[0]*reso[1]*reso[2] floats. */
OTM_API int otm_conv_filter(const int reso[3], int kind, double radius, int periodic, int adjoint,
                            const float* in, float* out, int where);
/* otm_conv_filter for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops): no counterpart in the reference.
 * in and out are device arrays of the current device that must not overlap;
 * the launches go on `stream` (a hipStream_t; null = the default stream) and
 * the call returns without waiting for them. The weight sums live in a
 * static device array of 2^23 floats (32 MiB per device, kept until the
 * process ends): larger cells get OTM_E_UNSUPPORTED. Every call on a device
 * shares that array, so calls whose launches overlap on one GPU race. */
OTM_API int otm_conv_filter_noalloc(const int reso[3], int kind, double radius, int periodic, int adjoint,
                                    const float* in, float* out, void* stream);

/* Design-density pipeline of runCustom (Framework.cu:163):
 * kappa = k1 + (k0-k1) * filter(rho)^3 with the Spline4 r=1.5 filter. */
typedef struct otm_density_params {
  int filter_kind;       /* 0 Spline4, 1 Linear, -1 none */
  double filter_radius;  /* 1.5 in runCustom */
  int filter_periodic;   /* 0 in runCustom (TensorExpression.h:1083) */
  double penal;          /* SIMP exponent; only 3 is supported */
} otm_density_params;
OTM_API void otm_density_params_init(otm_density_params* p);

/* Design density held by the context (var_tsexp_t rho, Framework.cu:155). */
OTM_API int otm_set_design(otm_ctx* ctx, const float* rho, int where);
OTM_API int otm_get_design(otm_ctx* ctx, float* rho, int where);

/* objective.eval() path for the density expression and heat_tensor_t
 * (homoExpression.h:89-105): filter + SIMP + update + heatMatrix. */
/* The first half of otm_eval: filter, SIMP map and operator rebuild for
 * the current design, no solve (the iterate stays the warm start). */
OTM_API int otm_apply_design(otm_ctx* ctx, const otm_density_params* p);
OTM_API int otm_eval(otm_ctx* ctx, const otm_density_params* p, double C[9]);

/* objective.backward() (homoExpression.h:106-120): sensitivity with dC,
 * chained through SIMP and the filter adjoint into d obj / d rho, kept in
 * the context; copied to drho if it is non-NULL (the call then returns
 * with drho written). With drho NULL the gradient stays in the context's
 * stream order for the calls that use it (otm_oc_step,
 * otm_symmetrize_design) and the call returns without waiting. */
OTM_API int otm_backward(otm_ctx* ctx, const double dC[9], float* drho, int where);

/* homo::OCOptimizer (OCOptimizer.h:7-20). */
OTM_API int otm_oc_create(int ne, double min_rho, double step_limit, double damp, int device, otm_oc** out);
OTM_API int otm_oc_destroy(otm_oc* oc);
/* Bisection control (the reference hard-codes 20 steps and 1e-4). */
OTM_API int otm_oc_set_bisection(otm_oc* oc, int max_steps, double vol_tol);
/* OCOptimizer::filterSens (OCOptimizer.cu:134-152): periodic, density
 * weighted, linear kernel of the given radius; in place on sens. */
OTM_API int otm_oc_filter_sens(otm_oc* oc, float* sens, const float* rho, int pitch, const int reso[3],
                               double radius, int where);
/* OCOptimizer::update (OCOptimizer.cu:38-67): Lagrange bisection, in place
 * on rho (ne contiguous floats). Reports bisection steps and final volume. */
OTM_API int otm_oc_update(otm_oc* oc, const float* sens, float* rho, double volratio, int where, int* steps,
                          double* volume);

/* One OC design iteration on the context's design density, using its
 * current d obj / d rho (after otm_backward): filterSens + update, on the
 * context's stream (the optimizer's buffers are used there; calls on one
 * optimizer must not overlap). Returns once the bisection has ended; the
 * updated density is in the context's stream order for the next call. */
OTM_API int otm_oc_step(otm_ctx* ctx, otm_oc* oc, double filter_radius, double volratio, int* steps,
                        double* volume);

/* homo::MMAOptimizer (mmaOptimizer.h:12-35): Svanberg's MMA in FP64 on the
 * device (mmaOpt.cu:834-1019 with the interior-point subsolver
 * mmaOpt.cu:429-830). m <= 4 constraints, n design variables. */
typedef struct otm_mma otm_mma;
/* MMAOptimizer(int m, int n, double a0, double a, double c, double d) */
OTM_API int otm_mma_create(int m, int n, double a0, double a, double c, double d, int device, otm_mma** out);
OTM_API int otm_mma_destroy(otm_mma* mma);
/* MMAOptimizer::setBound(float xmin, float xmax) and the array overload */
OTM_API int otm_mma_set_bounds(otm_mma* mma, double xmin, double xmax);
OTM_API int otm_mma_set_bounds_array(otm_mma* mma, const float* xmin, const float* xmax, int where);
/* Move limit of the subproblem (MMAOptimizer::update passes 1). */
OTM_API int otm_mma_set_move(otm_mma* mma, double move);
/* MMAOptimizer::update(int itn, float* x, float* dfdx, float* gval,
 * float** dgdx) (mmaOptimizer.cu:61-77): x is updated in place; dfdx is the
 * objective gradient, gval the m constraint values (host), dgdx m pointers
 * to constraint gradients. x, dfdx and dgdx[i] are `where` memory.
 * *newton receives the number of interior-point Newton steps. */
OTM_API int otm_mma_update(otm_mma* mma, int itn, float* x, const float* dfdx, const double* gval,
                           const float* const* dgdx, int where, int* newton);

/* HIP-event timing of the MMA passes (no counterpart in the reference):
 * kclass 0 = later trial-point passes of a line search, 1 = the first trial
 * of each Newton step (it also forms the elements' step bound), 2 =
 * start-of-level point pass; launches, milliseconds and algorithmic bytes
 * since enabling. */
OTM_API int otm_mma_profile_enable(otm_mma* mma, int enable);
OTM_API int otm_mma_profile_read(otm_mma* mma, int kclass, long long* launches, double* ms, double* bytes);

/* initDensity (Framework.cu:61-145) on the context's design, indexed like
 * cfg::InitWay: 0 random, 1 randcenter, 2 noise, 3 manual, 4 interp,
 * 5 rep_randcenter, 6 P, 7 G, 8 D, 9 IWP, 10 example.
 *  - random / randcenter / rep_randcenter: randTri (Framework.cuh:238-310),
 *    a random trigonometric series of 6P + 36P^2 terms (P = n_period; plus a
 *    random rotation of the centred coordinates for randcenter) mapped
 *    through sigmoid(., 15, t) with t bisected to the volume ratio;
 *    coefficients uniform in [-1, 1] from `seed` (the reference draws them
 *    from a clock-seeded cuRAND) or, for rep_randcenter, given in coeffs;
 *  - noise (Framework.cu:75-88): uniform values from `seed` (a counter-based
 *    generator), symmetrised, projected to {1e-9, 1}, scaled to the volume
 *    ratio, clamped to [0.001, 1];
 *  - P, G, D, IWP, example: the reference's closed forms;
 *  - manual / interp read OpenVDB files: OTM_E_UNSUPPORTED.
 * Then the field is symmetrised and clamped to [0.1, 1] (Framework.cu:
 * 140-144) and becomes the context's design. coeffs_out (host, may be
 * NULL) receives the randTri coefficients used (the reference writes them
 * to a file that rep_randcenter reads back). */
typedef struct otm_init_params {
  int initway;          /* cfg::InitWay */
  int seed;             /* random starts */
  double volratio;      /* HomoConfig.volRatio */
  int n_period;         /* HomoConfig.initperiod, 1..10 (default 10) */
  int sym;              /* OTM_SYM_* (HomoConfig.sym) */
  const float* coeffs;  /* randTri coefficients (host), NULL: drawn from seed */
  int ncoeffs;
} otm_init_params;
OTM_API void otm_init_params_init(otm_init_params* p);
OTM_API int otm_init_design(otm_ctx* ctx, const otm_init_params* p, float* coeffs_out);
/* otm_init_design with n_period 10, no symmetry and seeded coefficients. */
OTM_API int otm_init_density(otm_ctx* ctx, int initway, int seed, double volratio);

/* cfg::Symmetry (cmdline.h:9), same order. */
enum { OTM_SYM_REFLECT3 = 0, OTM_SYM_REFLECT6 = 1, OTM_SYM_ROTATE3 = 2, OTM_SYM_ROTATE2 = 3, OTM_SYM_NONE = 4 };
/* symmetrizeField (Framework.cuh:221-235): replaces every value by the mean
 * over its orbit under the symmetry group about the cell centre (reflect6
 * and rotate3 need a cubic cell). In place, host or device field. */
OTM_API int otm_symmetrize(const int reso[3], int sym, float* field, int where);
/* The same on the context's design density and, after otm_backward, on its
 * gradient (runCustom's OC branch, Framework.cu:223-224). */
OTM_API int otm_symmetrize_design(otm_ctx* ctx, int sym);

/* Grid_H::writeDensity binary format (grid.cpp:330-345): int32 reso[3],
 * then reso[0]*reso[1]*reso[2] float32 values, x fastest. otm_read_density
 * with rho == NULL only reads the resolution. */
OTM_API int otm_write_density(const char* path, const int reso[3], const float* rho, int where);
OTM_API int otm_read_density(const char* path, int reso[3], float* rho, int where);

/* Mean design density and mean SIMP-mapped conductivity of the last
 * otm_eval (rho_H.sum()/ne and rhop_H.sum()/ne in runCustom). */
OTM_API int otm_design_stats(otm_ctx* ctx, double* mean_rho, double* mean_kappa);

/* runCustom (Framework.cu:147-244): tensor-matching design loop, objective
 * sum over (00,11,22,01,21,02) of (C_ij - target)^2 - 1e-4, with OC
 * (VolumeGovernor + ConvergeChecker) or MMA (minimise volume subject to the
 * tensor error). */
enum { OTM_MODEL_MMA = 0, OTM_MODEL_OC = 1 };
/* Iteration log of the design loop (logIter, Framework.cu:7-59): files are
 * prefix + name (getPath). Every 5th iteration "rho" (the design) and "C"
 * (the tensor); every logrho-th "rho_%04d"; every logc-th a line of "Clog";
 * every logsens-th "sens_%04d" (the objective's design gradient); every
 * logobj-th a line of "objlog". Intervals 0 switch a file off. Density and
 * gradient files use the binary format of otm_write_density (the reference
 * writes OpenVDB, which this build does not link). */
typedef struct otm_log_params {
  const char* prefix;  /* HomoConfig.outprefix; NULL: no log */
  int logrho;          /* intervals of HomoConfig.logrho, logc, logsens, logobj */
  int logc;
  int logsens;
  int logobj;
} otm_log_params;
OTM_API void otm_log_params_init(otm_log_params* p);
/* One iteration's log: C (9 doubles), the objective, and the design
 * gradient sens (reso^3 floats on `where`, NULL when logsens is off). */
OTM_API int otm_log_iteration(otm_ctx* ctx, const otm_log_params* p, int iter, const double C[9], double obj,
                              const float* sens, int where);

/* Full reductions of a field with their gradient (reduce_tsexp_t and the
 * associative ops of TensorExpression.h:1187-1251, 1342-1402): kind 0 sum,
 * 1 p-norm (sum x^p)^(1/p), 2 log-sum-exp. The value is accumulated in FP64
 * (deterministic two-stage sums); grad (optional, n floats) receives
 * upstream * d value / d x_i. x and grad on `where`. */
enum { OTM_REDUCE_SUM = 0, OTM_REDUCE_PNORM = 1, OTM_REDUCE_LOGSUMEXP = 2 };
OTM_API int otm_field_reduce(const float* x, long long n, int kind, double p, int where, double* value, float* grad,
                             double upstream);

typedef struct otm_custom_params {
  otm_config base;     /* resolution, heat ratios, solver */
  int model;           /* OTM_MODEL_MMA or OTM_MODEL_OC (HomoConfig.model) */
  int initway;         /* HomoConfig.winit, see otm_init_density */
  int seed;            /* seed of the random initialisers */
  int max_iter;        /* HomoConfig.max_iter */
  double finthres;     /* HomoConfig.finthres */
  double volratio;     /* HomoConfig.volRatio (noise initialiser) */
  double target[6];    /* HomoConfig.target_tensor: xx yy zz xy yz xz */
  int sym;             /* HomoConfig.sym: OTM_SYM_* (default OTM_SYM_NONE) */
  otm_log_params log;  /* iteration log (default: none) */
} otm_custom_params;
typedef struct otm_custom_report {
  int iterations;
  int converged;
  double objective;    /* last objective value */
  double volume;       /* mean design density */
  double C[9];         /* last homogenized tensor */
  int unconverged_solves;  /* solves that ended above fem_rel_thres (the loop carried on) */
} otm_custom_report;
OTM_API void otm_custom_params_init(otm_custom_params* p);
/* rho0: optional host initial density (reso^3 floats); rho_out: host
 * reso^3 floats, the final design; report may be NULL. */
OTM_API int otm_run_custom(const otm_custom_params* p, const float* rho0, float* rho_out, otm_custom_report* report);

/* ---- batched form: many cells per launch -------------------------------
 * No counterpart in the reference (one hierarchy per process). A batch
 * solves the heatMatrix of many contexts together: every launch covers the
 * same operation for all member cells (one fine smoother launch for all
 * fine levels, one for all levels l, one per reduction), so small cells fill
 * the GPU together. Each cell keeps its own arithmetic: results are the
 * same whatever else is in the batch. The contexts must be on the batch's
 * device and updated (otm_update / otm_eval's SIMP map); a batch call
 * synchronises their streams first. */
typedef struct otm_batch otm_batch;
OTM_API int otm_batch_create(int device, otm_batch** out);
OTM_API int otm_batch_destroy(otm_batch* batch);
/* Homogenization_H::update (homogenization.cpp:76-81) of n contexts in one
 * launch sequence: kappa[i] is cell i's SIMP-mapped conductivity (row pitch
 * `pitch` floats, <= 0: reso[0]; host or device per `where`); every cell's
 * Galerkin hierarchy and coarsest inverse are rebuilt together. */
OTM_API int otm_batch_update(otm_batch* batch, otm_ctx* const* cells, int n, const float* const* kappa, int pitch,
                             int where);
/* heatMatrix (homogenization.cu:180-252) of n contexts: C receives 9
 * doubles per cell (row-major). status (may be NULL) receives a status code
 * per cell; without it the first failure is returned. */
OTM_API int otm_batch_heat_matrix(otm_batch* batch, otm_ctx* const* cells, int n, double* C, int* status);
/* Sensitivity (homogenization.cu:415-479) of n solved contexts: dC holds 9
 * doubles per cell, sens[i] receives cell i's field with row pitch `pitch`
 * floats (<= 0: reso[0]), host or device pointers per `where`. */
OTM_API int otm_batch_sensitivity(otm_batch* batch, otm_ctx* const* cells, int n, const double* dC,
                                  float* const* sens, int pitch, int where);
/* Launches issued and PCG steps taken by the batch so far. */
OTM_API int otm_batch_stats(const otm_batch* batch, long long* launches, int* steps);

/* ---- serving form: continuous batching -----------------------------------
 * No counterpart in the reference. A server owns a worker thread and one
 * batch. Submitted requests are admitted between PCG steps while the
 * member count (max_active) and the HBM budget (bytes of resident contexts,
 * active and cached) allow; hierarchies stay resident and are reused by
 * later requests of the same shape. Results do not depend on what else was
 * in flight. */
typedef struct otm_server otm_server;
typedef struct otm_request {
  otm_config cfg;           /* cell shape and solver settings (device must be the server's) */
  const float* kappa;       /* host, reso[0]*reso[1]*reso[2] SIMP-mapped conductivities (copied at submit) */
  double dC[9];             /* objective gradient for the sensitivity */
  int want_sens;            /* 1: compute the contracted sensitivity */
} otm_request;
typedef struct otm_result {
  int status;               /* OTM_OK or the failure code */
  double C[9];              /* kappa^H */
  int iters[3];             /* PCG iterations per load case */
  double relres[3];         /* final true relative residuals */
  double queue_ms;          /* submit -> admission */
  double latency_ms;        /* submit -> result */
} otm_result;
typedef struct otm_server_counters {
  long long completed;         /* requests finished */
  long long steps;             /* batched PCG steps run */
  long long contexts_created;  /* hierarchies allocated */
  long long contexts_evicted;  /* idle hierarchies freed to fit the budget */
  int active_peak;             /* most requests on the GPU at once */
  int queued;                  /* requests waiting for admission */
  double resident_bytes;       /* device bytes of held hierarchies */
  double budget_bytes;
} otm_server_counters;
OTM_API void otm_request_init(otm_request* r);
OTM_API int otm_server_create(int device, double hbm_budget_bytes, int max_active, otm_server** out);
OTM_API int otm_server_destroy(otm_server* server);
OTM_API int otm_server_submit(otm_server* server, const otm_request* req, long long* id);
/* Blocks until request `id` is done; sens (host, may be NULL) receives the
 * sensitivity field (x fastest) when it was requested. Returns the
 * request's status. */
OTM_API int otm_server_wait(otm_server* server, long long id, otm_result* res, float* sens);
OTM_API int otm_server_stats(otm_server* server, otm_server_counters* st);

/* Live kernel timing with HIP events recorded on the context's stream
 * around every launch of a kernel class. Classes: */
enum {
  OTM_PROF_FINE_GS = 0,      /* fine-level Gauss-Seidel colour sweep */
  OTM_PROF_FINE_APPLY = 1,   /* fine-level operator apply + p.q (PCG) */
  OTM_PROF_FINE_RESID = 2,   /* fine-level residual (V-cycle and PCG start) */
  OTM_PROF_COARSE_GS = 3,    /* stored-stencil Gauss-Seidel colour sweep */
  OTM_PROF_COARSE_RESID = 4, /* stored-stencil residual */
  OTM_PROF_TRANSFER = 5,     /* restriction and prolongation */
  OTM_PROF_COARSEST = 6,     /* coarsest dense solve */
  OTM_PROF_VECTOR = 7,       /* PCG vector updates and dots */
  OTM_PROF_ENERGY = 8,       /* kappa^H, sensitivity, element energies */
  OTM_PROF_SETUP = 9,        /* Galerkin assembly and coarsest factorisation */
  OTM_PROF_NCLASSES = 10
};
/* enable is a bit mask of classes to time (bit k = class k; 0 stops
 * timing); enabling clears the totals. Timed launches carry one event
 * pair each, so time only the classes of interest. */
OTM_API int otm_profile_enable(otm_ctx* ctx, int enable);
/* Launch count, summed kernel time (ms) and algorithmic bytes moved by the
 * launches of one class since timing was enabled. */
OTM_API int otm_profile_read(otm_ctx* ctx, int kclass, long long* launches, double* ms, double* bytes);

/* Device memory helpers for callers without their own HIP allocator. */
OTM_API int otm_device_alloc(int device, size_t bytes, void** ptr);
OTM_API int otm_device_free(void* ptr);
/* dst_where / src_where: OTM_HOST, OTM_DEVICE, or OTM_ANY when the caller
 * does not know (e.g. a pointer handed over by reference code that may be
 * managed memory); any OTM_ANY lets the runtime find out from the pointers. */
OTM_API int otm_memcpy(void* dst, const void* src, size_t bytes, int dst_where, int src_where);
OTM_API int otm_synchronize(otm_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
