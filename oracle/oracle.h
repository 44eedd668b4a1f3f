/*
 * oracle.h -- FP64 CPU restatement of the OpenTM thermal homogenization path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it. The product (opentm_amd/, include/opentm.h) never calls it.
 *
 * Every function restates the method of the reference (quanyuchen2000/OPENTM)
 * from its published behaviour; the file:line it follows is given per
 * function. Arrays are lexicographic with x fastest: index = x + nx*(y + ny*z).
 * Nodes and elements of a periodic nx*ny*nz cell are both indexed this way;
 * element (x,y,z) spans nodes (x..x+1, y..y+1, z..z+1) modulo the size.
 */
#ifndef OPENTM_ORACLE_H
#define OPENTM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Element template: Ke (8x8 row-major), feMu (8x3), disp (8x3) for an element
 * of half side `half_len` (the reference passes 1/(2N)). Local node order is
 * the reference's remapped order (templateMatrix.cpp:74-82): local index
 * l = bx + 2*by + 4*bz in grid axes. */
void orc_template(double half_len, double Ke[64], double feMu[24], double disp[24]);

/* Assembled periodic operator in ELL-27 form: A[o*n + v] is the coupling of
 * node v to node v + off(o), off(o) = (o%3-1, o/3%3-1, o/9-1). kappa_eff is
 * the per-element conductivity already including the 1e-9 floor. */
void orc_assemble(int nx, int ny, int nz, const double* kappa_eff,
                  const double Ke[64], double* A);

/* y = A x for the ELL-27 operator. */
void orc_matvec(int nx, int ny, int nz, const double* A, const double* x, double* y);

/* Load-case RHS f_v = sum_e kappa_eff_e * feMu[l(v,e)][k]. */
void orc_rhs(int nx, int ny, int nz, const double* kappa_eff,
             const double feMu[24], int k, double* f);

/* Jacobi-preconditioned CG on A x = b. pin != 0 keeps node 0 at zero (the
 * reference's Dirichlet pin); pin == 0 projects the constant null space out
 * of every residual. Returns iterations; *relres gets ||b-Ax||/||b||. */
int orc_cg(int nx, int ny, int nz, const double* A, const double* b, double* x,
           double rtol, int maxit, int pin, double* relres);

/* Full homogenization: kappa (SIMP-mapped, without floor) -> chi[3n],
 * C[9] (row-major 3x3), S[6n] element energies in the order
 * (00,01,02,11,12,22). chi/S may be NULL. Returns max CG iterations. */
int orc_homogenize(int nx, int ny, int nz, const double* kappa, double rtol,
                   int pin, double* chi, double C[9], double* S);

/* Element energy tensors S_e(i,j) from given chi (3n) without solving. */
void orc_element_energy(int nx, int ny, int nz, const double* chi, double* S6);

/* Contracted sensitivity sens_e = sum_{i<=j} S_e(i,j) (dC_ij + [i!=j] dC_ji). */
void orc_sensitivity(int n, const double* S6, const double dC[9], double* sens);

/* Radial convolution filters (TensorExpression.h:1053-1093, .cuh:303-384).
 * kind 0 = Spline4, kind 1 = Linear; radius r; periodic flag. Forward:
 * out_i = sum_j w_ij in_j / W_i over in-domain taps (all taps if periodic);
 * W (may be NULL) receives W_i. Adjoint: out_j = sum_i w_ij in_i / W_i. */
void orc_filter_fwd(int nx, int ny, int nz, int kind, double r, int periodic,
                    const double* in, double* out, double* W);
void orc_filter_adj(int nx, int ny, int nz, int kind, double r, int periodic,
                    const double* in, double* out);

/* OC sensitivity filter (OCOptimizer.cu:76-104, :134-152): periodic linear
 * radius r, density weighted: s_i <- sum_j w s_j rho_j / (rho_i sum_j w). */
void orc_oc_filter_sens(int nx, int ny, int nz, double r, const double* rho,
                        const double* sens, double* out);

/* OC update with Lagrange bisection (OCOptimizer.cu:12-67). Writes rho_new.
 * Returns the number of bisection steps taken; *g_out the final multiplier. */
int orc_oc_update(int n, const double* sens, const double* rho, double* rho_new,
                  double volratio, double min_rho, double step_limit, double damp,
                  int max_steps, double vol_tol, double* g_out);

/* One MMA update (mmaOpt.cu:834-1019 mmasub_ker with subsolv_g :429-830,
 * Svanberg's method-of-moving-asymptotes subproblem solved by a primal-dual
 * interior point). m constraints, n variables; dgdx is m x n row-major.
 * x is updated in place; xold1/xold2/low/upp carry state between calls.
 * epsimin is the last interior-point barrier level (1e-7 in the reference,
 * mmaOpt.cu:846). Returns the number of Newton iterations. */
int orc_mmasub(int m, int n, int itn, double* x, const double* xmin, const double* xmax,
               double* xold1, double* xold2, const double* df0dx, const double* gval,
               const double* dgdx, double* low, double* upp, double a0, const double* a,
               const double* c, const double* d, double move, double epsimin);

/* ---- multigrid pieces (mg_impl.h; orc_ = FP64, orc32_ = FP32 build) ----
 * Operators are ELL-27 as from orc_assemble; vectors hold nrhs right-hand
 * sides, load case major. */
/* Exact 8-colour Gauss-Seidel sweeps in the reference's colour order
 * (grid.cu:943-987, colours of grid.cpp:520-545; reverse = 1: 7..0). */
void orc_gs_ell(int nx, int ny, int nz, const double* A, const double* f, double* u, int sweeps, int reverse,
                int nrhs);
/* Brick (block) Gauss-Seidel: bricks of bx*by*bz nodes from the origin,
 * cut at the edges, exact inside, couplings leaving a brick frozen at uin. */
void orc_brick_gs_ell(int nx, int ny, int nz, const double* A, const double* f, const double* uin, double* uout,
                      int bx, int by, int bz, int sweeps, int reverse, int nrhs);
/* P^T A P onto the lattice of half the edges (trilinear P). */
void orc_galerkin_ell(int nx, int ny, int nz, const double* Af, double* Ac);
/* kappa^H by MG-preconditioned CG (same system as orc_homogenize). */
int orc_mg_homogenize(int nx, int ny, int nz, const double* kappa, double rtol, double* chi, double C[9], double* S,
                      int its[3], double relres[3]);
int orc32_mg_homogenize(int nx, int ny, int nz, const double* kappa, double rtol, double* chi, double C[9],
                        double* S, int its[3], double relres[3]);
void orc32_gs_ell(int nx, int ny, int nz, const float* A, const float* f, float* u, int sweeps, int reverse,
                  int nrhs);
void orc32_brick_gs_ell(int nx, int ny, int nz, const float* A, const float* f, const float* uin, float* uout, int bx,
                        int by, int bz, int sweeps, int reverse, int nrhs);
void orc32_galerkin_ell(int nx, int ny, int nz, const float* Af, float* Ac);

#ifdef __cplusplus
}
#endif
#endif
