/*
 * mg_impl.h -- geometric multigrid pieces of the CPU restatement, written
 * once and compiled twice: with REAL = double (mg64.c, the oracle's fast
 * FP64 solver for large cells) and with REAL = float (mg32.c, the CPU port
 * of the GPU's plain FP32 MG-PCG, timed as the CPU baseline). FN(name)
 * names the exported symbol of each build.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Every level is a periodic lattice with an ELL-27 operator A[o*n + v]
 * (coupling of node v to v + off(o), off(o) = (o%3-1, o/3%3-1, o/9-1)), the
 * layout orc_assemble writes. Vectors hold three right-hand sides, load
 * case major: v[k*n + i].
 *
 * Reference behaviour restated here:
 *  - Gauss-Seidel colours (Grid_H::gs_relaxation, grid.cu:943-987, and the
 *    colour sets of Grid_H::vgsid2lexpos_h, grid.cpp:520-545): node (x,y,z)
 *    of the unpadded lattice has padded position (x+1, y+1, z+1) and colour
 *    c = ((x+1)&1) + 2((y+1)&1) + 4((z+1)&1); a sweep relaxes colours 0..7
 *    in that order, u_v = (f_v - sum_{j != v} A_vj u_j) / A_vv (w_SOR = 1).
 *  - Restriction / prolongation (grid.cu:811-940): trilinear, weight
 *    prod_axes (1 - |d|/2) for the fine node at distance d in {-1,0,1} from
 *    2I, restriction the transpose (no 1/8 scaling).
 *  - Galerkin coarse operators P^T A P (grid_ex.cu:16-227 builds them from
 *    the element matrices; here by the stencil triple product, an
 *    independent formulation of the same operator).
 *  - The coarsest level is solved directly (grid.cpp:272-310 uses a dense
 *    decomposition): here a Cholesky factorisation, in FP64, of
 *    A + (tr A / n^2) 1 1^T, which equals A on mean-free vectors.
 * The V-cycle smooths twice before and after each coarse correction
 * (colours 0..7 before, 7..0 after, so the cycle is symmetric), and the
 * solver is MG-preconditioned CG on the mean-free subspace with FP64
 * inner products.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static long FN(vidx)(int x, int y, int z, int nx, int ny) { return (long)x + (long)nx * ((long)y + (long)ny * z); }
static int FN(wrap)(int a, int n) { return ((a % n) + n) % n; }

typedef struct {
  int nx, ny, nz;
  long n;
  REAL* A;        /* 27 n */
  REAL *u, *f, *r; /* 3 n each */
  double* L;      /* coarsest: Cholesky factor, n x n */
} FN(Lvl);

/* neighbour index of v through offset o */
static long FN(nb)(const FN(Lvl) * L, int x, int y, int z, int o) {
  return FN(vidx)(FN(wrap)(x + o % 3 - 1, L->nx), FN(wrap)(y + (o / 3) % 3 - 1, L->ny), FN(wrap)(z + o / 9 - 1, L->nz),
                  L->nx, L->ny);
}

/* One colour of the reference's 8-colour Gauss-Seidel, in place on u. */
static void FN(gs_colour)(const FN(Lvl) * L, const REAL* f, REAL* u, int c, int nrhs) {
  const int x0 = 1 - (c & 1), y0 = 1 - ((c >> 1) & 1), z0 = 1 - (c >> 2);
  const long n = L->n;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = z0; z < L->nz; z += 2)
    for (int y = y0; y < L->ny; y += 2)
      for (int x = x0; x < L->nx; x += 2) {
        const long v = FN(vidx)(x, y, z, L->nx, L->ny);
        for (int k = 0; k < nrhs; k++) {
          double s = (double)f[k * n + v];
          for (int o = 0; o < 27; o++)
            if (o != 13) s -= (double)L->A[(long)o * n + v] * (double)u[k * n + FN(nb)(L, x, y, z, o)];
          u[k * n + v] = (REAL)(s / (double)L->A[13 * n + v]);
        }
      }
}

static void FN(gs_sweep)(const FN(Lvl) * L, const REAL* f, REAL* u, int reverse, int nrhs) {
  for (int i = 0; i < 8; i++) FN(gs_colour)(L, f, u, reverse ? 7 - i : i, nrhs);
}

/* r = f - A u */
static void FN(residual)(const FN(Lvl) * L, const REAL* f, const REAL* u, REAL* r, int nrhs) {
  const long n = L->n;
#pragma omp parallel for schedule(static)
  for (long v = 0; v < n; v++) {
    const int x = (int)(v % L->nx), y = (int)((v / L->nx) % L->ny), z = (int)(v / ((long)L->nx * L->ny));
    long nbv[27];
    for (int o = 0; o < 27; o++) nbv[o] = FN(nb)(L, x, y, z, o);
    for (int k = 0; k < nrhs; k++) {
      double s = (double)f[k * n + v];
      for (int o = 0; o < 27; o++) s -= (double)L->A[(long)o * n + v] * (double)u[k * n + nbv[o]];
      r[k * n + v] = (REAL)s;
    }
  }
}

/* trilinear weight of fine offset d in {-1, 0, 1} */
static double FN(w1)(int d) { return d == 0 ? 1.0 : 0.5; }

/* fc = P^T r: coarse node I gathers the 27 fine nodes around 2I */
static void FN(restrict)(const FN(Lvl) * F, const FN(Lvl) * C, const REAL* r, REAL* fc, int nrhs) {
#pragma omp parallel for schedule(static)
  for (long I = 0; I < C->n; I++) {
    const int X = (int)(I % C->nx), Y = (int)((I / C->nx) % C->ny), Z = (int)(I / ((long)C->nx * C->ny));
    double s[3] = {0, 0, 0};
    for (int o = 0; o < 27; o++) {
      const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
      const double w = FN(w1)(dx) * FN(w1)(dy) * FN(w1)(dz);
      const long i = FN(vidx)(FN(wrap)(2 * X + dx, F->nx), FN(wrap)(2 * Y + dy, F->ny), FN(wrap)(2 * Z + dz, F->nz),
                              F->nx, F->ny);
      for (int k = 0; k < nrhs; k++) s[k] += w * (double)r[k * F->n + i];
    }
    for (int k = 0; k < nrhs; k++) fc[k * C->n + I] = (REAL)s[k];
  }
}

/* uf += P uc: fine node i takes the coarse nodes of its 2x2x2 parent cell */
static void FN(prolong_add)(const FN(Lvl) * F, const FN(Lvl) * C, const REAL* uc, REAL* uf, int nrhs) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < F->n; i++) {
    const int x = (int)(i % F->nx), y = (int)((i / F->nx) % F->ny), z = (int)(i / ((long)F->nx * F->ny));
    double s[3] = {0, 0, 0};
    for (int c = 0; c < 8; c++) {
      const int ax = c & 1, ay = (c >> 1) & 1, az = c >> 2;
      /* coarse node X = floor(x/2) + a, used when x is odd or a == 0 */
      if ((ax && !(x & 1)) || (ay && !(y & 1)) || (az && !(z & 1))) continue;
      const double w = ((x & 1) ? 0.5 : 1.0) * ((y & 1) ? 0.5 : 1.0) * ((z & 1) ? 0.5 : 1.0);
      const long I = FN(vidx)(FN(wrap)(x / 2 + ax, C->nx), FN(wrap)(y / 2 + ay, C->ny), FN(wrap)(z / 2 + az, C->nz),
                              C->nx, C->ny);
      for (int k = 0; k < nrhs; k++) s[k] += w * (double)uc[k * C->n + I];
    }
    for (int k = 0; k < nrhs; k++) uf[k * F->n + i] = (REAL)((double)uf[k * F->n + i] + s[k]);
  }
}

/* A_c = P^T A_f P as a stencil: for coarse node I, sum over the fine nodes i
 * of P(:, I), their fine couplings (i, j), and the coarse nodes J of P(j, :). */
static void FN(galerkin)(const FN(Lvl) * F, FN(Lvl) * C) {
  memset(C->A, 0, sizeof(REAL) * 27 * C->n);
#pragma omp parallel for schedule(dynamic, 64)
  for (long I = 0; I < C->n; I++) {
    const int X = (int)(I % C->nx), Y = (int)((I / C->nx) % C->ny), Z = (int)(I / ((long)C->nx * C->ny));
    double acc[27];
    for (int q = 0; q < 27; q++) acc[q] = 0.0;
    for (int oi = 0; oi < 27; oi++) {
      const int dx = oi % 3 - 1, dy = (oi / 3) % 3 - 1, dz = oi / 9 - 1;
      const double wi = FN(w1)(dx) * FN(w1)(dy) * FN(w1)(dz);
      const int fx = 2 * X + dx, fy = 2 * Y + dy, fz = 2 * Z + dz; /* unwrapped fine coordinates */
      const long i = FN(vidx)(FN(wrap)(fx, F->nx), FN(wrap)(fy, F->ny), FN(wrap)(fz, F->nz), F->nx, F->ny);
      for (int o = 0; o < 27; o++) {
        const double a = (double)F->A[(long)o * F->n + i];
        if (a == 0.0) continue;
        const int jx = fx + o % 3 - 1, jy = fy + (o / 3) % 3 - 1, jz = fz + o / 9 - 1;
        /* coarse parents of fine node j (unwrapped): offsets relative to I */
        for (int c = 0; c < 8; c++) {
          const int ax = c & 1, ay = (c >> 1) & 1, az = c >> 2;
          if ((ax && !(jx & 1)) || (ay && !(jy & 1)) || (az && !(jz & 1))) continue;
          const int Jx = (jx - (jx & 1)) / 2 + ax, Jy = (jy - (jy & 1)) / 2 + ay, Jz = (jz - (jz & 1)) / 2 + az;
          const double wj = ((jx & 1) ? 0.5 : 1.0) * ((jy & 1) ? 0.5 : 1.0) * ((jz & 1) ? 0.5 : 1.0);
          const int ox = Jx - X, oy = Jy - Y, oz = Jz - Z;
          if (ox < -1 || ox > 1 || oy < -1 || oy > 1 || oz < -1 || oz > 1) continue; /* cannot happen */
          acc[(ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)] += wi * a * wj;
        }
      }
    }
    for (int q = 0; q < 27; q++) C->A[(long)q * C->n + I] = (REAL)acc[q];
  }
}

/* Dense coarsest operator + rank-one regularisation, Cholesky in FP64. */
static int FN(coarsest_factor)(FN(Lvl) * C) {
  const long n = C->n;
  double* L = (double*)calloc((size_t)(n * n), sizeof(double));
  if (!L) return -1;
  double tr = 0.0;
  for (long v = 0; v < n; v++) {
    const int x = (int)(v % C->nx), y = (int)((v / C->nx) % C->ny), z = (int)(v / ((long)C->nx * C->ny));
    for (int o = 0; o < 27; o++) L[v * n + FN(nb)(C, x, y, z, o)] += (double)C->A[(long)o * n + v];
    tr += (double)C->A[13 * n + v];
  }
  /* symmetrise (the FP32 build's stencil is symmetric to rounding) */
  for (long i = 0; i < n; i++)
    for (long j = i + 1; j < n; j++) {
      const double s = 0.5 * (L[i * n + j] + L[j * n + i]);
      L[i * n + j] = L[j * n + i] = s;
    }
  const double alpha = tr / ((double)n * (double)n);
  for (long i = 0; i < n * n; i++) L[i] += alpha;
  for (long j = 0; j < n; j++) {
    double d = L[j * n + j];
    for (long k = 0; k < j; k++) d -= L[j * n + k] * L[j * n + k];
    if (!(d > 0.0)) {
      free(L);
      return -2;
    }
    d = sqrt(d);
    L[j * n + j] = d;
    for (long i = j + 1; i < n; i++) {
      double s = L[i * n + j];
      for (long k = 0; k < j; k++) s -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = s / d;
    }
  }
  C->L = L;
  return 0;
}

static void FN(coarsest_solve)(const FN(Lvl) * C, const REAL* f, REAL* u, int nrhs) {
  const long n = C->n;
  double* y = (double*)malloc(sizeof(double) * n);
  for (int k = 0; k < nrhs; k++) {
    for (long i = 0; i < n; i++) {
      double s = (double)f[k * n + i];
      for (long j = 0; j < i; j++) s -= C->L[i * n + j] * y[j];
      y[i] = s / C->L[i * n + i];
    }
    for (long i = n - 1; i >= 0; i--) {
      double s = y[i];
      for (long j = i + 1; j < n; j++) s -= C->L[j * n + i] * y[j];
      y[i] = s / C->L[i * n + i];
    }
    for (long i = 0; i < n; i++) u[k * n + i] = (REAL)y[i];
  }
  free(y);
}

typedef struct {
  int nl;
  FN(Lvl) lv[12];
} FN(Hier);

static void FN(hier_free)(FN(Hier) * H) {
  for (int l = 0; l < H->nl; l++) {
    if (l > 0) free(H->lv[l].A);
    free(H->lv[l].u);
    free(H->lv[l].f);
    free(H->lv[l].r);
    free(H->lv[l].L);
  }
  H->nl = 0;
}

/* Levels halve every edge while all edges are even and the halves are at
 * least `minc`; A0 (the fine ELL operator) is borrowed, not owned. */
static int FN(hier_build)(FN(Hier) * H, int nx, int ny, int nz, REAL* A0, int minc) {
  memset(H, 0, sizeof(*H));
  int d[3] = {nx, ny, nz};
  for (;;) {
    FN(Lvl)* L = &H->lv[H->nl];
    L->nx = d[0];
    L->ny = d[1];
    L->nz = d[2];
    L->n = (long)d[0] * d[1] * d[2];
    L->A = H->nl == 0 ? A0 : (REAL*)malloc(sizeof(REAL) * 27 * L->n);
    L->u = (REAL*)calloc((size_t)(3 * L->n), sizeof(REAL));
    L->f = (REAL*)calloc((size_t)(3 * L->n), sizeof(REAL));
    L->r = (REAL*)calloc((size_t)(3 * L->n), sizeof(REAL));
    if (H->nl > 0) FN(galerkin)(&H->lv[H->nl - 1], L);
    H->nl++;
    int can = H->nl < 12;
    for (int a = 0; a < 3; a++)
      if ((d[a] & 1) || d[a] / 2 < minc) can = 0;
    if (!can) break;
    for (int a = 0; a < 3; a++) d[a] /= 2;
  }
  if (H->lv[H->nl - 1].n > 4096) return -3;
  return FN(coarsest_factor)(&H->lv[H->nl - 1]);
}

/* u = M^{-1} f on level l (u overwritten), V(2,2). */
static void FN(vcycle)(FN(Hier) * H, int l, const REAL* f, REAL* u, int nrhs) {
  FN(Lvl)* L = &H->lv[l];
  if (l == H->nl - 1) {
    FN(coarsest_solve)(L, f, u, nrhs);
    return;
  }
  memset(u, 0, sizeof(REAL) * 3 * L->n);
  for (int s = 0; s < 2; s++) FN(gs_sweep)(L, f, u, 0, nrhs);
  FN(residual)(L, f, u, L->r, nrhs);
  FN(Lvl)* C = &H->lv[l + 1];
  FN(restrict)(L, C, L->r, C->f, nrhs);
  FN(vcycle)(H, l + 1, C->f, C->u, nrhs);
  FN(prolong_add)(L, C, C->u, u, nrhs);
  for (int s = 0; s < 2; s++) FN(gs_sweep)(L, f, u, 1, nrhs);
}

static double FN(dotk)(long n, const REAL* a, const REAL* b) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long i = 0; i < n; i++) s += (double)a[i] * (double)b[i];
  return s;
}
static void FN(remove_mean)(long n, REAL* a) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long i = 0; i < n; i++) s += (double)a[i];
  s /= (double)n;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) a[i] = (REAL)((double)a[i] - s);
}

/* MG-preconditioned CG of the three load cases (each its own recurrence)
 * on the mean-free subspace, from x (warm start) to relative residual rtol
 * or maxit iterations. The true residual is recomputed and checked when
 * the recurrence says it has converged (restart if it has not).
 * its[k] / relres[k] receive iterations and final true relative residuals. */
static void FN(mgpcg)(FN(Hier) * H, const REAL* b, REAL* x, double rtol, int maxit, int its[3], double relres[3]) {
  FN(Lvl)* L = &H->lv[0];
  const long n = L->n;
  REAL* r = (REAL*)malloc(sizeof(REAL) * 3 * n);
  REAL* z = (REAL*)malloc(sizeof(REAL) * 3 * n);
  REAL* p = (REAL*)malloc(sizeof(REAL) * 3 * n);
  REAL* q = (REAL*)malloc(sizeof(REAL) * 3 * n);
  double rz[3] = {0, 0, 0}, bn[3];
  int active[3] = {1, 1, 1}, it = 0;
  for (int k = 0; k < 3; k++) {
    its[k] = 0;
    bn[k] = sqrt(FN(dotk)(n, b + k * n, b + k * n));
    if (!(bn[k] > 0)) bn[k] = 1.0;
  }
  for (int restart = 0;; restart++) {
    FN(residual)(L, b, x, r, 3);
    int any = 0;
    for (int k = 0; k < 3; k++) {
      FN(remove_mean)(n, r + k * n);
      relres[k] = sqrt(FN(dotk)(n, r + k * n, r + k * n)) / bn[k];
      active[k] = relres[k] > rtol;
      any |= active[k];
    }
    if (!any || it >= maxit || restart >= 6) break;
    FN(vcycle)(H, 0, r, z, 3);
    for (int k = 0; k < 3; k++) {
      FN(remove_mean)(n, z + k * n);
      rz[k] = FN(dotk)(n, r + k * n, z + k * n);
    }
    memcpy(p, z, sizeof(REAL) * 3 * n);
    while (it < maxit) {
      /* q = -(A p): the residual of p with a zero right-hand side */
      memset(q, 0, sizeof(REAL) * 3 * n);
      FN(residual)(L, q, p, q, 3);
      int still = 0;
      for (int k = 0; k < 3; k++) {
        if (!active[k]) continue;
        const double pq = -FN(dotk)(n, p + k * n, q + k * n);
        const double alpha = rz[k] / pq;
        REAL* xk = x + k * n;
        REAL* rk = r + k * n;
        const REAL* pk = p + k * n;
        const REAL* qk = q + k * n;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
          xk[i] = (REAL)((double)xk[i] + alpha * (double)pk[i]);
          rk[i] = (REAL)((double)rk[i] + alpha * (double)qk[i]);
        }
        its[k]++;
        const double rr = sqrt(FN(dotk)(n, rk, rk)) / bn[k];
        if (getenv("ORC_VERBOSE")) {
          fprintf(stderr, "[oracle mg] it %d load case %d relres %.3e\n", it + 1, k, rr);
          fflush(stderr);
        }
        if (rr <= rtol) active[k] = 0;
        still |= active[k];
      }
      it++;
      if (!still) break;
      FN(vcycle)(H, 0, r, z, 3);
      for (int k = 0; k < 3; k++) {
        if (!active[k]) continue;
        FN(remove_mean)(n, z + k * n);
        const double rzn = FN(dotk)(n, r + k * n, z + k * n);
        const double beta = rzn / rz[k];
        rz[k] = rzn;
        REAL* pk = p + k * n;
        const REAL* zk = z + k * n;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) pk[i] = (REAL)((double)zk[i] + beta * (double)pk[i]);
      }
    }
  }
  free(r);
  free(z);
  free(p);
  free(q);
}

/* ---- exported entry points ------------------------------------------------ */

/* Exact multicolour Gauss-Seidel sweeps (reverse = 0: the reference's
 * colour order 0..7, 1: 7..0) in place on u, nrhs right-hand sides. */
void FN(gs_ell)(int nx, int ny, int nz, const REAL* A, const REAL* f, REAL* u, int sweeps, int reverse, int nrhs) {
  FN(Lvl) L;
  memset(&L, 0, sizeof(L));
  L.nx = nx;
  L.ny = ny;
  L.nz = nz;
  L.n = (long)nx * ny * nz;
  L.A = (REAL*)A;
  for (int s = 0; s < sweeps; s++) FN(gs_sweep)(&L, f, u, reverse, nrhs);
}

/* Block Gauss-Seidel over bricks of bx*by*bz nodes from the origin (cut at
 * the lattice edge): inside a brick the sweeps are exact in the colour
 * order; a coupling to a node outside the brick (in unwrapped coordinates)
 * reads its value from uin, frozen. The result goes to uout. */
void FN(brick_gs_ell)(int nx, int ny, int nz, const REAL* A, const REAL* f, const REAL* uin, REAL* uout, int bx,
                      int by, int bz, int sweeps, int reverse, int nrhs) {
  const long n = (long)nx * ny * nz;
  memcpy(uout, uin, sizeof(REAL) * (size_t)nrhs * (size_t)n);
  const int nbx = (nx + bx - 1) / bx, nby = (ny + by - 1) / by, nbz = (nz + bz - 1) / bz;
#pragma omp parallel for schedule(dynamic, 1)
  for (long b = 0; b < (long)nbx * nby * nbz; b++) {
    const int ox = (int)(b % nbx) * bx, oy = (int)((b / nbx) % nby) * by, oz = (int)(b / ((long)nbx * nby)) * bz;
    const int ex = nx - ox < bx ? nx - ox : bx, ey = ny - oy < by ? ny - oy : by, ez = nz - oz < bz ? nz - oz : bz;
    for (int s = 0; s < sweeps; s++)
      for (int i = 0; i < 8; i++) {
        const int c = reverse ? 7 - i : i;
        const int px = 1 - (c & 1), py = 1 - ((c >> 1) & 1), pz = 1 - (c >> 2); /* coordinate parities */
        for (int z = oz; z < oz + ez; z++) {
          if ((z & 1) != pz) continue;
          for (int y = oy; y < oy + ey; y++) {
            if ((y & 1) != py) continue;
            for (int x = ox; x < ox + ex; x++) {
              if ((x & 1) != px) continue;
              const long v = FN(vidx)(x, y, z, nx, ny);
              for (int k = 0; k < nrhs; k++) {
                double acc = (double)f[k * n + v];
                for (int o = 0; o < 27; o++) {
                  if (o == 13) continue;
                  const int X = x + o % 3 - 1, Y = y + (o / 3) % 3 - 1, Z = z + o / 9 - 1;
                  const int in = X >= ox && X < ox + ex && Y >= oy && Y < oy + ey && Z >= oz && Z < oz + ez;
                  const long j = FN(vidx)(FN(wrap)(X, nx), FN(wrap)(Y, ny), FN(wrap)(Z, nz), nx, ny);
                  acc -= (double)A[(long)o * n + v] * (double)(in ? uout[k * n + j] : uin[k * n + j]);
                }
                uout[k * n + v] = (REAL)(acc / (double)A[13 * n + v]);
              }
            }
          }
        }
      }
  }
}

/* Galerkin operator P^T A P of a level (edges nx, ny, nz, all even) onto
 * the level of half the edges. */
void FN(galerkin_ell)(int nx, int ny, int nz, const REAL* Af, REAL* Ac) {
  FN(Lvl) F, C;
  memset(&F, 0, sizeof(F));
  memset(&C, 0, sizeof(C));
  F.nx = nx;
  F.ny = ny;
  F.nz = nz;
  F.n = (long)nx * ny * nz;
  F.A = (REAL*)Af;
  C.nx = nx / 2;
  C.ny = ny / 2;
  C.nz = nz / 2;
  C.n = (long)C.nx * C.ny * C.nz;
  C.A = Ac;
  FN(galerkin)(&F, &C);
}

/* kappa^H by MG-preconditioned CG (the same system as orc_homogenize, a
 * faster solver for large cells): kappa is the SIMP-mapped conductivity
 * (the 1e-9 floor is added inside), chi (3n, may be NULL) the mean-free
 * fluctuation fields, S (6n, may be NULL) the element energies. its and
 * relres receive per-load-case iterations and final true relative
 * residuals. Returns 0, or a negative code (coarsest level too large or
 * not positive). */
int FN(mg_homogenize)(int nx, int ny, int nz, const double* kappa, double rtol, double* chi_out, double C[9],
                      double* S_out, int its[3], double relres[3]) {
  const long n = (long)nx * ny * nz;
  double Ke[64], feMu[24], disp[24];
  orc_template(0.5 / nx, Ke, feMu, disp);
  double* ke = (double*)malloc(sizeof(double) * n);
  for (long e = 0; e < n; e++) ke[e] = 1e-9 + kappa[e];
  double* Ad = (double*)malloc(sizeof(double) * 27 * n);
  orc_assemble(nx, ny, nz, ke, Ke, Ad);
  REAL* A0;
  if (sizeof(REAL) == sizeof(double)) {
    A0 = (REAL*)Ad;
  } else {
    A0 = (REAL*)malloc(sizeof(REAL) * 27 * n);
    for (long i = 0; i < 27 * n; i++) A0[i] = (REAL)Ad[i];
    free(Ad);
    Ad = NULL;
  }
  FN(Hier) H;
  const int st = FN(hier_build)(&H, nx, ny, nz, A0, 4);
  if (st != 0) {
    FN(hier_free)(&H);
    free(A0);
    free(ke);
    return st;
  }
  REAL* b = (REAL*)malloc(sizeof(REAL) * 3 * n);
  REAL* x = (REAL*)calloc((size_t)(3 * n), sizeof(REAL));
  double* fk = (double*)malloc(sizeof(double) * n);
  for (int k = 0; k < 3; k++) {
    orc_rhs(nx, ny, nz, ke, feMu, k, fk);
    for (long i = 0; i < n; i++) b[k * n + i] = (REAL)fk[i];
  }
  free(fk);
  FN(mgpcg)(&H, b, x, rtol, 500, its, relres);
  double* chi = chi_out ? chi_out : (double*)malloc(sizeof(double) * 3 * n);
  for (int k = 0; k < 3; k++) {
    FN(remove_mean)(n, x + k * n);
    for (long i = 0; i < n; i++) chi[k * n + i] = (double)x[k * n + i];
  }
  double* S = S_out ? S_out : (double*)malloc(sizeof(double) * 6 * n);
  orc_element_energy(nx, ny, nz, chi, S);
  double c6[6];
  for (int c = 0; c < 6; c++) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long e = 0; e < n; e++) s += ke[e] * S[(long)c * n + e];
    c6[c] = s;
  }
  C[0] = c6[0]; C[1] = c6[1]; C[2] = c6[2];
  C[3] = c6[1]; C[4] = c6[3]; C[5] = c6[4];
  C[6] = c6[2]; C[7] = c6[4]; C[8] = c6[5];
  if (!S_out) free(S);
  if (!chi_out) free(chi);
  free(b);
  free(x);
  FN(hier_free)(&H);
  free(A0);
  free(ke);
  return 0;
}
