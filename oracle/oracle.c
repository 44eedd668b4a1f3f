/*
 * oracle.c -- FP64 CPU restatement of the OpenTM thermal homogenization path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path. Loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by
 * the product library.
 *
 * Conventions (see oracle.h): lexicographic x-fastest arrays; element
 * (x,y,z) spans nodes (x..x+1, y..y+1, z..z+1) mod size; local node
 * l = bx + 2 by + 4 bz.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RHO_PENAL_MIN 1e-9 /* grid.h:429 rhoPenalMin */

static inline int wrap(int a, int n) { a %= n; return a < 0 ? a + n : a; }
static inline long idx3(int x, int y, int z, int nx, int ny) {
  return (long)x + (long)nx * ((long)y + (long)ny * z);
}

/* ---- element template ---------------------------------------------------
 * Restates templateMatrix.cpp:21-87: 3x3x3 Gauss-Legendre quadrature of the
 * trilinear hex with corners at +-a, Kmu = sum w qxyz^T qxyz, feMu = sum w
 * qxyz^T (heat matrix = identity), then the node remap {1,3,2,0,5,7,6,4}
 * (templateMatrix.cpp:74-82) and disp = Ke[1:,1:]^-1 feMu[1:] with node 0
 * fixed at zero (templateMatrix.cpp:85). */
static void inv3(const double J[9], double Ji[9], double* det) {
  double d = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
             J[2] * (J[3] * J[7] - J[4] * J[6]);
  *det = d;
  Ji[0] = (J[4] * J[8] - J[5] * J[7]) / d;
  Ji[1] = (J[2] * J[7] - J[1] * J[8]) / d;
  Ji[2] = (J[1] * J[5] - J[2] * J[4]) / d;
  Ji[3] = (J[5] * J[6] - J[3] * J[8]) / d;
  Ji[4] = (J[0] * J[8] - J[2] * J[6]) / d;
  Ji[5] = (J[2] * J[3] - J[0] * J[5]) / d;
  Ji[6] = (J[3] * J[7] - J[4] * J[6]) / d;
  Ji[7] = (J[1] * J[6] - J[0] * J[7]) / d;
  Ji[8] = (J[0] * J[4] - J[1] * J[3]) / d;
}

/* Dense solve with partial pivoting, A (n x n, row-major) overwritten,
 * B (n x m) overwritten by the solution. */
static void dense_solve(int n, int m, double* A, double* B) {
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int i = k + 1; i < n; i++)
      if (fabs(A[i * n + k]) > fabs(A[p * n + k])) p = i;
    if (p != k) {
      for (int j = 0; j < n; j++) { double t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t; }
      for (int j = 0; j < m; j++) { double t = B[k * m + j]; B[k * m + j] = B[p * m + j]; B[p * m + j] = t; }
    }
    for (int i = k + 1; i < n; i++) {
      double f = A[i * n + k] / A[k * n + k];
      for (int j = k; j < n; j++) A[i * n + j] -= f * A[k * n + j];
      for (int j = 0; j < m; j++) B[i * m + j] -= f * B[k * m + j];
    }
  }
  for (int k = n - 1; k >= 0; k--) {
    for (int j = 0; j < m; j++) {
      double s = B[k * m + j];
      for (int i = k + 1; i < n; i++) s -= A[k * n + i] * B[i * m + j];
      B[k * m + j] = s / A[k * n + k];
    }
  }
}

void orc_template(double a, double Ke[64], double feMu[24], double disp[24]) {
  const double g = sqrt(3.0 / 5.0);
  const double xs[3] = {-g, 0.0, g};
  const double ws[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
  /* corner coordinates of the reference element order (templateMatrix.cpp:50-58) */
  const double Jr[8][3] = {{-a, -a, -a}, {a, -a, -a}, {a, a, -a}, {-a, a, -a},
                           {-a, -a, a},  {a, -a, a},  {a, a, a},  {-a, a, a}};
  double K[64] = {0}, F[24] = {0};
  for (int ii = 0; ii < 3; ii++)
    for (int jj = 0; jj < 3; jj++)
      for (int kk = 0; kk < 3; kk++) {
        double x = xs[ii], y = xs[jj], z = xs[kk];
        double q[3][8] = {
            {-((y - 1) * (z - 1)) / 8, ((y - 1) * (z - 1)) / 8, -((y + 1) * (z - 1)) / 8,
             ((y + 1) * (z - 1)) / 8, ((y - 1) * (z + 1)) / 8, -((y - 1) * (z + 1)) / 8,
             ((y + 1) * (z + 1)) / 8, -((y + 1) * (z + 1)) / 8},
            {-((x - 1) * (z - 1)) / 8, ((x + 1) * (z - 1)) / 8, -((x + 1) * (z - 1)) / 8,
             ((x - 1) * (z - 1)) / 8, ((x - 1) * (z + 1)) / 8, -((x + 1) * (z + 1)) / 8,
             ((x + 1) * (z + 1)) / 8, -((x - 1) * (z + 1)) / 8},
            {-((x - 1) * (y - 1)) / 8, ((x + 1) * (y - 1)) / 8, -((x + 1) * (y + 1)) / 8,
             ((x - 1) * (y + 1)) / 8, ((x - 1) * (y - 1)) / 8, -((x + 1) * (y - 1)) / 8,
             ((x + 1) * (y + 1)) / 8, -((x - 1) * (y + 1)) / 8}};
        double J[9] = {0}, Ji[9], det;
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) {
            double s = 0;
            for (int n = 0; n < 8; n++) s += q[r][n] * Jr[n][c];
            J[r * 3 + c] = s;
          }
        inv3(J, Ji, &det);
        double qx[3][8];
        for (int r = 0; r < 3; r++)
          for (int n = 0; n < 8; n++)
            qx[r][n] = Ji[r * 3 + 0] * q[0][n] + Ji[r * 3 + 1] * q[1][n] + Ji[r * 3 + 2] * q[2][n];
        double w = det * ws[ii] * ws[jj] * ws[kk];
        for (int m = 0; m < 8; m++) {
          for (int n = 0; n < 8; n++)
            K[m * 8 + n] += w * (qx[0][m] * qx[0][n] + qx[1][m] * qx[1][n] + qx[2][m] * qx[2][n]);
          for (int c = 0; c < 3; c++) F[m * 3 + c] += w * qx[c][m];
        }
      }
  const int map[8] = {1, 3, 2, 0, 5, 7, 6, 4};
  for (int i = 0; i < 8; i++) {
    for (int j = 0; j < 8; j++) Ke[map[i] * 8 + map[j]] = K[i * 8 + j];
    for (int c = 0; c < 3; c++) feMu[map[i] * 3 + c] = F[i * 3 + c];
  }
  double A7[49], B7[21];
  for (int i = 0; i < 7; i++) {
    for (int j = 0; j < 7; j++) A7[i * 7 + j] = Ke[(i + 1) * 8 + (j + 1)];
    for (int c = 0; c < 3; c++) B7[i * 3 + c] = feMu[(i + 1) * 3 + c];
  }
  dense_solve(7, 3, A7, B7);
  for (int c = 0; c < 3; c++) disp[c] = 0.0;
  for (int i = 0; i < 7; i++)
    for (int c = 0; c < 3; c++) disp[(i + 1) * 3 + c] = B7[i * 3 + c];
}

/* ---- assembly ------------------------------------------------------------
 * Element-by-element assembly of sum_e kappa_e Ke over the periodic cell.
 * The reference applies the same operator matrix-free, node by node
 * (grid.cu:318-471 smoother, grid.cu:1027-1164 residual), with weights
 * rhoPenalMin + kappa (grid.cu:405, :1088). */
static inline int off_index(int dx, int dy, int dz) { return (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1); }

void orc_assemble(int nx, int ny, int nz, const double* ke, const double Ke[64], double* A) {
  long n = (long)nx * ny * nz;
  memset(A, 0, sizeof(double) * 27 * n);
  /* node-centric gather so the loop parallelises without write conflicts */
#pragma omp parallel for schedule(static)
  for (long v = 0; v < n; v++) {
    int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((long)nx * ny));
    for (int e = 0; e < 8; e++) {
      int ex = (e & 1) ? x : x - 1, ey = (e & 2) ? y : y - 1, ez = (e & 4) ? z : z - 1;
      /* local index of v in this element */
      int lv = ((e & 1) ? 0 : 1) + ((e & 2) ? 0 : 2) + ((e & 4) ? 0 : 4);
      double k = ke[idx3(wrap(ex, nx), wrap(ey, ny), wrap(ez, nz), nx, ny)];
      for (int m = 0; m < 8; m++) {
        int dx = ex + (m & 1) - x, dy = ey + ((m >> 1) & 1) - y, dz = ez + ((m >> 2) & 1) - z;
        A[(long)off_index(dx, dy, dz) * n + v] += k * Ke[lv * 8 + m];
      }
    }
  }
}

void orc_matvec(int nx, int ny, int nz, const double* A, const double* xv, double* yv) {
  long n = (long)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (long v = 0; v < n; v++) {
    int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((long)nx * ny));
    double s = 0;
    for (int o = 0; o < 27; o++) {
      int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
      s += A[(long)o * n + v] * xv[idx3(wrap(x + dx, nx), wrap(y + dy, ny), wrap(z + dz, nz), nx, ny)];
    }
    yv[v] = s;
  }
}

/* RHS of load case k: grid.cu:1282-1342 (enforce_unit_macro_strain_kernel_H):
 * f_v = sum over the 8 incident elements of (rhoPenalMin + kappa_e) *
 * feMu[7 - ei][k], ei the element's position around the vertex. */
void orc_rhs(int nx, int ny, int nz, const double* ke, const double feMu[24], int k, double* f) {
  long n = (long)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (long v = 0; v < n; v++) {
    int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((long)nx * ny));
    double s = 0;
    for (int e = 0; e < 8; e++) {
      int ex = (e & 1) ? x : x - 1, ey = (e & 2) ? y : y - 1, ez = (e & 4) ? z : z - 1;
      int lv = 7 - e;
      s += ke[idx3(wrap(ex, nx), wrap(ey, ny), wrap(ez, nz), nx, ny)] * feMu[lv * 3 + k];
    }
    f[v] = s;
  }
}

/* ---- CG ------------------------------------------------------------------
 * The reference solves with a stationary MG V-cycle (MG.cpp:80-142) and a
 * Dirichlet pin at the origin node (grid.cu:183-188). The oracle solves the
 * same linear system to a much tighter tolerance with Jacobi-PCG, either
 * with the pin (pin=1) or on the mean-zero subspace (pin=0). */
static double dot(long n, const double* a, const double* b) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}
static void remove_mean(long n, double* a) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long i = 0; i < n; i++) s += a[i];
  s /= (double)n;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) a[i] -= s;
}

int orc_cg(int nx, int ny, int nz, const double* A, const double* b, double* x,
           double rtol, int maxit, int pin, double* relres) {
  long n = (long)nx * ny * nz;
  double* r = (double*)malloc(sizeof(double) * n);
  double* z = (double*)malloc(sizeof(double) * n);
  /* p starts at zero: the first update p = z + 0 * p must not see garbage */
  double* p = (double*)calloc((size_t)n, sizeof(double));
  double* q = (double*)malloc(sizeof(double) * n);
  const double* diag = A + 13 * n;
  double bn = sqrt(dot(n, b, b));
  if (bn == 0) bn = 1;
  if (pin) x[0] = 0;
  orc_matvec(nx, ny, nz, A, x, q);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) r[i] = b[i] - q[i];
  if (pin) r[0] = 0; else remove_mean(n, r);
  int it = 0;
  double rz = 0, res = sqrt(dot(n, r, r)) / bn;
  while (res > rtol && it < maxit) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) z[i] = r[i] / diag[i];
    if (pin) z[0] = 0; else remove_mean(n, z);
    double rzn = dot(n, r, z);
    double beta = it == 0 ? 0.0 : rzn / rz;
    rz = rzn;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    orc_matvec(nx, ny, nz, A, p, q);
    if (pin) q[0] = 0;
    double alpha = rz / dot(n, p, q);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    it++;
    /* recompute the true residual every 50 steps to avoid drift */
    if (it % 50 == 0) {
      orc_matvec(nx, ny, nz, A, x, q);
#pragma omp parallel for schedule(static)
      for (long i = 0; i < n; i++) r[i] = b[i] - q[i];
      if (pin) r[0] = 0; else remove_mean(n, r);
    }
    res = sqrt(dot(n, r, r)) / bn;
  }
  /* final true residual */
  orc_matvec(nx, ny, nz, A, x, q);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) r[i] = b[i] - q[i];
  if (pin) r[0] = 0; else remove_mean(n, r);
  if (relres) *relres = sqrt(dot(n, r, r)) / bn;
  free(r); free(z); free(p); free(q);
  return it;
}

/* ---- homogenized tensor and element energies -----------------------------
 * homogenization.cu:13-153 (heatMatrix_kernel_opt): per element the
 * macro-minus-fluctuation nodal field d_k = disp[:,k] - chi_k(nodes), energy
 * S_e(i,j) = d_i^T Ke d_j, C_ij = sum_e (rhoPenalMin + kappa_e) S_e(i,j) / |Y|
 * with |Y| = 1 (homogenization.cu:207). */
static void element_d(int nx, int ny, int ex, int ey, int ez, int nz_, long n,
                      const double* chi, const double disp[24], double d[3][8]) {
  for (int m = 0; m < 8; m++) {
    long v = idx3(wrap(ex + (m & 1), nx), wrap(ey + ((m >> 1) & 1), ny), wrap(ez + ((m >> 2) & 1), nz_), nx, ny);
    for (int k = 0; k < 3; k++) d[k][m] = disp[m * 3 + k] - chi[(long)k * n + v];
  }
}

void orc_element_energy(int nx, int ny, int nz, const double* chi, double* S6) {
  long n = (long)nx * ny * nz;
  double Ke[64], feMu[24], disp[24];
  orc_template(0.5 / nx, Ke, feMu, disp);
#pragma omp parallel for schedule(static)
  for (long e = 0; e < n; e++) {
    int x = (int)(e % nx), y = (int)((e / nx) % ny), z = (int)(e / ((long)nx * ny));
    double d[3][8], kd[3][8];
    element_d(nx, ny, x, y, z, nz, n, chi, disp, d);
    for (int k = 0; k < 3; k++)
      for (int a = 0; a < 8; a++) {
        double s = 0;
        for (int b = 0; b < 8; b++) s += Ke[a * 8 + b] * d[k][b];
        kd[k][a] = s;
      }
    int c = 0;
    for (int i = 0; i < 3; i++)
      for (int j = i; j < 3; j++) {
        double s = 0;
        for (int a = 0; a < 8; a++) s += d[i][a] * kd[j][a];
        S6[(long)c * n + e] = s;
        c++;
      }
  }
}

int orc_homogenize(int nx, int ny, int nz, const double* kappa, double rtol, int pin,
                   double* chi_out, double C[9], double* S_out) {
  long n = (long)nx * ny * nz;
  double Ke[64], feMu[24], disp[24];
  orc_template(0.5 / nx, Ke, feMu, disp);
  double* ke = (double*)malloc(sizeof(double) * n);
  for (long e = 0; e < n; e++) ke[e] = RHO_PENAL_MIN + kappa[e];
  double* A = (double*)malloc(sizeof(double) * 27 * n);
  orc_assemble(nx, ny, nz, ke, Ke, A);
  double* chi = chi_out ? chi_out : (double*)malloc(sizeof(double) * 3 * n);
  double* f = (double*)malloc(sizeof(double) * n);
  int maxits = 0;
  for (int k = 0; k < 3; k++) {
    orc_rhs(nx, ny, nz, ke, feMu, k, f);
    double* x = chi + (long)k * n;
    memset(x, 0, sizeof(double) * n);
    double rr;
    int it = orc_cg(nx, ny, nz, A, f, x, rtol, 20000, pin, &rr);
    if (it > maxits) maxits = it;
    /* report the mean-zero representative (translateForce, grid.cpp:654-664) */
    remove_mean(n, x);
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
  free(f); free(A); free(ke);
  return maxits;
}

/* homogenization.cu:255-412 (Sensitivity_kernel_opt_2_H) with penalty
 * exponent 1: sens_e = sum_{i<=j} S_e(i,j) * (dC_ij + [i != j] dC_ji). */
void orc_sensitivity(int n, const double* S6, const double dC[9], double* sens) {
  const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
  double w[6];
  for (int c = 0; c < 6; c++) {
    int i = ii[c], j = jj[c];
    w[c] = dC[i * 3 + j] + (i != j ? dC[j * 3 + i] : 0.0);
  }
#pragma omp parallel for schedule(static)
  for (long e = 0; e < n; e++) {
    double s = 0;
    for (int c = 0; c < 6; c++) s += S6[(long)c * n + e] * w[c];
    sens[e] = s;
  }
}

/* ---- filters ----------------------------------------------------------------
 * Radial weights (TensorExpression.h:1053-1086): Spline4 1-6q+8q^{3/2}-3q^2,
 * Linear 1-sqrt(q), with q = |off|^2/r^2 < 1, over a (2h+1)^3 window with
 * h = int(r + 0.5) (TensorExpression.h:1074-1093). The normalisation by the
 * total window weight cancels in the renormalised convolution and is omitted. */
static double radial_w(int kind, double r, int dx, int dy, int dz) {
  double q = (double)(dx * dx + dy * dy + dz * dz) / (r * r);
  if (q >= 1.0) return 0.0;
  double s = sqrt(q);
  if (kind == 0) return 1.0 - 6.0 * q + 8.0 * q * s - 3.0 * q * q;
  return 1.0 - s;
}

void orc_filter_fwd(int nx, int ny, int nz, int kind, double r, int periodic,
                    const double* in, double* out, double* W) {
  long n = (long)nx * ny * nz;
  int h = (int)(r + 0.5);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) {
    int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long)nx * ny));
    double s = 0, ws = 0;
    for (int dz = -h; dz <= h; dz++)
      for (int dy = -h; dy <= h; dy++)
        for (int dx = -h; dx <= h; dx++) {
          int X = x + dx, Y = y + dy, Z = z + dz;
          int inside = X >= 0 && X < nx && Y >= 0 && Y < ny && Z >= 0 && Z < nz;
          if (!inside && !periodic) continue;
          double w = radial_w(kind, r, dx, dy, dz);
          s += w * in[idx3(wrap(X, nx), wrap(Y, ny), wrap(Z, nz), nx, ny)];
          ws += w;
        }
    out[i] = s / ws;
    if (W) W[i] = ws;
  }
}

void orc_filter_adj(int nx, int ny, int nz, int kind, double r, int periodic,
                    const double* in, double* out) {
  long n = (long)nx * ny * nz;
  double* W = (double*)malloc(sizeof(double) * n);
  double* one = (double*)malloc(sizeof(double) * n);
  double* tmp = (double*)malloc(sizeof(double) * n);
  for (long i = 0; i < n; i++) one[i] = 1.0;
  orc_filter_fwd(nx, ny, nz, kind, r, periodic, one, tmp, W);
  int h = (int)(r + 0.5);
#pragma omp parallel for schedule(static)
  for (long j = 0; j < n; j++) {
    int x = (int)(j % nx), y = (int)((j / nx) % ny), z = (int)(j / ((long)nx * ny));
    double s = 0;
    for (int dz = -h; dz <= h; dz++)
      for (int dy = -h; dy <= h; dy++)
        for (int dx = -h; dx <= h; dx++) {
          int X = x - dx, Y = y - dy, Z = z - dz;
          int inside = X >= 0 && X < nx && Y >= 0 && Y < ny && Z >= 0 && Z < nz;
          if (!inside && !periodic) continue;
          long i = idx3(wrap(X, nx), wrap(Y, ny), wrap(Z, nz), nx, ny);
          s += radial_w(kind, r, dx, dy, dz) * in[i] / W[i];
        }
    out[j] = s;
  }
  free(W); free(one); free(tmp);
}

void orc_oc_filter_sens(int nx, int ny, int nz, double r, const double* rho,
                        const double* sens, double* out) {
  long n = (long)nx * ny * nz;
  int h = (int)(r + 0.5);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) {
    int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long)nx * ny));
    double s = 0, ws = 0;
    for (int dz = -h; dz <= h; dz++)
      for (int dy = -h; dy <= h; dy++)
        for (int dx = -h; dx <= h; dx++) {
          long j = idx3(wrap(x + dx, nx), wrap(y + dy, ny), wrap(z + dz, nz), nx, ny);
          double w = radial_w(1, r, dx, dy, dz);
          s += sens[j] * rho[j] * w;
          ws += w;
        }
    out[i] = s / (ws * rho[i]);
  }
}

static void oc_apply(long n, const double* sens, const double* rho, double* rn, double g,
                     double min_rho, double step, double damp) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) {
    double B = -sens[i] / g;
    if (B < 0) B = 0.01;
    double r0 = rho[i];
    double v = pow(B, damp) * r0;
    if (v - r0 < -step) v = r0 - step;
    if (v - r0 > step) v = r0 + step;
    if (v < min_rho) v = min_rho;
    if (v > 1) v = 1;
    rn[i] = v;
  }
}

int orc_oc_update(int n, const double* sens, const double* rho, double* rho_new,
                  double volratio, double min_rho, double step_limit, double damp,
                  int max_steps, double vol_tol, double* g_out) {
  double mx = 0;
  for (long i = 0; i < n; i++) if (fabs(sens[i]) > mx) mx = fabs(sens[i]);
  double lo = 0, hi = mx, g = 0;
  int it;
  for (it = 0; it < max_steps; it++) {
    g = 0.5 * (hi + lo);
    oc_apply(n, sens, rho, rho_new, g, min_rho, step_limit, damp);
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long i = 0; i < n; i++) s += rho_new[i];
    double vol = s / n;
    if (vol < volratio - vol_tol) hi = g;
    else if (vol > volratio + vol_tol) lo = g;
    else { it++; break; }
  }
  if (g_out) *g_out = g;
  return it;
}

/* ---- MMA -------------------------------------------------------------------
 * Restates mmasub_ker (mmaOpt.cu:834-1019) and subsolv_g (mmaOpt.cu:429-830):
 * asymptote update with asyinit 0.5, asyincr 1.2, asydecr 0.7 and the
 * [0.01, 10] (xmax-xmin) clamps; alfa/beta with albefa 0.1 and the move
 * limit; p0/q0/P/Q with 1e-3 and raa0 = 1e-5 regularisation; then the
 * interior point with epsi from 1 down to epsimin = 1e-7, Newton steps on the
 * (m+1) system [Alam a; a^T -zet/z] (solveLinearHost, mmaOpt.cpp:48-73),
 * step length from the 1.01 fraction-to-boundary rule and a halving line
 * search on the residual norm (at most 50 halvings, 200 Newton steps per
 * epsi level). */
static void small_solve(int k, double* A, double* b) {
  /* Gaussian elimination with partial pivoting, A k x k row-major */
  for (int c = 0; c < k; c++) {
    int p = c;
    for (int r = c + 1; r < k; r++)
      if (fabs(A[r * k + c]) > fabs(A[p * k + c])) p = r;
    if (p != c) {
      for (int j = 0; j < k; j++) { double t = A[c * k + j]; A[c * k + j] = A[p * k + j]; A[p * k + j] = t; }
      double t = b[c]; b[c] = b[p]; b[p] = t;
    }
    for (int r = c + 1; r < k; r++) {
      double f = A[r * k + c] / A[c * k + c];
      for (int j = c; j < k; j++) A[r * k + j] -= f * A[c * k + j];
      b[r] -= f * b[c];
    }
  }
  for (int r = k - 1; r >= 0; r--) {
    double s = b[r];
    for (int j = r + 1; j < k; j++) s -= A[r * k + j] * b[j];
    b[r] = s / A[r * k + r];
  }
}

typedef struct {
  int m, n;
  const double *low, *upp, *alfa, *beta, *p0, *q0, *P, *Q, *a, *b, *c, *d;
  double a0;
} mma_sub;

/* residual of the perturbed KKT system at (x..s); returns the norm, *mx the max */
static double mma_residual(const mma_sub* S, double epsi, const double* x, const double* y, double z,
                           const double* lam, const double* xsi, const double* eta, const double* mu,
                           double zet, const double* s, double* mx) {
  int m = S->m, n = S->n;
  double ss = 0, mm = 0;
  double gvec[16] = {0};
  for (int j = 0; j < n; j++) {
    double ux1 = S->upp[j] - x[j], xl1 = x[j] - S->low[j];
    double plam = S->p0[j], qlam = S->q0[j];
    for (int i = 0; i < m; i++) {
      plam += lam[i] * S->P[(long)i * n + j];
      qlam += lam[i] * S->Q[(long)i * n + j];
      gvec[i] += S->P[(long)i * n + j] / ux1 + S->Q[(long)i * n + j] / xl1;
    }
    double dpsidx = plam / (ux1 * ux1) - qlam / (xl1 * xl1);
    double rex = dpsidx - xsi[j] + eta[j];
    double rexsi = xsi[j] * (x[j] - S->alfa[j]) - epsi;
    double reeta = eta[j] * (S->beta[j] - x[j]) - epsi;
    ss += rex * rex + rexsi * rexsi + reeta * reeta;
    if (fabs(rex) > mm) mm = fabs(rex);
    if (fabs(rexsi) > mm) mm = fabs(rexsi);
    if (fabs(reeta) > mm) mm = fabs(reeta);
  }
  double alam = 0;
  for (int i = 0; i < m; i++) alam += S->a[i] * lam[i];
  double rez = S->a0 - zet - alam, rezet = zet * z - epsi;
  ss += rez * rez + rezet * rezet;
  if (fabs(rez) > mm) mm = fabs(rez);
  if (fabs(rezet) > mm) mm = fabs(rezet);
  for (int i = 0; i < m; i++) {
    double rey = S->c[i] + S->d[i] * y[i] - mu[i] - lam[i];
    double relam = gvec[i] - S->a[i] * z - y[i] + s[i] - S->b[i];
    double remu = mu[i] * y[i] - epsi;
    double res = lam[i] * s[i] - epsi;
    ss += rey * rey + relam * relam + remu * remu + res * res;
    if (fabs(rey) > mm) mm = fabs(rey);
    if (fabs(relam) > mm) mm = fabs(relam);
    if (fabs(remu) > mm) mm = fabs(remu);
    if (fabs(res) > mm) mm = fabs(res);
  }
  *mx = mm;
  return sqrt(ss);
}

/* Newton steps of the subsolver and those whose step length the design
 * variables' bounds set (not the small variables'): a test statistic. */
static long orc_mma_steps = 0, orc_mma_elem_bound_steps = 0;
static int mma_subsolv(const mma_sub* S, double epsimin, double* xout) {
  int m = S->m, n = S->n;
  double* x = (double*)malloc(sizeof(double) * n);
  double* xsi = (double*)malloc(sizeof(double) * n);
  double* eta = (double*)malloc(sizeof(double) * n);
  double* dx = (double*)malloc(sizeof(double) * n);
  double* xo = (double*)malloc(sizeof(double) * n);
  double* xsio = (double*)malloc(sizeof(double) * n);
  double* etao = (double*)malloc(sizeof(double) * n);
  double* dxsi = (double*)malloc(sizeof(double) * n);
  double* deta = (double*)malloc(sizeof(double) * n);
  double* delx = (double*)malloc(sizeof(double) * n);
  double* diagx = (double*)malloc(sizeof(double) * n);
  double* GG = (double*)malloc(sizeof(double) * (size_t)m * n);
  double y[16], lam[16], mu[16], s[16], z = 1, zet = 1;
  for (int j = 0; j < n; j++) {
    x[j] = 0.5 * (S->alfa[j] + S->beta[j]);
    xsi[j] = 1.0 / (x[j] - S->alfa[j]); if (xsi[j] < 1) xsi[j] = 1;
    eta[j] = 1.0 / (S->beta[j] - x[j]); if (eta[j] < 1) eta[j] = 1;
  }
  for (int i = 0; i < m; i++) {
    y[i] = 1; lam[i] = 1; s[i] = 1;
    mu[i] = 0.5 * S->c[i] > 1 ? 0.5 * S->c[i] : 1;
  }
  int total = 0;
  double epsi = 1;
  while (epsi > epsimin) {
    double residumax;
    double residunorm = mma_residual(S, epsi, x, y, z, lam, xsi, eta, mu, zet, s, &residumax);
    int ittt = 0;
    while (residumax > 0.9 * epsi && ittt < 200) {
      ittt++;
      total++;
      double gvec[16] = {0}, GGdd[16] = {0}, Alam[256] = {0};
      for (int j = 0; j < n; j++) {
        double ux1 = S->upp[j] - x[j], xl1 = x[j] - S->low[j];
        double ux2 = ux1 * ux1, xl2 = xl1 * xl1, ux3 = ux1 * ux2, xl3 = xl1 * xl2;
        double plam = S->p0[j], qlam = S->q0[j];
        for (int i = 0; i < m; i++) {
          double Pi = S->P[(long)i * n + j], Qi = S->Q[(long)i * n + j];
          plam += lam[i] * Pi;
          qlam += lam[i] * Qi;
          gvec[i] += Pi / ux1 + Qi / xl1;
          GG[(long)i * n + j] = Pi / ux2 - Qi / xl2;
        }
        double dpsidx = plam / ux2 - qlam / xl2;
        delx[j] = dpsidx - epsi / (x[j] - S->alfa[j]) + epsi / (S->beta[j] - x[j]);
        diagx[j] = 2 * (plam / ux3 + qlam / xl3) + xsi[j] / (x[j] - S->alfa[j]) + eta[j] / (S->beta[j] - x[j]);
        for (int i = 0; i < m; i++) {
          GGdd[i] += GG[(long)i * n + j] * delx[j] / diagx[j];
          for (int k = 0; k < m; k++) Alam[i * m + k] += GG[(long)i * n + j] * GG[(long)k * n + j] / diagx[j];
        }
      }
      double alam = 0;
      for (int i = 0; i < m; i++) alam += S->a[i] * lam[i];
      double delz = S->a0 - alam - epsi / z;
      int K = m + 1;
      double AA[289], bb[17];
      for (int i = 0; i < m; i++) {
        double dely = S->c[i] + S->d[i] * y[i] - lam[i] - epsi / y[i];
        double dellam = gvec[i] - S->a[i] * z - y[i] - S->b[i] + epsi / lam[i];
        double diagy = S->d[i] + mu[i] / y[i];
        double diaglamyi = s[i] / lam[i] + 1.0 / diagy;
        bb[i] = dellam + dely / diagy - GGdd[i];
        for (int k = 0; k < m; k++) AA[i * K + k] = Alam[i * m + k] + (i == k ? diaglamyi : 0.0);
        AA[i * K + m] = S->a[i];
        AA[m * K + i] = S->a[i];
      }
      AA[m * K + m] = -zet / z;
      bb[m] = delz;
      small_solve(K, AA, bb);
      double dlam[16], dy[16], dmu[16], ds[16];
      double dz = bb[m];
      for (int i = 0; i < m; i++) dlam[i] = bb[i];
      for (int j = 0; j < n; j++) {
        double g = 0;
        for (int i = 0; i < m; i++) g += GG[(long)i * n + j] * dlam[i];
        dx[j] = -delx[j] / diagx[j] - g / diagx[j];
        dxsi[j] = -xsi[j] + epsi / (x[j] - S->alfa[j]) - xsi[j] * dx[j] / (x[j] - S->alfa[j]);
        deta[j] = -eta[j] + epsi / (S->beta[j] - x[j]) + eta[j] * dx[j] / (S->beta[j] - x[j]);
      }
      for (int i = 0; i < m; i++) {
        double dely = S->c[i] + S->d[i] * y[i] - lam[i] - epsi / y[i];
        double diagy = S->d[i] + mu[i] / y[i];
        dy[i] = -dely / diagy + dlam[i] / diagy;
        dmu[i] = -mu[i] + epsi / y[i] - mu[i] * dy[i] / y[i];
        ds[i] = -s[i] + epsi / lam[i] - s[i] * dlam[i] / lam[i];
      }
      double dzet = -zet + epsi / z - zet * dz / z;
      /* fraction-to-boundary step */
      double stm = 1.0;
      double stm_small;
      for (int i = 0; i < m; i++) {
        double v;
        v = -1.01 * dy[i] / y[i]; if (v > stm) stm = v;
        v = -1.01 * dlam[i] / lam[i]; if (v > stm) stm = v;
        v = -1.01 * dmu[i] / mu[i]; if (v > stm) stm = v;
        v = -1.01 * ds[i] / s[i]; if (v > stm) stm = v;
      }
      { double v = -1.01 * dz / z; if (v > stm) stm = v; v = -1.01 * dzet / zet; if (v > stm) stm = v; }
      stm_small = stm;
      for (int j = 0; j < n; j++) {
        double v;
        v = -1.01 * dxsi[j] / xsi[j]; if (v > stm) stm = v;
        v = -1.01 * deta[j] / eta[j]; if (v > stm) stm = v;
        v = -1.01 * dx[j] / (x[j] - S->alfa[j]); if (v > stm) stm = v;
        v = 1.01 * dx[j] / (S->beta[j] - x[j]); if (v > stm) stm = v;
      }
      double steg = 1.0 / stm;
      orc_mma_steps++;
      if (stm > stm_small) orc_mma_elem_bound_steps++;
      memcpy(xo, x, sizeof(double) * n);
      memcpy(xsio, xsi, sizeof(double) * n);
      memcpy(etao, eta, sizeof(double) * n);
      double yo[16], lamo[16], muo[16], so[16], zo = z, zeto = zet;
      memcpy(yo, y, sizeof(double) * m); memcpy(lamo, lam, sizeof(double) * m);
      memcpy(muo, mu, sizeof(double) * m); memcpy(so, s, sizeof(double) * m);
      int itto = 0;
      double resinew = 2 * residunorm, rmax = residumax;
      while (resinew > residunorm && itto < 50) {
        itto++;
        for (int j = 0; j < n; j++) {
          x[j] = xo[j] + steg * dx[j];
          xsi[j] = xsio[j] + steg * dxsi[j];
          eta[j] = etao[j] + steg * deta[j];
        }
        for (int i = 0; i < m; i++) {
          y[i] = yo[i] + steg * dy[i];
          lam[i] = lamo[i] + steg * dlam[i];
          mu[i] = muo[i] + steg * dmu[i];
          s[i] = so[i] + steg * ds[i];
        }
        z = zo + steg * dz;
        zet = zeto + steg * dzet;
        resinew = mma_residual(S, epsi, x, y, z, lam, xsi, eta, mu, zet, s, &rmax);
        steg /= 2;
      }
      residunorm = resinew;
      residumax = rmax;
    }
    epsi *= 0.1;
  }
  memcpy(xout, x, sizeof(double) * n);
  free(x); free(xsi); free(eta); free(dx); free(xo); free(xsio); free(etao);
  free(dxsi); free(deta); free(delx); free(diagx); free(GG);
  return total;
}

void orc_mma_stats(long* steps, long* elem_bound_steps) {
  *steps = orc_mma_steps;
  *elem_bound_steps = orc_mma_elem_bound_steps;
}
void orc_mma_stats_reset(void) { orc_mma_steps = orc_mma_elem_bound_steps = 0; }

int orc_mmasub(int m, int n, int itn, double* x, const double* xmin, const double* xmax,
               double* xold1, double* xold2, const double* df0dx, const double* gval,
               const double* dgdx, double* low, double* upp, double a0, const double* a,
               const double* c, const double* d, double move, double epsimin) {
  const double raa0 = 1e-5, albefa = 0.1;
  const double asyinit = 0.5, asyincr = 1.2, asydecr = 0.7;
  if (m < 1 || m > 15) return -1;
  double* alfa = (double*)malloc(sizeof(double) * n);
  double* beta = (double*)malloc(sizeof(double) * n);
  double* p0 = (double*)malloc(sizeof(double) * n);
  double* q0 = (double*)malloc(sizeof(double) * n);
  double* P = (double*)malloc(sizeof(double) * (size_t)m * n);
  double* Q = (double*)malloc(sizeof(double) * (size_t)m * n);
  double b[16] = {0};
  for (int j = 0; j < n; j++) {
    double dxm = xmax[j] - xmin[j];
    if (itn <= 2) {
      low[j] = x[j] - asyinit * dxm;
      upp[j] = x[j] + asyinit * dxm;
    } else {
      double zzz = (x[j] - xold1[j]) * (xold1[j] - xold2[j]);
      double f = zzz > 0 ? asyincr : (zzz < 0 ? asydecr : 1.0);
      low[j] = x[j] - f * (xold1[j] - low[j]);
      upp[j] = x[j] + f * (upp[j] - xold1[j]);
      double lmin = x[j] - 10 * dxm, lmax = x[j] - 0.01 * dxm;
      double umin = x[j] + 0.01 * dxm, umax = x[j] + 10 * dxm;
      if (low[j] < lmin) low[j] = lmin;
      if (low[j] > lmax) low[j] = lmax;
      if (upp[j] > umax) upp[j] = umax;
      if (upp[j] < umin) upp[j] = umin;
    }
    double z1 = low[j] + albefa * (x[j] - low[j]), z2 = x[j] - move * dxm;
    double al = z1 > z2 ? z1 : z2;
    alfa[j] = al > xmin[j] ? al : xmin[j];
    double z3 = upp[j] - albefa * (upp[j] - x[j]), z4 = x[j] + move * dxm;
    double be = z3 < z4 ? z3 : z4;
    beta[j] = be < xmax[j] ? be : xmax[j];
    double xmami = dxm > 1e-5 ? dxm : 1e-5;
    double ux1 = upp[j] - x[j], xl1 = x[j] - low[j];
    double ux2 = ux1 * ux1, xl2 = xl1 * xl1;
    double dp = df0dx[j] > 0 ? df0dx[j] : 0, dq = -df0dx[j] > 0 ? -df0dx[j] : 0;
    double pq0 = 1e-3 * (dp + dq) + raa0 / xmami;
    p0[j] = (dp + pq0) * ux2;
    q0[j] = (dq + pq0) * xl2;
    for (int i = 0; i < m; i++) {
      double g = dgdx[(long)i * n + j];
      double pp = g > 0 ? g : 0, qq = -g > 0 ? -g : 0;
      double pq = 1e-3 * (pp + qq) + raa0 / xmami;
      P[(long)i * n + j] = (pp + pq) * ux2;
      Q[(long)i * n + j] = (qq + pq) * xl2;
      b[i] += P[(long)i * n + j] / ux1 + Q[(long)i * n + j] / xl1;
    }
  }
  for (int i = 0; i < m; i++) b[i] -= gval[i];
  mma_sub S = {m, n, low, upp, alfa, beta, p0, q0, P, Q, a, b, c, d, a0};
  double* xnew = (double*)malloc(sizeof(double) * n);
  int its = mma_subsolv(&S, epsimin, xnew);
  memcpy(xold2, xold1, sizeof(double) * n);
  memcpy(xold1, x, sizeof(double) * n);
  memcpy(x, xnew, sizeof(double) * n);
  free(alfa); free(beta); free(p0); free(q0); free(P); free(Q); free(xnew);
  return its;
}
