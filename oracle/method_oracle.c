/*
 * CPU oracle for the METHOD time-step hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Serial C restatement of the reference algorithm; every function names the
 * reference file:line it follows (paths relative to Project/CPU/ of
 * AlexJamesWright/METHOD). Only tests/, the smoke check and bench.py's CPU
 * baseline may load this library, and only as the checker.
 *
 * What pins each part (besides the reference runs kept as fixtures under
 * tests/golden/, which compare every field after whole steps):
 *   weno3/5        exact on linear data, reference KATs (test_weno.cc);
 *                  brute force against the Jiang-Shu closed form.
 *   flux vectors   stationary state -> zero net flux; x/y/z symmetry.
 *   primsToAll     closed form (test_srmhd.cc:228-282).
 *   C2P            round trip cons(prims) -> prims to 1e-12, residual at the
 *                  root, hydro (B = 0) and rest (v = 0) special cases.
 *   BCs            index patterns (test_boundaryConds.cc).
 *   dt             closed form (test_simulation.cc:34-72).
 */
#include "method_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ sizes */

typedef struct {
  int Nx, Ny, Nz, dims;
  int is, ie, js, je, ks, ke;
  long ntot;
  double dx, dy, dz;
} geom;

/* Data::initData (simData.cc:61-97) and Simulation::Simulation (simulation.cc:34-36). */
static geom mkgeom(const oc_config *c) {
  geom g;
  g.Nx = c->nx + 2 * c->Ng;
  g.Ny = c->ny + 2 * c->Ng;
  g.Nz = c->nz + 2 * c->Ng;
  g.dims = 3;
  double zmin = c->zmin, zmax = c->zmax, ymin = c->ymin, ymax = c->ymax;
  if (c->nz == 0) {
    g.Nz = 1;
    zmin = -1e20;
    zmax = 1e20;
    g.dims = 2;
  }
  if (c->ny == 0) {
    g.Nz = g.Ny = 1;
    zmin = ymin = -1e20;
    zmax = ymax = 1e20;
    g.dims = 1;
  }
  g.is = c->Ng; g.ie = g.Nx - c->Ng;
  g.js = c->Ng; g.je = g.Ny - c->Ng;
  g.ks = c->Ng; g.ke = g.Nz - c->Ng;
  if (g.dims < 3) { g.ks = 0; g.ke = 1; }
  if (g.dims < 2) { g.js = 0; g.je = 1; }
  g.ntot = (long)g.Nx * g.Ny * g.Nz;
  g.dx = (c->xmax - c->xmin) / c->nx;
  g.dy = (ymax - ymin) / c->ny;
  g.dz = (zmax - zmin) / c->nz;
  return g;
}

#define IX(v, i, j, k) ((long)(v) * g.ntot + ((long)(i) * g.Ny + (j)) * g.Nz + (k))

int oc_ncons(int model) { return model == OC_SRMHD ? 9 : 14; }
int oc_nprims(int model) { return model == OC_SRMHD ? 8 : 11; }
int oc_naux(int model) { return model == OC_SRMHD ? 13 : 17; }

/* -------------------------------------------------------------------- BCs */

/* Periodic / Outflow / Flow::apply (boundaryConds.cc:5-145,449-730): per-axis
 * copies, x first, then y when Ny > 1, then z when Nz > 1. */
void oc_apply_bcs(const oc_config *c, double *a, int nvar) {
  geom g = mkgeom(c);
  int Ng = c->Ng, nx = c->nx, ny = c->ny, nz = c->nz;
  for (int v = 0; v < nvar; v++)
    for (int i = 0; i < Ng; i++)
      for (int j = 0; j < g.Ny; j++)
        for (int k = 0; k < g.Nz; k++) {
          if (c->periodic[0]) {
            a[IX(v, i, j, k)] = a[IX(v, nx + i, j, k)];
            a[IX(v, nx + Ng + i, j, k)] = a[IX(v, Ng + i, j, k)];
          } else {
            a[IX(v, i, j, k)] = a[IX(v, Ng, j, k)];
            a[IX(v, nx + Ng + i, j, k)] = a[IX(v, nx + Ng - 1, j, k)];
          }
        }
  if (g.Ny > 1)
    for (int v = 0; v < nvar; v++)
      for (int i = 0; i < g.Nx; i++)
        for (int j = 0; j < Ng; j++)
          for (int k = 0; k < g.Nz; k++) {
            if (c->periodic[1]) {
              a[IX(v, i, j, k)] = a[IX(v, i, ny + j, k)];
              a[IX(v, i, ny + Ng + j, k)] = a[IX(v, i, Ng + j, k)];
            } else {
              a[IX(v, i, j, k)] = a[IX(v, i, Ng, k)];
              a[IX(v, i, ny + Ng + j, k)] = a[IX(v, i, ny + Ng - 1, k)];
            }
          }
  if (g.Nz > 1)
    for (int v = 0; v < nvar; v++)
      for (int i = 0; i < g.Nx; i++)
        for (int j = 0; j < g.Ny; j++)
          for (int k = 0; k < Ng; k++) {
            if (c->periodic[2]) {
              a[IX(v, i, j, k)] = a[IX(v, i, j, nz + k)];
              a[IX(v, i, j, nz + Ng + k)] = a[IX(v, i, j, Ng + k)];
            } else {
              a[IX(v, i, j, k)] = a[IX(v, i, j, Ng)];
              a[IX(v, i, j, nz + Ng + k)] = a[IX(v, i, j, nz + Ng - 1)];
            }
          }
}

static void apply_bcs_all(const oc_config *c, double *cons, double *prims, double *aux) {
  oc_apply_bcs(c, cons, oc_ncons(c->model));
  oc_apply_bcs(c, prims, oc_nprims(c->model));
  oc_apply_bcs(c, aux, oc_naux(c->model));
}

/* ----------------------------------------------------------------- SRMHD */

/* SRMHD::primsToAll (srmhd.cc:573-673). */
static void srmhd_prims_to_all(const oc_config *c, double *cons, double *prims, double *aux) {
  geom g = mkgeom(c);
  double gam = c->gamma;
  for (int i = 0; i < g.Nx; i++)
    for (int j = 0; j < g.Ny; j++)
      for (int k = 0; k < g.Nz; k++) {
#define P(v) prims[IX(v, i, j, k)]
#define U(v) cons[IX(v, i, j, k)]
#define A(v) aux[IX(v, i, j, k)]
        U(5) = P(5);
        U(6) = P(6);
        U(7) = P(7);
        A(11) = P(5) * P(5) + P(6) * P(6) + P(7) * P(7);
        U(8) = 0;
        A(9) = P(1) * P(1) + P(2) * P(2) + P(3) * P(3);
        A(1) = 1.0 / sqrt(1 - A(9));
        A(4) = A(1) * (P(1) * P(5) + P(2) * P(6) + P(3) * P(7));
        A(5) = P(5) / A(1) + A(4) * P(1);
        A(6) = P(6) / A(1) + A(4) * P(2);
        A(7) = P(7) / A(1) + A(4) * P(3);
        A(8) = (P(5) * P(5) + P(6) * P(6) + P(7) * P(7) + A(4) * A(4)) / (A(1) * A(1));
        A(0) = 1 + P(4) / P(0) * (gam / (gam - 1));
        A(2) = P(4) / (P(0) * (gam - 1));
        A(3) = sqrt(A(2) * gam * (gam - 1) / A(0));
        U(0) = P(0) * A(1);
        U(1) = (P(0) * A(0) + A(8)) * A(1) * A(1) * P(1) - A(4) * A(5);
        U(2) = (P(0) * A(0) + A(8)) * A(1) * A(1) * P(2) - A(4) * A(6);
        U(3) = (P(0) * A(0) + A(8)) * A(1) * A(1) * P(3) - A(4) * A(7);
        A(12) = U(1) * U(1) + U(2) * U(2) + U(3) * U(3);
        A(10) = P(5) * U(1) + P(6) * U(2) + P(7) * U(3);
        U(4) = (P(0) * A(0) + A(8)) * A(1) * A(1) - (P(4) + A(8) / 2.0) - A(4) * A(4) - U(0);
#undef P
#undef U
#undef A
      }
}

/* SRMHD::fluxVector (srmhd.cc:72-206). With quirks on, the z-direction flux
 * of phi is By as at srmhd.cc:200; off, it is Bz. */
static void srmhd_flux_vector(const oc_config *c, const double *cons, const double *prims, const double *aux,
                              double *f, int dir) {
  geom g = mkgeom(c);
  int vn = 1 + dir;      /* normal velocity prim index */
  int bn = 5 + dir;      /* normal B prim index */
  for (int i = 0; i < g.Nx; i++)
    for (int j = 0; j < g.Ny; j++)
      for (int k = 0; k < g.Nz; k++) {
#define P(v) prims[IX(v, i, j, k)]
#define U(v) cons[IX(v, i, j, k)]
#define A(v) aux[IX(v, i, j, k)]
#define F(v) f[IX(v, i, j, k)]
        double vdir = P(vn), Bdir = P(bn), W = A(1);
        F(0) = U(0) * vdir;
        for (int m = 0; m < 3; m++) {
          double val = U(1 + m) * vdir;
          /* the x form groups (p + bsq/2), the y and z forms add left to right */
          if (m == dir) val = (dir == 0) ? val + (P(4) + A(8) * 0.5) : val + P(4) + A(8) * 0.5;
          F(1 + m) = val - A(5 + m) * Bdir / W;
        }
        F(4) = (U(4) + P(4) + A(8) * 0.5) * vdir - A(4) * Bdir / W;
        for (int m = 0; m < 3; m++) {
          if (m == dir)
            F(5 + m) = U(8);
          else
            F(5 + m) = P(5 + m) * vdir - Bdir * P(1 + m);
        }
        F(8) = (dir == 2 && c->quirks) ? P(6) : Bdir;
#undef P
#undef U
#undef A
#undef F
      }
}

/* SRMHD::sourceTerm (srmhd.cc:233-250). */
static void srmhd_source(const oc_config *c, const double *cons, double *source) {
  geom g = mkgeom(c);
  for (int v = 0; v < 9; v++)
    for (long n = 0; n < g.ntot; n++)
      source[v * g.ntot + n] = (v == 8) ? -cons[8 * g.ntot + n] / (c->cp * c->cp) : 0.0;
}

/* Residual of srmhd.cc:269-295 and its analytic Jacobian; x[0] = v^2, x[1] = Z. */
static void srmhd_residual(double D, double tau, double Bsq, double Ssq, double BS, double gam,
                           const double *x, double *f, double *J) {
  double v2 = x[0], Z = x[1];
  double s = 1.0 - v2, sq = sqrt(s);
  double gf = (gam - 1.0) / gam;
  double p = gf * (Z * s - D * sq);
  double ZB = Z + Bsq, BS2 = BS * BS, iZ2 = 1.0 / (Z * Z);
  f[0] = ZB * ZB * v2 - (2.0 * Z + Bsq) * BS2 * iZ2 - Ssq;
  f[1] = ZB - p - 0.5 * Bsq * s - 0.5 * BS2 * iZ2 - D - tau;
  if (J) {
    double BS2iZ3 = BS2 * iZ2 / Z;
    J[0] = ZB * ZB;
    J[1] = 2.0 * ZB * (v2 + BS2iZ3);
    J[2] = gf * (Z - 0.5 * D / sq) + 0.5 * Bsq;
    J[3] = 1.0 - gf * s + BS2iZ3;
  }
}

#define SRMHD_NEWTON_TOL 1e-12
#define SRMHD_NEWTON_MAXIT 60

/* Newton on the 2x2 C2P system of srmhd.cc:269-295 (the reference solves the
 * same system with cminpack hybrd1, srmhd.cc:446-447). Converged when the
 * scaled residual max(|f0|/(Z+B^2)^2, |f1|/(Z+B^2)) <= 1e-12. Steps that leave
 * the physical region (v^2 in [0,1), Z > 0) are pulled back half way. */
int oc_srmhd_newton(double D, double Sx, double Sy, double Sz, double tau, double Bx, double By, double Bz,
                    double gamma, double *x) {
  double BS = Bx * Sx + By * Sy + Bz * Sz;
  double Bsq = Bx * Bx + By * By + Bz * Bz;
  double Ssq = Sx * Sx + Sy * Sy + Sz * Sz;
  if (!(x[0] >= 0.0 && x[0] < 1.0)) x[0] = 0.0;
  if (!(x[1] > 0.0)) x[1] = D + tau + Bsq;
  for (int it = 0; it <= SRMHD_NEWTON_MAXIT; it++) {
    double f[2], J[4];
    srmhd_residual(D, tau, Bsq, Ssq, BS, gamma, x, f, J);
    double ZB = x[1] + Bsq;
    double r = fmax(fabs(f[0]) / (ZB * ZB), fabs(f[1]) / ZB);
    if (r <= SRMHD_NEWTON_TOL) return it;
    if (!(r == r) || it == SRMHD_NEWTON_MAXIT) break;
    double det = J[0] * J[3] - J[1] * J[2];
    double d0 = (J[3] * f[0] - J[1] * f[1]) / det;
    double d1 = (J[0] * f[1] - J[2] * f[0]) / det;
    double v2 = x[0] - d0, Z = x[1] - d1;
    if (v2 < 0.0) v2 = 0.5 * x[0];
    if (v2 >= 1.0) v2 = 0.5 * (x[0] + 1.0);
    if (Z <= 0.0) Z = 0.5 * x[1];
    x[0] = v2;
    x[1] = Z;
  }
  return -1;
}

/* SRMHD::getPrimitiveVars (srmhd.cc:390-562): guess (v^2, rho h / (1 - v^2))
 * from the old prims/aux (srmhd.cc:438-442), root, then the recovery formulas
 * of srmhd.cc:509-557. A cell whose solve fails is retried from the mean
 * root of its converged interior face neighbours (srmhd.cc:465-507; the
 * reference uses the last cell's arguments there, which is reproduced with
 * quirks on). */
static long srmhd_c2p(const oc_config *c, const double *cons, double *prims, double *aux, oc_stats *st) {
  geom g = mkgeom(c);
  double gam = c->gamma;
  double *sol = (double *)malloc(sizeof(double) * 2 * g.ntot);
  char *ok = (char *)calloc(g.ntot, 1);
  long fails = 0;
  int li = g.ie - 1, lj = g.je - 1, lk = g.ke - 1; /* last interior cell */
  for (int i = g.is; i < g.ie; i++)
    for (int j = g.js; j < g.je; j++)
      for (int k = g.ks; k < g.ke; k++) {
        long n = ((long)i * g.Ny + j) * g.Nz + k;
#define P(v) prims[IX(v, i, j, k)]
#define U(v) cons[IX(v, i, j, k)]
#define A(v) aux[IX(v, i, j, k)]
        double x[2];
        x[0] = P(1) * P(1) + P(2) * P(2) + P(3) * P(3);
        x[1] = P(0) * A(0) / (1 - x[0]);
        int its = oc_srmhd_newton(U(0), U(1), U(2), U(3), U(4), U(5), U(6), U(7), gam, x);
        if (st) {
          st->cells++;
          if (its >= 0) {
            st->iterations += its;
            if (its > st->max_iterations) st->max_iterations = its;
          }
        }
        if (its < 0) {
          fails++;
        } else {
          ok[n] = 1;
          sol[2 * n] = x[0];
          sol[2 * n + 1] = x[1];
        }
#undef P
#undef U
#undef A
      }
  if (fails) {
    for (int i = g.is; i < g.ie; i++)
      for (int j = g.js; j < g.je; j++)
        for (int k = g.ks; k < g.ke; k++) {
          long n = ((long)i * g.Ny + j) * g.Nz + k;
          if (ok[n]) continue;
          int nb[6][3] = {{i - 1, j, k}, {i, j - 1, k}, {i, j, k - 1}, {i + 1, j, k}, {i, j + 1, k}, {i, j, k + 1}};
          double x[2] = {0.0, 0.0};
          int cnt = 0;
          for (int q = 0; q < 6; q++) {
            int a = nb[q][0], b = nb[q][1], d = nb[q][2];
            if (a < g.is || a >= g.ie || b < g.js || b >= g.je || d < g.ks || d >= g.ke) continue;
            long m = ((long)a * g.Ny + b) * g.Nz + d;
            if (!ok[m]) continue;
            x[0] += sol[2 * m];
            x[1] += sol[2 * m + 1];
            cnt++;
          }
          if (cnt) {
            x[0] /= cnt;
            x[1] /= cnt;
          }
          int si = c->quirks ? li : i, sj = c->quirks ? lj : j, sk = c->quirks ? lk : k;
          int its = oc_srmhd_newton(cons[IX(0, si, sj, sk)], cons[IX(1, si, sj, sk)], cons[IX(2, si, sj, sk)],
                                    cons[IX(3, si, sj, sk)], cons[IX(4, si, sj, sk)], cons[IX(5, si, sj, sk)],
                                    cons[IX(6, si, sj, sk)], cons[IX(7, si, sj, sk)], gam, x);
          if (st && its < 0) st->failures++;
          sol[2 * n] = x[0];
          sol[2 * n + 1] = x[1];
        }
  }
  for (int i = g.is; i < g.ie; i++)
    for (int j = g.js; j < g.je; j++)
      for (int k = g.ks; k < g.ke; k++) {
        long n = ((long)i * g.Ny + j) * g.Nz + k;
#define P(v) prims[IX(v, i, j, k)]
#define U(v) cons[IX(v, i, j, k)]
#define A(v) aux[IX(v, i, j, k)]
        double s0 = sol[2 * n], s1 = sol[2 * n + 1];
        P(5) = U(5);
        P(6) = U(6);
        P(7) = U(7);
        A(10) = U(5) * U(1) + U(6) * U(2) + U(7) * U(3);
        A(11) = U(5) * U(5) + U(6) * U(6) + U(7) * U(7);
        A(12) = U(1) * U(1) + U(2) * U(2) + U(3) * U(3);
        A(1) = 1 / sqrt(1 - s0);
        P(0) = U(0) / A(1);
        A(0) = s1 / (P(0) * A(1) * A(1));
        P(4) = (A(0) - 1) * P(0) * (gam - 1) / gam;
        A(2) = P(4) / (P(0) * (gam - 1));
        P(1) = (U(5) * A(10) + U(1) * s1) / (s1 * (A(11) + s1));
        P(2) = (U(6) * A(10) + U(2) * s1) / (s1 * (A(11) + s1));
        P(3) = (U(7) * A(10) + U(3) * s1) / (s1 * (A(11) + s1));
        A(9) = P(1) * P(1) + P(2) * P(2) + P(3) * P(3);
        A(3) = sqrt(A(2) * gam * (gam - 1) / A(0));
        A(4) = A(1) * (U(5) * P(1) + U(6) * P(2) + U(7) * P(3));
        A(5) = U(5) / A(1) + A(4) * P(1);
        A(6) = U(6) / A(1) + A(4) * P(2);
        A(7) = U(7) / A(1) + A(4) * P(3);
        A(8) = (P(5) * P(5) + P(6) * P(6) + P(7) * P(7) + A(4) * A(4)) / (A(1) * A(1));
#undef P
#undef U
#undef A
      }
  if (st) st->failures += fails;
  free(sol);
  free(ok);
  return fails;
}

/* ----------------------------------------------------------------- SRRMHD */

#define SRRMHD_TOL 1.0e-12
#define SRRMHD_EPS 1.0e-4
#define SRRMHD_MAXITER 50

/* Data::sigmaFunc for constant conductivity (simData.cc:114-119). */
static double sigma_of(const oc_config *c) { return c->sigma; }

/* SRRMHD::primsToAll (srrmhd.cc:358-510): qch from the 4th-order divergence
 * of E where the stencil fits, zero (with J) elsewhere. */
static void srrmhd_prims_to_all(const oc_config *c, double *cons, double *prims, double *aux) {
  geom g = mkgeom(c);
  double gam = c->gamma;
  for (int i = 0; i < g.Nx; i++)
    for (int j = 0; j < g.Ny; j++)
      for (int k = 0; k < g.Nz; k++) {
#define P(v) prims[IX(v, i, j, k)]
#define U(v) cons[IX(v, i, j, k)]
#define A(v) aux[IX(v, i, j, k)]
#define PE(v, a, b, d) prims[IX(v, a, b, d)]
        U(11) = U(12) = 0.0;
        U(5) = P(5);
        U(6) = P(6);
        U(7) = P(7);
        U(8) = P(8);
        U(9) = P(9);
        U(10) = P(10);
        A(7) = P(5) * P(5) + P(6) * P(6) + P(7) * P(7);
        A(8) = P(8) * P(8) + P(9) * P(9) + P(10) * P(10);
        A(9) = P(1) * P(1) + P(2) * P(2) + P(3) * P(3);
        A(11) = P(1) * U(8) + P(2) * U(9) + P(3) * U(10);
        A(1) = 1.0 / sqrt(1 - A(9));
        U(0) = P(0) * A(1);
        A(2) = P(4) / (P(0) * (gam - 1));
        A(0) = 1 + A(2) * gam;
        A(3) = sqrt((A(2) * gam * (gam - 1)) / A(0));
        A(10) = P(0) * A(0) * A(1) * A(1);
        double sigma = sigma_of(c);
        int have = 0;
        if (i > 1 && j > 1 && k > 1 && i < g.Nx - 2 && j < g.Ny - 2 && k < g.Nz - 2) {
          U(13) = (-PE(8, i + 2, j, k) + 8 * PE(8, i + 1, j, k) - 8 * PE(8, i - 1, j, k) + PE(8, i - 2, j, k)) /
                      (12 * g.dx) +
                  (-PE(9, i, j + 2, k) + 8 * PE(9, i, j + 1, k) - 8 * PE(9, i, j - 1, k) + PE(9, i, j - 2, k)) /
                      (12 * g.dy) +
                  (-PE(10, i, j, k + 2) + 8 * PE(10, i, j, k + 1) - 8 * PE(10, i, j, k - 1) + PE(10, i, j, k - 2)) /
                      (12 * g.dz);
          have = 1;
        } else if (i > 1 && j > 1 && i < g.Nx - 2 && j < g.Ny - 2 && c->nz == 0) {
          U(13) = (-PE(8, i + 2, j, k) + 8 * PE(8, i + 1, j, k) - 8 * PE(8, i - 1, j, k) + PE(8, i - 2, j, k)) /
                      (12 * g.dx) +
                  (-PE(9, i, j + 2, k) + 8 * PE(9, i, j + 1, k) - 8 * PE(9, i, j - 1, k) + PE(9, i, j - 2, k)) /
                      (12 * g.dy);
          have = 1;
        } else if (i > 1 && i < g.Nx - 2 && c->ny == 0) {
          U(13) = (-PE(8, i + 2, j, k) + 8 * PE(8, i + 1, j, k) - 8 * PE(8, i - 1, j, k) + PE(8, i - 2, j, k)) /
                  (12 * g.dx);
          have = 1;
        }
        if (have) {
          A(4) = U(13) * P(1) + A(1) * sigma * (U(8) + P(2) * U(7) - P(3) * U(6) - A(11) * P(1));
          A(5) = U(13) * P(2) + A(1) * sigma * (U(9) + P(3) * U(5) - P(1) * U(7) - A(11) * P(2));
          A(6) = U(13) * P(3) + A(1) * sigma * (U(10) + P(1) * U(6) - P(2) * U(5) - A(11) * P(3));
        } else {
          U(13) = A(4) = A(5) = A(6) = 0.0;
        }
        U(1) = A(10) * P(1) + U(9) * U(7) - U(10) * U(6);
        U(2) = A(10) * P(2) + U(10) * U(5) - U(8) * U(7);
        U(3) = A(10) * P(3) + U(8) * U(6) - U(9) * U(5);
        U(4) = A(10) - P(4) + 0.5 * (A(7) + A(8)) - U(0);
        A(12) = A(10) * P(1);
        A(13) = A(10) * P(2);
        A(14) = A(10) * P(3);
        A(15) = A(12) * A(12) + A(13) * A(13) + A(14) * A(14);
        A(16) = A(10) - P(4) - U(0);
#undef P
#undef U
#undef A
#undef PE
      }
}

/* SRRMHD::fluxVector (srrmhd.cc:65-192). With quirks on, the z-direction Sy
 * flux is overwritten by the Sz formula and the Sz flux is left as whatever
 * the caller's buffer held (srrmhd.cc:163-171): `f` then must hold the
 * y-direction flux on entry, as the reference's shared buffer does. */
static void srrmhd_flux_vector(const oc_config *c, const double *cons, const double *prims, const double *aux,
                               double *f, int dir) {
  geom g = mkgeom(c);
  for (int i = 0; i < g.Nx; i++)
    for (int j = 0; j < g.Ny; j++)
      for (int k = 0; k < g.Nz; k++) {
#define P(v) prims[IX(v, i, j, k)]
#define U(v) cons[IX(v, i, j, k)]
#define A(v) aux[IX(v, i, j, k)]
#define F(v) f[IX(v, i, j, k)]
        if (dir == 0) {
          F(0) = U(0) * P(1);
          F(1) = A(10) * P(1) * P(1) + P(4) - U(8) * U(8) - U(5) * U(5) + 0.5 * (A(7) + A(8));
          F(2) = A(10) * P(1) * P(2) - U(8) * U(9) - U(5) * U(6);
          F(3) = A(10) * P(1) * P(3) - U(8) * U(10) - U(5) * U(7);
          F(4) = U(1) - U(0) * P(1);
          F(5) = U(12);
          F(6) = -U(10);
          F(7) = U(9);
          F(8) = U(11);
          F(9) = U(7);
          F(10) = -U(6);
          F(11) = U(8);
          F(12) = U(5);
          F(13) = A(4);
        } else if (dir == 1) {
          F(0) = U(0) * P(2);
          F(1) = A(10) * P(1) * P(2) - U(9) * U(8) - U(6) * U(5);
          F(2) = A(10) * P(2) * P(2) + P(4) - U(9) * U(9) - U(6) * U(6) + 0.5 * (A(7) + A(8));
          F(3) = A(10) * P(2) * P(3) - U(9) * U(10) - U(6) * U(7);
          F(4) = U(2) - U(0) * P(2);
          F(5) = U(10);
          F(6) = U(12);
          F(7) = -U(8);
          F(8) = -U(7);
          F(9) = U(11);
          F(10) = U(5);
          F(11) = U(9);
          F(12) = U(6);
          F(13) = A(5);
        } else {
          F(0) = U(0) * P(3);
          F(1) = A(10) * P(1) * P(3) - U(10) * U(8) - U(7) * U(5);
          if (c->quirks) {
            F(2) = A(10) * P(3) * P(3) + P(4) - U(10) * U(10) - U(7) * U(7) + 0.5 * (A(7) + A(8));
          } else {
            F(2) = A(10) * P(2) * P(3) - U(10) * U(9) - U(7) * U(6);
            F(3) = A(10) * P(3) * P(3) + P(4) - U(10) * U(10) - U(7) * U(7) + 0.5 * (A(7) + A(8));
          }
          F(4) = U(3) - U(0) * P(3);
          F(5) = -U(9);
          F(6) = U(8);
          F(7) = U(12);
          F(8) = U(6);
          F(9) = -U(5);
          F(10) = U(11);
          F(11) = U(10);
          F(12) = U(7);
          F(13) = A(6);
        }
#undef P
#undef U
#undef A
#undef F
      }
}

/* SRRMHD::sourceTermSingleCell (srrmhd.cc:194-207). */
static void srrmhd_source_cell(double cp, const double *u, const double *a, double *s) {
  s[0] = s[1] = s[2] = s[3] = s[4] = s[5] = s[6] = s[7] = s[13] = 0.0;
  s[8] = -a[4];
  s[9] = -a[5];
  s[10] = -a[6];
  s[11] = u[13] - u[11] / (cp * cp);
  s[12] = -u[12] / (cp * cp);
}

/* Residual of the secant C2P (srrmhd.cc:512-537). */
static double srrmhd_residual(double Z, double StildeSq, double D, double tauTilde, double gamma) {
  double vsq, W, rho, h, p, resid;
  vsq = StildeSq / (Z * Z);
  if (vsq >= 1.0 || Z < 0) return 1.0e6;
  W = 1 / sqrt(1 - vsq);
  rho = D / W;
  h = Z / (rho * W * W);
  p = (gamma - 1) * (h - rho) / gamma;
  if (rho < 0 || p < 0 || W < 1 || h < 0) return 1.0e6;
  resid = (1 - (gamma - 1) / (W * W * gamma)) * Z + ((gamma - 1) / (W * gamma) - 1) * D - tauTilde;
  return resid;
}

/* Secant iteration of srrmhd.cc:539-581. Returns iterations, or -1 when it
 * fails (the reference throws there); Z holds the best iterate either way. */
static int srrmhd_secant(double *Z, double StildeSq, double D, double tauTilde, double gamma) {
  double x0 = *Z, x1 = x0 + SRRMHD_EPS, x2;
  double f0 = srrmhd_residual(x0, StildeSq, D, tauTilde, gamma);
  double f1 = srrmhd_residual(x1, StildeSq, D, tauTilde, gamma);
  double bestX = x0, bestF = f0;
  for (int iter = 0; iter < SRRMHD_MAXITER; iter++) {
    if (fabs(f0) < SRRMHD_TOL) {
      *Z = x0;
      return iter;
    }
    x2 = x1 - f1 * (x1 - x0) / (f1 - f0);
    x1 = x0;
    x0 = x2;
    f1 = f0;
    f0 = srrmhd_residual(x0, StildeSq, D, tauTilde, gamma);
    if (fabs(f0) < fabs(bestF)) {
      bestX = x0;
      bestF = f0;
    }
  }
  *Z = bestX;
  return -1;
}

/* SRRMHD::getPrimitiveVarsSingleCell (srrmhd.cc:256-312): a[10] holds the
 * guess Z on entry. Returns secant iterations or -1. */
static int srrmhd_c2p_cell(double gam, double sigma, const double *u, double *p, double *a) {
  p[5] = u[5]; p[6] = u[6]; p[7] = u[7];
  p[8] = u[8]; p[9] = u[9]; p[10] = u[10];
  a[7] = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
  a[8] = u[8] * u[8] + u[9] * u[9] + u[10] * u[10];
  a[12] = u[1] - (u[9] * u[7] - u[10] * u[6]);
  a[13] = u[2] - (u[10] * u[5] - u[8] * u[7]);
  a[14] = u[3] - (u[8] * u[6] - u[9] * u[5]);
  a[15] = a[12] * a[12] + a[13] * a[13] + a[14] * a[14];
  a[16] = u[4] - 0.5 * (a[7] + a[8]);
  int its = srrmhd_secant(&a[10], a[15], u[0], a[16], gam);
  a[9] = a[15] / (a[10] * a[10]);
  a[1] = 1.0 / sqrt(1 - a[9]);
  p[0] = u[0] / a[1];
  a[0] = a[10] / (p[0] * a[1] * a[1]);
  a[2] = (a[0] - 1) / gam;
  a[3] = sqrt((a[2] * gam * (gam - 1)) / a[0]);
  p[4] = p[0] * a[2] * (gam - 1);
  p[1] = a[12] / a[10];
  p[2] = a[13] / a[10];
  p[3] = a[14] / a[10];
  a[11] = p[1] * u[8] + p[2] * u[9] + p[3] * u[10];
  a[4] = u[13] * p[1] + a[1] * sigma * (u[8] + (p[2] * u[7] - p[3] * u[6]) - a[11] * p[1]);
  a[5] = u[13] * p[2] + a[1] * sigma * (u[9] + (p[3] * u[5] - p[1] * u[7]) - a[11] * p[2]);
  a[6] = u[13] * p[3] + a[1] * sigma * (u[10] + (p[1] * u[6] - p[2] * u[5]) - a[11] * p[3]);
  return its;
}

/* SRRMHD::getPrimitiveVars (srrmhd.cc:314-356): guess Z = aux[10] of the cell. */
static long srrmhd_c2p(const oc_config *c, const double *cons, double *prims, double *aux, oc_stats *st) {
  geom g = mkgeom(c);
  long fails = 0;
  double u[14], p[11], a[17];
  for (int i = g.is; i < g.ie; i++)
    for (int j = g.js; j < g.je; j++)
      for (int k = g.ks; k < g.ke; k++) {
        for (int v = 0; v < 14; v++) u[v] = cons[IX(v, i, j, k)];
        a[10] = aux[IX(10, i, j, k)];
        int its = srrmhd_c2p_cell(c->gamma, sigma_of(c), u, p, a);
        if (st) {
          st->cells++;
          if (its >= 0) {
            st->iterations += its;
            if (its > st->max_iterations) st->max_iterations = its;
          }
        }
        if (its < 0) fails++;
        for (int v = 0; v < 11; v++) prims[IX(v, i, j, k)] = p[v];
        for (int v = 0; v < 17; v++) aux[IX(v, i, j, k)] = a[v];
      }
  if (st) st->failures += fails;
  return fails;
}

/* ------------------------------------------------------------ model API */

void oc_prims_to_all(const oc_config *c, double *cons, double *prims, double *aux) {
  if (c->model == OC_SRMHD)
    srmhd_prims_to_all(c, cons, prims, aux);
  else
    srrmhd_prims_to_all(c, cons, prims, aux);
}

void oc_set(const oc_config *c, double *cons, double *prims, double *aux) {
  oc_prims_to_all(c, cons, prims, aux);
  apply_bcs_all(c, cons, prims, aux);
}

void oc_flux_vector(const oc_config *c, const double *cons, const double *prims, const double *aux, double *f,
                    int dir) {
  if (c->model == OC_SRMHD)
    srmhd_flux_vector(c, cons, prims, aux, f, dir);
  else
    srrmhd_flux_vector(c, cons, prims, aux, f, dir);
}

void oc_source(const oc_config *c, const double *cons, const double *prims, const double *aux, double *source) {
  (void)prims;
  if (c->model == OC_SRMHD) {
    srmhd_source(c, cons, source);
    return;
  }
  geom g = mkgeom(c);
  double u[14], a[17], s[14];
  for (long n = 0; n < g.ntot; n++) {
    for (int v = 0; v < 14; v++) u[v] = cons[v * g.ntot + n];
    for (int v = 0; v < 17; v++) a[v] = aux[v * g.ntot + n];
    srrmhd_source_cell(c->cp, u, a, s);
    for (int v = 0; v < 14; v++) source[v * g.ntot + n] = s[v];
  }
}

long oc_c2p(const oc_config *c, const double *cons, double *prims, double *aux, oc_stats *st) {
  if (c->model == OC_SRMHD) return srmhd_c2p(c, cons, prims, aux, st);
  return srrmhd_c2p(c, cons, prims, aux, st);
}

/* ------------------------------------------------------------------ WENO */

/* weno3_upwind (wenoUpwinds.cc:5-33): Jiang-Shu weights alpha = d/(eps + beta^2). */
double oc_weno3_upwind(double vec0, double vec1, double vec2) {
  double alpha0 = 0, alpha1 = 0, beta0 = 0, beta1 = 0, stencil0 = 0, stencil1 = 0, w0, w1;
  double epsilon = 1e-16, alphaSum;
  beta0 += 1.0 * vec1 * vec1;
  beta0 += -2.0 * vec0 * vec1;
  beta0 += 1.0 * vec0 * vec0;
  alpha0 = (1.0 / 3.0) / (epsilon + beta0 * beta0);
  beta1 += 1.0 * vec2 * vec2;
  beta1 += -2.0 * vec1 * vec2;
  beta1 += 1.0 * vec1 * vec1;
  alpha1 = (2.0 / 3.0) / (epsilon + beta1 * beta1);
  alphaSum = alpha0 + alpha1;
  w0 = alpha0 / alphaSum;
  w1 = alpha1 / alphaSum;
  stencil0 += 1.5 * vec1;
  stencil0 += -0.5 * vec0;
  stencil1 += 0.5 * vec2;
  stencil1 += 0.5 * vec1;
  return (w0 * stencil0 + w1 * stencil1);
}

/* weno5_upwind (wenoUpwinds.cc:36-84). The smoothness indicators are the
 * Jiang-Shu forms 13/12 (.)^2 + 1/4 (.)^2 expanded into monomials, as the
 * reference writes them; coefficients are the exact fractions. */
double oc_weno5_upwind(double vec0, double vec1, double vec2, double vec3, double vec4) {
  double alpha0, alpha1, alpha2, beta0 = 0, beta1 = 0, beta2 = 0, stencil0 = 0, stencil1 = 0, stencil2 = 0;
  double w0, w1, w2, epsilon = 1e-16, alphaSum;
  beta0 += (10.0 / 3.0) * vec2 * vec2;
  beta0 += (-31.0 / 3.0) * vec1 * vec2;
  beta0 += (25.0 / 3.0) * vec1 * vec1;
  beta0 += (11.0 / 3.0) * vec0 * vec2;
  beta0 += (-19.0 / 3.0) * vec0 * vec1;
  beta0 += (4.0 / 3.0) * vec0 * vec0;
  alpha0 = 0.1 / (epsilon + beta0 * beta0);
  beta1 += (4.0 / 3.0) * vec3 * vec3;
  beta1 += (-13.0 / 3.0) * vec2 * vec3;
  beta1 += (13.0 / 3.0) * vec2 * vec2;
  beta1 += (5.0 / 3.0) * vec1 * vec3;
  beta1 += (-13.0 / 3.0) * vec1 * vec2;
  beta1 += (4.0 / 3.0) * vec1 * vec1;
  alpha1 = 0.6 / (epsilon + beta1 * beta1);
  beta2 += (4.0 / 3.0) * vec4 * vec4;
  beta2 += (-19.0 / 3.0) * vec3 * vec4;
  beta2 += (25.0 / 3.0) * vec3 * vec3;
  beta2 += (11.0 / 3.0) * vec2 * vec4;
  beta2 += (-31.0 / 3.0) * vec2 * vec3;
  beta2 += (10.0 / 3.0) * vec2 * vec2;
  alpha2 = 0.3 / (epsilon + beta2 * beta2);
  alphaSum = alpha0 + alpha1 + alpha2;
  w0 = alpha0 / alphaSum;
  w1 = alpha1 / alphaSum;
  w2 = alpha2 / alphaSum;
  stencil0 += (11.0 / 6.0) * vec2;
  stencil0 += (-7.0 / 6.0) * vec1;
  stencil0 += (1.0 / 3.0) * vec0;
  stencil1 += (1.0 / 3.0) * vec3;
  stencil1 += (5.0 / 6.0) * vec2;
  stencil1 += (-1.0 / 6.0) * vec1;
  stencil2 += (-1.0 / 6.0) * vec4;
  stencil2 += (5.0 / 6.0) * vec3;
  stencil2 += (1.0 / 3.0) * vec2;
  return (w0 * stencil0 + w1 * stencil1 + w2 * stencil2);
}

/* Upwind/downwind stencil gathers of Weno3/Weno5 (weno.cc:123-229): the
 * upwind value at index i uses arr[i-shift .. i-shift+order-1], the
 * downwind one arr[i+shift-1 .. i+shift-order], over i in [shift, N-shift). */
static double recon_up(int order, const double *a, long i0, long stride) {
  int shift = (order + 1) / 2;
  const double *p = a + i0 - shift * stride;
  if (order == 3) return oc_weno3_upwind(p[0], p[stride], p[2 * stride]);
  return oc_weno5_upwind(p[0], p[stride], p[2 * stride], p[3 * stride], p[4 * stride]);
}

static double recon_down(int order, const double *a, long i0, long stride) {
  int shift = (order + 1) / 2;
  const double *p = a + i0 + (shift - 1) * stride;
  if (order == 3) return oc_weno3_upwind(p[0], p[-stride], p[-2 * stride]);
  return oc_weno5_upwind(p[0], p[-stride], p[-2 * stride], p[-3 * stride], p[-4 * stride]);
}

/* FVS::fluxReconstruction (fluxVectorSplitting.cc:4-56): Lax-Friedrichs
 * split with alpha = 1, upwind WENO on f+, downwind on f-, summed; entries
 * outside [shift, N-shift) along dir are zero. */
/* FVS::fluxReconstruction with `nv` variables (the reference's vars
 * argument, fluxVectorSplitting.cc:13-14). */
static void flux_recon_n(const oc_config *c, const double *cons, const double *f, double *frecon, int dir, int nv) {
  geom g = mkgeom(c);
  long n = g.ntot * nv;
  double *fp = (double *)calloc(n, sizeof(double));
  double *fm = (double *)calloc(n, sizeof(double));
  const double alpha = 1.0;
  for (long q = 0; q < n; q++) {
    fp[q] = 0.5 * (f[q] + alpha * cons[q]);
    fm[q] = 0.5 * (f[q] - alpha * cons[q]);
  }
  int order = c->weno_order, shift = (order + 1) / 2;
  long stride = dir == 0 ? (long)g.Ny * g.Nz : (dir == 1 ? g.Nz : 1);
  int N = dir == 0 ? g.Nx : (dir == 1 ? g.Ny : g.Nz);
  memset(frecon, 0, sizeof(double) * n);
  for (int v = 0; v < nv; v++)
    for (int i = 0; i < g.Nx; i++)
      for (int j = 0; j < g.Ny; j++)
        for (int k = 0; k < g.Nz; k++) {
          int pos = dir == 0 ? i : (dir == 1 ? j : k);
          if (pos < shift || pos >= N - shift) continue;
          long q = IX(v, i, j, k);
          double up = recon_up(order, fp, q, stride);
          double dn = recon_down(order, fm, q, stride);
          frecon[q] = up + dn;
        }
  free(fp);
  free(fm);
}

void oc_flux_reconstruction(const oc_config *c, const double *cons, const double *f, double *frecon, int dir) {
  flux_recon_n(c, cons, f, frecon, dir, oc_ncons(c->model));
}

/* FVS::F (fluxVectorSplitting.cc:58-131): per direction fluxVector then
 * reconstruction, fnet over interior = sum_dir (f[i+1]/d - f[i]/d). The flux
 * vector buffer is shared between directions, as the reference's d->f is. */
void oc_fvs_F(const oc_config *c, const double *cons, const double *prims, const double *aux, double *fnet) {
  geom g = mkgeom(c);
  int nv = oc_ncons(c->model);
  long n = g.ntot * nv;
  double *f = (double *)calloc(n, sizeof(double));
  double *fr[3] = {NULL, NULL, NULL};
  for (int dir = 0; dir < g.dims; dir++) {
    fr[dir] = (double *)malloc(sizeof(double) * n);
    oc_flux_vector(c, cons, prims, aux, f, dir);
    oc_flux_reconstruction(c, cons, f, fr[dir], dir);
  }
  double dd[3] = {g.dx, g.dy, g.dz};
  long st[3] = {(long)g.Ny * g.Nz, g.Nz, 1};
  int Ng = c->Ng;
  int i1 = g.Nx - Ng, j0 = g.dims >= 2 ? Ng : 0, j1 = g.dims >= 2 ? g.Ny - Ng : 1;
  int k0 = g.dims == 3 ? Ng : 0, k1 = g.dims == 3 ? g.Nz - Ng : 1;
  for (int v = 0; v < nv; v++)
    for (int i = Ng; i < i1; i++)
      for (int j = j0; j < j1; j++)
        for (int k = k0; k < k1; k++) {
          long q = IX(v, i, j, k);
          double acc = 0.0;
          for (int dir = 0; dir < g.dims; dir++) {
            double term = fr[dir][q + st[dir]] / dd[dir] - fr[dir][q] / dd[dir];
            acc = dir == 0 ? term : acc + term;
          }
          fnet[q] = acc;
        }
  for (int dir = 0; dir < 3; dir++) free(fr[dir]);
  free(f);
}

/* ------------------------------------------------------------ time step */

/* Simulation::updateTime's dt policy (simulation.cc:113-124), alpha = 1. */
double oc_timestep(const oc_config *c, double t, int iters, double end_time) {
  geom g = mkgeom(c);
  double sd = sqrt((double)g.dims);
  double dtX = c->cfl * g.dx / (1.0 * sd);
  double dtY = c->cfl * g.dy / (1.0 * sd);
  double dtZ = c->cfl * g.dz / (1.0 * sd);
  double dt = (dtX <= dtY && dtX <= dtZ) ? dtX : ((dtY < dtZ) ? dtY : dtZ);
  if (iters < 15) dt *= 0.1;
  if (iters >= 15 && iters < 30) dt *= 0.25;
  if (t + dt > end_time) dt = end_time - t;
  return dt;
}

/* Largest |characteristic speed| along axis d of an ideal SRMHD state: the
 * fast magnetosonic speed a (a^2 = cs^2 + va^2 - cs^2 va^2, va^2 = b^2/(rho h + b^2))
 * taken as isotropic in the fluid frame and carried to the lab frame
 * (Gammie, McKinney & Toth 2003, ApJ 589, 444, sec. 3.3, flat space). The
 * reference has no such bound: its dt and flux split use the light speed
 * (simulation.cc:39-41, fluxVectorSplitting.cc:10). */
double oc_srmhd_max_speed(double rho, double vd, double v2, double h, double c, double b2) {
  double cs2 = c * c;
  double va2 = b2 / (rho * h + b2);
  double a2 = cs2 + va2 - cs2 * va2;
  double den = 1.0 - v2 * a2;
  double disc = a2 * (1.0 - v2) * (den - vd * vd * (1.0 - a2));
  double sq = sqrt(disc > 0.0 ? disc : 0.0);
  double b = vd * (1.0 - a2);
  double lp = fabs(b + sq), lm = fabs(b - sq);
  return (lp > lm ? lp : lm) / den;
}

/* Maximum of oc_srmhd_max_speed over interior cells, per axis (0 for axes
 * the grid does not have). aux entries: h = 0, c = 3, b^2 = 8, v^2 = 9. */
void oc_max_wavespeed(const oc_config *c, const double *prims, const double *aux, double out[3]) {
  geom g = mkgeom(c);
  out[0] = out[1] = out[2] = 0.0;
  for (int i = g.is; i < g.ie; i++)
    for (int j = g.js; j < g.je; j++)
      for (int k = g.ks; k < g.ke; k++)
        for (int d = 0; d < g.dims; d++) {
          double s = oc_srmhd_max_speed(prims[IX(0, i, j, k)], prims[IX(1 + d, i, j, k)], aux[IX(9, i, j, k)],
                                        aux[IX(0, i, j, k)], aux[IX(3, i, j, k)], aux[IX(8, i, j, k)]);
          if (s > out[d]) out[d] = s;
        }
}

/* TimeIntegrator::finalise (timeInt.h:117-132): C2P, then BCs. */
static long finalise(const oc_config *c, double *cons, double *prims, double *aux, oc_stats *st) {
  long fails = oc_c2p(c, cons, prims, aux, st);
  apply_bcs_all(c, cons, prims, aux);
  return fails;
}

/* RK2::predictorStep, correctorStep, step (RK2.cc:31-100). */
long oc_rk2_step(const oc_config *c, double *cons, double *prims, double *aux, double dt, oc_stats *st) {
  geom g = mkgeom(c);
  int nc = oc_ncons(c->model), np = oc_nprims(c->model), na = oc_naux(c->model);
  double *p1cons = (double *)calloc(g.ntot * nc, sizeof(double));
  double *p1prims = (double *)malloc(sizeof(double) * g.ntot * np);
  double *p1aux = (double *)malloc(sizeof(double) * g.ntot * na);
  double *args = (double *)calloc(g.ntot * nc, sizeof(double));
  memcpy(p1prims, prims, sizeof(double) * g.ntot * np);
  memcpy(p1aux, aux, sizeof(double) * g.ntot * na);
  memcpy(p1cons, cons, sizeof(double) * g.ntot * nc);

  oc_fvs_F(c, cons, prims, aux, args);
  for (int v = 0; v < nc; v++)
    for (int i = g.is; i < g.ie; i++)
      for (int j = g.js; j < g.je; j++)
        for (int k = g.ks; k < g.ke; k++) p1cons[IX(v, i, j, k)] = cons[IX(v, i, j, k)] - dt * args[IX(v, i, j, k)];
  long fails = finalise(c, p1cons, p1prims, p1aux, st);

  oc_fvs_F(c, p1cons, p1prims, p1aux, args);
  for (int v = 0; v < nc; v++)
    for (int i = g.is; i < g.ie; i++)
      for (int j = g.js; j < g.je; j++)
        for (int k = g.ks; k < g.ke; k++) {
          long q = IX(v, i, j, k);
          cons[q] = 0.5 * (cons[q] + p1cons[q] - dt * args[q]);
        }
  fails += finalise(c, cons, prims, aux, st);
  free(p1cons);
  free(p1prims);
  free(p1aux);
  free(args);
  return fails;
}

/* The model extension's source added to `src` over the interior
 * (RKSplit::setSource, rkSplit.cc:15-25; RKPlus::rhs, RKPlus.cc:26-38). */
static void add_extension_source(const oc_config *c, const double *cons, const double *prims, const double *aux,
                                 double *src) {
  if (!c->regime) return;
  geom g = mkgeom(c);
  int nc = oc_ncons(c->model);
  double *ext = (double *)calloc(g.ntot * nc, sizeof(double));
  oc_regime_source(c, cons, prims, aux, ext);
  for (int v = 0; v < nc; v++)
    for (int i = g.is; i < g.ie; i++)
      for (int j = g.js; j < g.je; j++)
        for (int k = g.ks; k < g.ke; k++) src[IX(v, i, j, k)] += ext[IX(v, i, j, k)];
  free(ext);
}

/* RKSplit::step (rkSplit.cc:31-53): RK2::step, then cons += dt * source over
 * the interior (source from Model::sourceTerm on the whole grid), then
 * finalise. */
long oc_rksplit_step(const oc_config *c, double *cons, double *prims, double *aux, double dt, oc_stats *st) {
  long fails = oc_rk2_step(c, cons, prims, aux, dt, st);
  geom g = mkgeom(c);
  int nc = oc_ncons(c->model);
  double *src = (double *)calloc(g.ntot * nc, sizeof(double));
  oc_source(c, cons, prims, aux, src);
  add_extension_source(c, cons, prims, aux, src);
  for (int v = 0; v < nc; v++)
    for (int i = g.is; i < g.ie; i++)
      for (int j = g.js; j < g.je; j++)
        for (int k = g.ks; k < g.ke; k++) cons[IX(v, i, j, k)] += dt * src[IX(v, i, j, k)];
  free(src);
  fails += finalise(c, cons, prims, aux, st);
  return fails;
}

/* --------------------------------------------------------------- REGIME */

/* The REGIME resistive extension of ideal SRMHD (REGIME.cc; Wright & Hawke
 * 2019), restated on the padded grid with the reference's loop ranges. Per
 * cell: E = -v x B; the charge density q = div E by fourth-order central
 * differences where the stencil fits ([2, N-2) on every active axis, zero
 * elsewhere, REGIME.cc:186-212); alpha = 1 / ((q^2+s^2)((q^2+s^2) + s^2 B^2));
 * dw/ds_b (12 x 3, REGIME.cc:226-290) times alpha; K = the centred difference
 * of the stiff E-flux reconstruction (fluxReconstruction of E with flux
 * (0, Bz, -By) etc., REGIME.cc:293-370); per axis M = -df/dw . dw/ds_b
 * (8 x 3, REGIME.cc:467-774) and D = M K; the source is the centred
 * difference of D (REGIME.cc:92-138). */
#define DW(l, m) dw[(l) * 3 + (m)]
static void regime_dwdsb(double q, double sig, const double *P, const double *E, double alpha, double *dw) {
  double qsigsq = q * q + sig * sig, sigcu = sig * sig * sig, qch = q;
  for (int n = 0; n < 36; n++) dw[n] = 0.0;
  DW(1, 0) = -qch * (qch * qch + sig * sig * (1 + P[5] * P[5]));
  DW(1, 1) = -sig * (P[5] * P[6] * qch * sig - P[7] * qsigsq);
  DW(1, 2) = -sig * (P[5] * P[7] * qch * sig + P[6] * qsigsq);
  DW(2, 0) = -sig * (P[5] * P[6] * qch * sig + P[7] * qsigsq);
  DW(2, 1) = -qch * (qch * qch + sig * sig * (1 + P[6] * P[6]));
  DW(2, 2) = -sig * (P[6] * P[7] * qch * sig - P[5] * qsigsq);
  DW(3, 0) = -sig * (P[5] * P[7] * qch * sig - P[6] * qsigsq);
  DW(3, 1) = -sig * (P[6] * P[7] * qch * sig + P[5] * qsigsq);
  DW(3, 2) = -qch * (qch * qch + sig * sig * (1 + P[7] * P[7]));
  DW(5, 0) = -P[5] * sigcu * E[0];
  DW(5, 1) = -P[6] * sigcu * E[0] - sig * qsigsq * P[3];
  DW(5, 2) = -P[7] * sigcu * E[0] + sig * qsigsq * P[2];
  DW(6, 0) = -P[5] * sigcu * E[1] + sig * qsigsq * P[3];
  DW(6, 1) = -P[6] * sigcu * E[1];
  DW(6, 2) = -P[7] * sigcu * E[1] - sig * qsigsq * P[1];
  DW(7, 0) = -P[5] * sigcu * E[2] - sig * qsigsq * P[2];
  DW(7, 1) = -P[6] * sigcu * E[2] + sig * qsigsq * P[1];
  DW(7, 2) = -P[7] * sigcu * E[2];
  DW(8, 0) = -P[5] * P[5] * sigcu - sig * qsigsq;
  DW(8, 1) = -P[5] * P[6] * sigcu;
  DW(8, 2) = -P[5] * P[7] * sigcu;
  DW(9, 0) = -P[6] * P[5] * sigcu;
  DW(9, 1) = -P[6] * P[6] * sigcu - sig * qsigsq;
  DW(9, 2) = -P[6] * P[7] * sigcu;
  DW(10, 0) = -P[7] * P[5] * sigcu;
  DW(10, 1) = -P[7] * P[6] * sigcu;
  DW(10, 2) = -P[7] * P[7] * sigcu - sig * qsigsq;
  for (int n = 0; n < 36; n++) dw[n] *= alpha;
}

/* df/dw of axis `dir` (8 x 12, the entries REGIME::set_Dx/Dy/Dz set), then
 * D = M K with M = -df/dw . dw/ds_b formed as the reference forms it. */
static void regime_D(int dir, double gam, const double *P, const double *E, const double *dw, const double *K,
                     double *D) {
  double J[8][12];
  memset(J, 0, sizeof J);
  double M[8][3];
  memset(M, 0, sizeof M);
  int a = 1 + dir; /* the velocity / momentum index of the axis */
  if (dir == 0) {
    J[0][0] = P[1]; J[0][1] = P[0];
    J[1][4] = 1; J[1][5] = -P[5]; J[1][6] = P[6]; J[1][7] = P[7]; J[1][8] = -E[0]; J[1][9] = E[1]; J[1][10] = E[2];
    J[2][5] = -P[6]; J[2][6] = -P[5]; J[2][8] = -E[1]; J[2][9] = -E[0];
    J[3][5] = -P[7]; J[3][7] = -P[5]; J[3][8] = -E[2]; J[3][10] = -E[0];
    J[4][1] = gam * P[4] / (gam - 1); J[4][4] = gam * P[1] / (gam - 1);
    J[4][6] = -E[2]; J[4][7] = E[1]; J[4][9] = P[7]; J[4][10] = -P[6];
    J[6][10] = -1;
    J[7][9] = 1;
  } else if (dir == 1) {
    J[0][0] = P[2]; J[0][2] = P[0];
    J[1][5] = -P[6]; J[1][6] = -P[5]; J[1][8] = -E[1]; J[1][9] = -E[0];
    J[2][4] = 1; J[2][5] = P[5]; J[2][6] = -P[6]; J[2][7] = P[7]; J[2][8] = E[0]; J[2][9] = -E[1]; J[2][10] = E[2];
    J[3][6] = -P[7]; J[3][7] = -P[6]; J[3][9] = -E[2]; J[3][10] = -E[1];
    J[4][2] = gam * P[4] / (gam - 1); J[4][4] = gam * P[2] / (gam - 1);
    J[4][5] = E[2]; J[4][7] = -E[0]; J[4][8] = -P[7]; J[4][10] = P[5];
    J[5][10] = 1;
    J[7][8] = -1;
  } else {
    J[0][0] = P[3]; J[0][3] = P[0];
    J[1][5] = -P[7]; J[1][7] = -P[5]; J[1][8] = -E[2]; J[1][10] = -E[0];
    J[2][6] = -P[7]; J[2][7] = -P[6]; J[2][9] = -E[2]; J[2][10] = -E[1];
    J[3][4] = 1; J[3][5] = P[5]; J[3][6] = P[6]; J[3][7] = -P[7]; J[3][8] = E[0]; J[3][9] = E[1]; J[3][10] = -E[2];
    J[4][3] = gam * P[4] / (gam - 1); J[4][4] = gam * P[3] / (gam - 1);
    J[4][5] = -E[1]; J[4][6] = E[0]; J[4][8] = P[6]; J[4][9] = -P[5];
    J[5][9] = -1;
    J[6][8] = 1;
  }
  for (int m = 0; m < 3; m++) {
    M[0][m] -= J[0][a] * DW(a, m);
    M[4][m] -= J[4][a] * DW(a, m);
  }
  /* the two induction rows with a single unit entry */
  if (dir == 0) {
    for (int m = 0; m < 3; m++) M[6][m] -= (-DW(10, m));
    for (int m = 0; m < 3; m++) M[7][m] -= DW(9, m);
  } else if (dir == 1) {
    for (int m = 0; m < 3; m++) M[5][m] -= DW(10, m);
    for (int m = 0; m < 3; m++) M[7][m] -= (-DW(8, m));
  } else {
    for (int m = 0; m < 3; m++) M[5][m] -= (-DW(9, m));
    for (int m = 0; m < 3; m++) M[6][m] -= DW(8, m);
  }
  for (int l = 1; l < 5; l++)
    for (int m = 0; m < 3; m++)
      for (int n = 5; n < 11; n++) M[l][m] -= J[l][n] * DW(n, m);
  for (int l = 0; l < 8; l++) {
    D[l] = 0.0;
    for (int m = 0; m < 3; m++) D[l] += M[l][m] * K[m];
  }
}
#undef DW

/* set_vars and set_K (REGIME.cc:141-224,293-370): E, q, alpha and K of every
 * padded cell from the prims array as given (ghosts included). */
static void regime_fields(const oc_config *c, const double *prims, double *E, double *q, double *alpha, double *K) {
  geom g = mkgeom(c);
  long ntot = g.ntot;
  int N[3] = {g.Nx, g.Ny, g.Nz};
  long st[3] = {(long)g.Ny * g.Nz, g.Nz, 1};
  double dd[3] = {g.dx, g.dy, g.dz};
  double sig = c->sigma;
  double *fb = (double *)calloc(3 * ntot, sizeof(double));
  double *fr = (double *)calloc(3 * ntot, sizeof(double));
  memset(q, 0, sizeof(double) * ntot);
  memset(K, 0, sizeof(double) * 3 * ntot);
  /* set_vars (REGIME.cc:141-224) */
  for (long n = 0; n < ntot; n++) {
    const double *P0 = prims + n;
#define PV(v) P0[(long)(v) * ntot]
    E[0 * ntot + n] = -(PV(2) * PV(7) - PV(3) * PV(6));
    E[1 * ntot + n] = -(PV(3) * PV(5) - PV(1) * PV(7));
    E[2 * ntot + n] = -(PV(1) * PV(6) - PV(2) * PV(5));
  }
  for (int i = 0; i < g.Nx; i++)
    for (int j = 0; j < g.Ny; j++)
      for (int k = 0; k < g.Nz; k++) {
        int pos[3] = {i, j, k};
        int in = 1;
        for (int d = 0; d < g.dims; d++) in &= pos[d] >= 2 && pos[d] < N[d] - 2;
        if (!in) continue;
        long n = ((long)i * g.Ny + j) * g.Nz + k;
        double acc = 0.0;
        for (int d = 0; d < g.dims; d++) {
          const double *Ed = E + d * ntot;
          double t = (-Ed[n + 2 * st[d]] + 8 * Ed[n + st[d]] - 8 * Ed[n - st[d]] + Ed[n - 2 * st[d]]) / (12 * dd[d]);
          acc = d == 0 ? t : acc + t;
        }
        q[n] = acc;
      }
  for (long n = 0; n < ntot; n++) {
    const double *P0 = prims + n;
    double qs = q[n] * q[n] + sig * sig;
    alpha[n] = 1 / (qs * (qs + sig * sig * (PV(5) * PV(5) + PV(6) * PV(6) + PV(7) * PV(7))));
  }
  /* set_K (REGIME.cc:293-370): the stiff E fluxes reconstructed per axis */
  for (int d = 0; d < g.dims; d++) {
    for (long n = 0; n < ntot; n++) {
      const double *P0 = prims + n;
      double b[3];
      if (d == 0) { b[0] = 0; b[1] = PV(7); b[2] = -PV(6); }
      else if (d == 1) { b[0] = -PV(7); b[1] = 0; b[2] = PV(5); }
      else { b[0] = PV(6); b[1] = -PV(5); b[2] = 0; }
      for (int v = 0; v < 3; v++) fb[v * ntot + n] = b[v];
    }
    flux_recon_n(c, E, fb, fr, d, 3);
    for (int v = 0; v < 3; v++)
      for (int i = 0; i < g.Nx; i++)
        for (int j = 0; j < g.Ny; j++)
          for (int k = 0; k < g.Nz; k++) {
            int pos[3] = {i, j, k};
            if (pos[d] < 1 || pos[d] >= N[d] - 1) continue;
            long n = ((long)i * g.Ny + j) * g.Nz + k;
            K[v * ntot + n] += fr[v * ntot + n + st[d]] / (2 * dd[d]) - fr[v * ntot + n - st[d]] / (2 * dd[d]);
          }
  }
#undef PV
  free(fb);
  free(fr);
}

void oc_regime_fields(const oc_config *c, const double *prims, double *E, double *q, double *alpha, double *dwdsb,
                      double *K) {
  geom g = mkgeom(c);
  long ntot = g.ntot;
  regime_fields(c, prims, E, q, alpha, K);
  for (long n = 0; n < ntot; n++) {
    double P[8], Ec[3], dw[36];
    for (int v = 0; v < 8; v++) P[v] = prims[(long)v * ntot + n];
    for (int v = 0; v < 3; v++) Ec[v] = E[v * ntot + n];
    regime_dwdsb(q[n], c->sigma, P, Ec, alpha[n], dw);
    for (int l = 0; l < 36; l++) dwdsb[l * ntot + n] = dw[l];
  }
}

void oc_regime_source(const oc_config *c, const double *cons, const double *prims, const double *aux, double *src) {
  (void)cons;
  (void)aux;
  geom g = mkgeom(c);
  long ntot = g.ntot;
  int N[3] = {g.Nx, g.Ny, g.Nz};
  long st[3] = {(long)g.Ny * g.Nz, g.Nz, 1};
  double dd[3] = {g.dx, g.dy, g.dz};
  double sig = c->sigma, gam = c->gamma;
  double *E = (double *)calloc(3 * ntot, sizeof(double));
  double *q = (double *)calloc(ntot, sizeof(double));
  double *alpha = (double *)calloc(ntot, sizeof(double));
  double *K = (double *)calloc(3 * ntot, sizeof(double));
  double *D = (double *)calloc(8 * ntot, sizeof(double));
  regime_fields(c, prims, E, q, alpha, K);
#define PV(v) P0[(long)(v) * ntot]
  /* sourceExtension (REGIME.cc:83-138): D per axis, then its centred
   * difference (= on the first axis, += on the others) */
  for (int d = 0; d < g.dims; d++) {
    for (long n = 0; n < ntot; n++) {
      const double *P0 = prims + n;
      double P[8], Ec[3], dw[36], Kc[3], Dc[8];
      for (int v = 0; v < 8; v++) P[v] = PV(v);
      for (int v = 0; v < 3; v++) {
        Ec[v] = E[v * ntot + n];
        Kc[v] = K[v * ntot + n];
      }
      regime_dwdsb(q[n], sig, P, Ec, alpha[n], dw);
      regime_D(d, gam, P, Ec, dw, Kc, Dc);
      for (int l = 0; l < 8; l++) D[l * ntot + n] = Dc[l];
    }
    for (int v = 0; v < 8; v++)
      for (int i = 0; i < g.Nx; i++)
        for (int j = 0; j < g.Ny; j++)
          for (int k = 0; k < g.Nz; k++) {
            int pos[3] = {i, j, k};
            if (pos[d] < 1 || pos[d] >= N[d] - 1) continue;
            long n = ((long)i * g.Ny + j) * g.Nz + k;
            double t = (D[v * ntot + n + st[d]] - D[v * ntot + n - st[d]]) / (2 * dd[d]);
            src[v * ntot + n] = d == 0 ? t : src[v * ntot + n] + t;
          }
  }
#undef PV
  free(E);
  free(q);
  free(alpha);
  free(K);
  free(D);
}

/* ------------------------------------------------------- RKPlus family */

typedef struct {
  double *cons, *prims, *aux;
} ostate;

static ostate state_new(const oc_config *c) {
  geom g = mkgeom(c);
  ostate s;
  s.cons = (double *)calloc(g.ntot * oc_ncons(c->model), sizeof(double));
  s.prims = (double *)calloc(g.ntot * oc_nprims(c->model), sizeof(double));
  s.aux = (double *)calloc(g.ntot * oc_naux(c->model), sizeof(double));
  return s;
}

static void state_free(ostate *s) {
  free(s->cons);
  free(s->prims);
  free(s->aux);
}

/* the guess arrays of a stage: prims and aux of another state, every cell */
static void copy_guess(const oc_config *c, ostate *dst, const double *prims, const double *aux) {
  geom g = mkgeom(c);
  memcpy(dst->prims, prims, sizeof(double) * g.ntot * oc_nprims(c->model));
  memcpy(dst->aux, aux, sizeof(double) * g.ntot * oc_naux(c->model));
}

/* RKPlus::rhs (RKPlus.cc:15-50) without a model extension: source - F over
 * the interior. */
static void rkplus_rhs(const oc_config *c, const double *cons, const double *prims, const double *aux,
                       double *rhs) {
  geom g = mkgeom(c);
  int nc = oc_ncons(c->model);
  double *flux = (double *)calloc(g.ntot * nc, sizeof(double));
  double *src = (double *)calloc(g.ntot * nc, sizeof(double));
  oc_fvs_F(c, cons, prims, aux, flux);
  oc_source(c, cons, prims, aux, src);
  add_extension_source(c, cons, prims, aux, src);
  for (int v = 0; v < nc; v++)
    for (int i = g.is; i < g.ie; i++)
      for (int j = g.js; j < g.je; j++)
        for (int k = g.ks; k < g.ke; k++) rhs[IX(v, i, j, k)] = src[IX(v, i, j, k)] - flux[IX(v, i, j, k)];
  free(flux);
  free(src);
}

/* out = sum_m w[m] x[m] over the interior, n terms (x may alias out). */
static void lincomb(const oc_config *c, double *out, int n, const double *const *x, const double *w) {
  geom g = mkgeom(c);
  int nc = oc_ncons(c->model);
  for (int v = 0; v < nc; v++)
    for (int i = g.is; i < g.ie; i++)
      for (int j = g.js; j < g.je; j++)
        for (int k = g.ks; k < g.ke; k++) {
          long q = IX(v, i, j, k);
          double acc = w[0] * x[0][q];
          for (int m = 1; m < n; m++) acc += w[m] * x[m][q];
          out[q] = acc;
        }
}

/* One step of RK2B, RK3, RK4 or RK4_10 (RKPlus.cc:56-726). Each stage forms
 * its state from earlier states and dt * rhs, then finalises it; the guess of
 * a stage's cons2prim is the previous stage's prims (the final stage's are
 * the step's own). RK4_10's second preparation (RKPlus.cc:612-640) resets the
 * stage prims and aux to the step's own while replacing the cons, so the next
 * rhs reads those prims with the new cons, as the reference does; with quirks
 * off the new stage state is finalised (cons2prim from those prims) first. */
long oc_rkplus_step(const oc_config *c, double *cons, double *prims, double *aux, double dt, int method,
                    oc_stats *st) {
  geom g = mkgeom(c);
  int nc = oc_ncons(c->model);
  size_t cb = sizeof(double) * g.ntot * nc;
  long fails = 0;
  double *r1 = (double *)calloc(g.ntot * nc, sizeof(double));
  double *r2 = (double *)calloc(g.ntot * nc, sizeof(double));
  ostate u1 = state_new(c), u2 = state_new(c), u3 = state_new(c), u4 = state_new(c);
  if (method == OC_RK2B) {
    /* RK2B::stage1/stage2/step (RKPlus.cc:83-136) */
    copy_guess(c, &u1, prims, aux);
    rkplus_rhs(c, cons, prims, aux, r1);
    { const double *x[2] = {cons, r1}; double w[2] = {1.0, dt}; lincomb(c, u1.cons, 2, x, w); }
    fails += finalise(c, u1.cons, u1.prims, u1.aux, st);
    rkplus_rhs(c, u1.cons, u1.prims, u1.aux, r2);
    for (int v = 0; v < nc; v++)
      for (int i = g.is; i < g.ie; i++)
        for (int j = g.js; j < g.je; j++)
          for (int k = g.ks; k < g.ke; k++) {
            long q = IX(v, i, j, k);
            cons[q] = 0.5 * (cons[q] + u1.cons[q] + dt * r2[q]);
          }
    fails += finalise(c, cons, prims, aux, st);
  } else if (method == OC_RK3) {
    /* RK3::stage1/2/3/step (RKPlus.cc:177-274) */
    copy_guess(c, &u1, prims, aux);
    rkplus_rhs(c, cons, prims, aux, r1);
    { const double *x[2] = {cons, r1}; double w[2] = {1.0, dt}; lincomb(c, u1.cons, 2, x, w); }
    fails += finalise(c, u1.cons, u1.prims, u1.aux, st);
    copy_guess(c, &u2, u1.prims, u1.aux);
    rkplus_rhs(c, u1.cons, u1.prims, u1.aux, r1);
    { const double *x[3] = {cons, u1.cons, r1}; double w[3] = {0.75, 0.25, 0.25 * dt}; lincomb(c, u2.cons, 3, x, w); }
    fails += finalise(c, u2.cons, u2.prims, u2.aux, st);
    rkplus_rhs(c, u2.cons, u2.prims, u2.aux, r1);
    { const double *x[3] = {cons, u2.cons, r1}; double w[3] = {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0 * dt};
      lincomb(c, cons, 3, x, w); }
    fails += finalise(c, cons, prims, aux, st);
  } else if (method == OC_RK4) {
    /* RK4::stage1..5/step (RKPlus.cc:330-470): the five-stage fourth-order
     * strong-stability-preserving scheme with the reference's coefficients */
    copy_guess(c, &u1, prims, aux);
    rkplus_rhs(c, cons, prims, aux, r1);
    { const double *x[2] = {cons, r1}; double w[2] = {1.0, 0.391752226571890 * dt}; lincomb(c, u1.cons, 2, x, w); }
    fails += finalise(c, u1.cons, u1.prims, u1.aux, st);
    copy_guess(c, &u2, u1.prims, u1.aux);
    rkplus_rhs(c, u1.cons, u1.prims, u1.aux, r1);
    { const double *x[3] = {cons, u1.cons, r1};
      double w[3] = {0.444370493651235, 0.555629506348765, 0.368410593050371 * dt}; lincomb(c, u2.cons, 3, x, w); }
    fails += finalise(c, u2.cons, u2.prims, u2.aux, st);
    copy_guess(c, &u3, u2.prims, u2.aux);
    rkplus_rhs(c, u2.cons, u2.prims, u2.aux, r1);
    { const double *x[3] = {cons, u2.cons, r1};
      double w[3] = {0.620101851488403, 0.379898148511597, 0.251891774271694 * dt}; lincomb(c, u3.cons, 3, x, w); }
    fails += finalise(c, u3.cons, u3.prims, u3.aux, st);
    copy_guess(c, &u4, u3.prims, u3.aux);
    rkplus_rhs(c, u3.cons, u3.prims, u3.aux, r1); /* rhs4, kept for stage 5 */
    { const double *x[3] = {cons, u3.cons, r1};
      double w[3] = {0.178079954393132, 0.821920045606868, 0.544974750228521 * dt}; lincomb(c, u4.cons, 3, x, w); }
    fails += finalise(c, u4.cons, u4.prims, u4.aux, st);
    rkplus_rhs(c, u4.cons, u4.prims, u4.aux, r2);
    { const double *x[5] = {u2.cons, u3.cons, u4.cons, r1, r2};
      double w[5] = {0.517231671970585, 0.096059710526147, 0.386708617503269, 0.063692468666290 * dt,
                     0.226007483236906 * dt};
      lincomb(c, cons, 5, x, w); }
    fails += finalise(c, cons, prims, aux, st);
  } else {
    /* RK4_10::prepare1/prepare2/stageRepeat/stageFinal/step (RKPlus.cc:586-726):
     * ten stages of dt/6, the fourth-order ten-stage strong-stability-
     * preserving scheme */
    memcpy(u1.cons, cons, cb);
    memcpy(u2.cons, cons, cb);
    copy_guess(c, &u1, prims, aux);
    for (int s = 0; s < 9; s++) {
      if (s == 5) {
        copy_guess(c, &u1, prims, aux);
        for (long n = 0; n < g.ntot * nc; n++) {
          u2.cons[n] = 1.0 / 25.0 * u2.cons[n] + 9.0 / 25.0 * u1.cons[n];
          u1.cons[n] = 15.0 * u2.cons[n] - 5.0 * u1.cons[n];
        }
        /* without the quirk the new stage state gets its own primitives */
        if (!c->quirks) fails += finalise(c, u1.cons, u1.prims, u1.aux, st);
      }
      rkplus_rhs(c, u1.cons, u1.prims, u1.aux, r1);
      { const double *x[2] = {u1.cons, r1}; double w[2] = {1.0, 1.0 / 6.0 * dt}; lincomb(c, u1.cons, 2, x, w); }
      fails += finalise(c, u1.cons, u1.prims, u1.aux, st);
    }
    rkplus_rhs(c, u1.cons, u1.prims, u1.aux, r1);
    { const double *x[3] = {u2.cons, u1.cons, r1}; double w[3] = {1.0, 3.0 / 5.0, 1.0 / 10.0 * dt};
      lincomb(c, cons, 3, x, w); }
    fails += finalise(c, cons, prims, aux, st);
  }
  state_free(&u1);
  state_free(&u2);
  state_free(&u3);
  state_free(&u4);
  free(r1);
  free(r2);
  return fails;
}

/* ----------------------------------------------------------- SSP2 IMEX */

/* IMEX SSP2(222) constants as the reference's SSP2 object ends up holding them
 * (IMEX2Args.h:24-26: the default constructor's values survive the assignment
 * in SSP2::SSP2, SSP2.cc:16, because operator= does not copy them). */
#define SSP2_GAM 0.2928932188134525
#define SSP2_OM2GAM 0.4142135623730950

/* Stage equation of IMEX2Residual1/2 (SSP2.cc:194-264) reduced to its stiff
 * part. Only E, psi and phi have sources (srrmhd.cc:194-207), so every other
 * component of the root equals the explicit part c; psi and phi are linear:
 *   x_psi (1 + dtg/cp^2) = c_psi + dtg x_qch,   x_phi (1 + dtg/cp^2) = c_phi,
 * and E solves E + dtg J(x) = c_E, where J (srrmhd.cc:303-311) depends on x
 * through the primitive recovery. The recovery's own equation
 * (srrmhd.cc:512-537) is solved jointly, in y = (Ex, Ey, Ez, Z). */
static int imex_residual(const double *c, double dtg, double gam, double sigma, const double *y, double *F) {
  const double Ex = y[0], Ey = y[1], Ez = y[2], Z = y[3];
  const double Bx = c[5], By = c[6], Bz = c[7];
  const double Sbx = c[1] - (Ey * Bz - Ez * By);
  const double Sby = c[2] - (Ez * Bx - Ex * Bz);
  const double Sbz = c[3] - (Ex * By - Ey * Bx);
  const double tauBar = c[4] - 0.5 * (Bx * Bx + By * By + Bz * Bz + Ex * Ex + Ey * Ey + Ez * Ez);
  if (!(Z > 0)) return -1;
  const double vx = Sbx / Z, vy = Sby / Z, vz = Sbz / Z;
  const double vsq = vx * vx + vy * vy + vz * vz;
  if (!(vsq < 1.0)) return -1;
  const double W = 1.0 / sqrt(1.0 - vsq);
  const double D = c[0];
  F[3] = (1 - (gam - 1) / (W * W * gam)) * Z + ((gam - 1) / (W * gam) - 1) * D - tauBar;
  const double vE = vx * Ex + vy * Ey + vz * Ez;
  const double q = c[13];
  const double Jx = q * vx + W * sigma * (Ex + (vy * Bz - vz * By) - vE * vx);
  const double Jy = q * vy + W * sigma * (Ey + (vz * Bx - vx * Bz) - vE * vy);
  const double Jz = q * vz + W * sigma * (Ez + (vx * By - vy * Bx) - vE * vz);
  F[0] = Ex - c[8] + dtg * Jx;
  F[1] = Ey - c[9] + dtg * Jy;
  F[2] = Ez - c[10] + dtg * Jz;
  return 0;
}

/* Newton on y = (E, Z) with a forward-difference Jacobian (steps 1e-7 of
 * max(|y_j|, 1e-3)) and Gaussian elimination with partial pivoting; steps
 * that leave the physical region are halved. Converged when the largest step
 * component is <= 1e-14 max(1, |y|). Returns iterations or -1. */
static int imex_solve(const double *c, double dtg, double gam, double sigma, double *y) {
  double F[4], Fp[4], Jm[4][5];
  for (int it = 0; it < 60; it++) {
    if (imex_residual(c, dtg, gam, sigma, y, F)) return -1;
    for (int j = 0; j < 4; j++) {
      double yj = y[j], h = 1e-7 * fmax(fabs(yj), 1e-3);
      y[j] = yj + h;
      if (imex_residual(c, dtg, gam, sigma, y, Fp)) {
        h = -h;
        y[j] = yj + h;
        if (imex_residual(c, dtg, gam, sigma, y, Fp)) return -1;
      }
      y[j] = yj;
      for (int i = 0; i < 4; i++) Jm[i][j] = (Fp[i] - F[i]) / h;
    }
    for (int i = 0; i < 4; i++) Jm[i][4] = -F[i];
    for (int k = 0; k < 4; k++) {
      int piv = k;
      for (int i = k + 1; i < 4; i++)
        if (fabs(Jm[i][k]) > fabs(Jm[piv][k])) piv = i;
      if (piv != k)
        for (int j = 0; j < 5; j++) {
          double t = Jm[k][j];
          Jm[k][j] = Jm[piv][j];
          Jm[piv][j] = t;
        }
      if (Jm[k][k] == 0.0) return -1;
      for (int i = k + 1; i < 4; i++) {
        double m = Jm[i][k] / Jm[k][k];
        for (int j = k; j < 5; j++) Jm[i][j] -= m * Jm[k][j];
      }
    }
    double d[4];
    for (int k = 3; k >= 0; k--) {
      double sum = Jm[k][4];
      for (int j = k + 1; j < 4; j++) sum -= Jm[k][j] * d[j];
      d[k] = sum / Jm[k][k];
    }
    double ymax = 1.0, dmax = 0.0;
    for (int j = 0; j < 4; j++) {
      ymax = fmax(ymax, fabs(y[j]));
      dmax = fmax(dmax, fabs(d[j]));
    }
    double lam = 1.0, yn[4];
    for (int tries = 0; tries < 40; tries++) {
      for (int j = 0; j < 4; j++) yn[j] = y[j] + lam * d[j];
      if (imex_residual(c, dtg, gam, sigma, yn, Fp) == 0) break;
      lam *= 0.5;
    }
    for (int j = 0; j < 4; j++) y[j] = yn[j];
    if (dmax <= 1e-14 * ymax) return it + 1;
  }
  return -1;
}

/* One implicit stage for one cell: c holds the explicit part (cons for stage
 * 1, cons - dt (flux1 - om2gam source1) for stage 2), Zg the starting guess
 * of Z; x receives the stage value. */
static int imex_stage_cell(const double *c, double dt, double gam, double sigma, double cp, double Zg, double *x) {
  const double dtg = dt * SSP2_GAM;
  double y[4] = {c[8], c[9], c[10], Zg};
  int its = imex_solve(c, dtg, gam, sigma, y);
  for (int v = 0; v < 14; v++) x[v] = c[v];
  x[8] = y[0];
  x[9] = y[1];
  x[10] = y[2];
  const double k = 1.0 + dtg / (cp * cp);
  x[11] = (c[11] + dtg * x[13]) / k;
  x[12] = c[12] / k;
  return its;
}

/* SSP2::step (SSP2.cc:53-173) with the SRRMHD model. */
long oc_ssp2_step(const oc_config *c, double *cons, double *prims, double *aux, double dt, oc_stats *st) {
  geom g = mkgeom(c);
  if (c->model != OC_SRRMHD) return -1;
  const int nc = 14;
  const long n = g.ntot;
  double *U1 = (double *)calloc(n * nc, sizeof(double));
  double *U2 = (double *)calloc(n * nc, sizeof(double));
  double *s1 = (double *)calloc(n * nc, sizeof(double));
  double *f1 = (double *)calloc(n * nc, sizeof(double));
  double *s2 = (double *)calloc(n * nc, sizeof(double));
  double *f2 = (double *)calloc(n * nc, sizeof(double));
  long fails = 0;
  double cc[14], x[14];
  /* stage one: x - cons - dt gam psi(x) = 0, guess from the cell's aux Z */
  for (int i = g.is; i < g.ie; i++)
    for (int j = g.js; j < g.je; j++)
      for (int k = g.ks; k < g.ke; k++) {
        for (int v = 0; v < nc; v++) cc[v] = cons[IX(v, i, j, k)];
        if (imex_stage_cell(cc, dt, c->gamma, c->sigma, c->cp, aux[IX(10, i, j, k)], x) < 0) fails++;
        for (int v = 0; v < nc; v++) U1[IX(v, i, j, k)] = x[v];
      }
  fails += finalise(c, U1, prims, aux, st);
  oc_source(c, U1, prims, aux, s1);
  oc_fvs_F(c, U1, prims, aux, f1);
  /* stage two: x - cons + dt (flux1 - om2gam source1 - gam psi(x)) = 0 */
  for (int i = g.is; i < g.ie; i++)
    for (int j = g.js; j < g.je; j++)
      for (int k = g.ks; k < g.ke; k++) {
        for (int v = 0; v < nc; v++) {
          long q = IX(v, i, j, k);
          cc[v] = cons[q] - dt * (f1[q] - SSP2_OM2GAM * s1[q]);
        }
        if (imex_stage_cell(cc, dt, c->gamma, c->sigma, c->cp, aux[IX(10, i, j, k)], x) < 0) fails++;
        for (int v = 0; v < nc; v++) U2[IX(v, i, j, k)] = x[v];
      }
  fails += finalise(c, U2, prims, aux, st);
  oc_source(c, U2, prims, aux, s2);
  oc_fvs_F(c, U2, prims, aux, f2);
  for (int v = 0; v < nc; v++)
    for (int i = g.is; i < g.ie; i++)
      for (int j = g.js; j < g.je; j++)
        for (int k = g.ks; k < g.ke; k++) {
          long q = IX(v, i, j, k);
          cons[q] += -0.5 * dt * (f1[q] + f2[q] - s1[q] - s2[q]);
        }
  fails += finalise(c, cons, prims, aux, st);
  free(U1);
  free(U2);
  free(s1);
  free(f1);
  free(s2);
  free(f2);
  return fails;
}
