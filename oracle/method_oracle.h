/*
 * CPU oracle for the METHOD time-step hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A serial C restatement of the reference's algorithm (AlexJamesWright/METHOD,
 * Project/CPU/Src), used by tests/ as the checker and by bench.py as the CPU
 * baseline. Nothing under method_amd/ or include/ may link, load or call it:
 * the product computes on the GPU and fails loudly without its HIP library.
 *
 * Layout is the reference's: [var][i][j][k] over the ghost-padded grid,
 * k fastest (simData.h:10), Nx = nx + 2 Ng, 2-D when nz == 0 (Nz = 1).
 */
#ifndef METHOD_ORACLE_H
#define METHOD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OC_SRMHD = 0, OC_SRRMHD = 1 };

typedef struct {
  int nx, ny, nz, Ng;
  double xmin, xmax, ymin, ymax, zmin, zmax;
  double gamma, sigma, cp, cfl;
  int model;        /* OC_SRMHD | OC_SRRMHD */
  int weno_order;   /* 3 | 5 */
  int periodic[3];  /* per axis: 1 periodic, 0 outflow */
  int quirks;       /* 1: reproduce the reference's flux-index quirks */
  int regime;       /* 1: the REGIME model extension's source (SRMHD) */
} oc_config;

typedef struct {
  long cells;        /* C2P cells solved */
  long iterations;   /* Newton (or secant) steps summed over cells */
  long failures;     /* cells that did not converge */
  long max_iterations;
} oc_stats;

int oc_ncons(int model);
int oc_nprims(int model);
int oc_naux(int model);

/* Boundary conditions (boundaryConds.cc:5-145,449-730) on one array. */
void oc_apply_bcs(const oc_config *c, double *a, int nvar);

/* Model::primsToAll over every cell (srmhd.cc:573-673, srrmhd.cc:358-510). */
void oc_prims_to_all(const oc_config *c, double *cons, double *prims, double *aux);

/* Simulation::set: primsToAll then Bcs::apply (simulation.cc:94-96). */
void oc_set(const oc_config *c, double *cons, double *prims, double *aux);

/* Model::fluxVector (srmhd.cc:72-206, srrmhd.cc:65-192). */
void oc_flux_vector(const oc_config *c, const double *cons, const double *prims, const double *aux,
                    double *f, int dir);

/* Scalar WENO kernels (wenoUpwinds.cc:5-84). */
double oc_weno3_upwind(double v0, double v1, double v2);
double oc_weno5_upwind(double v0, double v1, double v2, double v3, double v4);

/* FVS::fluxReconstruction and FVS::F (fluxVectorSplitting.cc:4-131). */
void oc_flux_reconstruction(const oc_config *c, const double *cons, const double *f, double *frecon,
                            int dir);
void oc_fvs_F(const oc_config *c, const double *cons, const double *prims, const double *aux,
              double *fnet);

/* Model::sourceTerm (srmhd.cc:233-250, srrmhd.cc:209-254). */
void oc_source(const oc_config *c, const double *cons, const double *prims, const double *aux,
               double *source);

/* Model::getPrimitiveVars over interior cells; prims/aux hold the guess on
 * entry (srmhd.cc:390-562, srrmhd.cc:314-356). Returns failures. */
long oc_c2p(const oc_config *c, const double *cons, double *prims, double *aux, oc_stats *st);

/* Single-cell SRMHD C2P root: Newton on the 2x2 system of srmhd.cc:269-295.
 * x[0] = v^2, x[1] = rho h W^2 hold the guess on entry and the root on exit.
 * Returns the number of Newton steps, or -1 if it did not converge. */
int oc_srmhd_newton(double D, double Sx, double Sy, double Sz, double tau, double Bx, double By,
                    double Bz, double gamma, double *x);

/* Simulation::updateTime's dt (simulation.cc:113-124). */
double oc_timestep(const oc_config *c, double t, int iters, double end_time);

/* Fast magnetosonic bound of one SRMHD state along an axis, and its maximum
 * over the interior per axis (no counterpart in the reference, whose dt uses
 * the light speed). */
double oc_srmhd_max_speed(double rho, double vd, double v2, double h, double c, double b2);
void oc_max_wavespeed(const oc_config *c, const double *prims, const double *aux, double out[3]);

/* One RK2 step (RK2.cc:31-100) or one SSP2 IMEX step (SSP2.cc:53-173). */
long oc_rk2_step(const oc_config *c, double *cons, double *prims, double *aux, double dt, oc_stats *st);
long oc_ssp2_step(const oc_config *c, double *cons, double *prims, double *aux, double dt, oc_stats *st);
/* One RKSplit step (rkSplit.cc:31-53): RK2, split source, finalise. */
long oc_rksplit_step(const oc_config *c, double *cons, double *prims, double *aux, double dt, oc_stats *st);
/* REGIME::sourceExtension (REGIME.cc:83-138) of the given state, every cell
 * it writes (variables 0..7; variable 8 and the outer layer stay zero). */
void oc_regime_source(const oc_config *c, const double *cons, const double *prims, const double *aux, double *src);
/* REGIME's per-cell fields (set_vars, set_dwdsb, set_K: REGIME.cc:141-370) of
 * the prims array as given: E [3], q, alpha, dwdsb [12][3] and K [3], each
 * [..][cell] over the padded grid. */
void oc_regime_fields(const oc_config *c, const double *prims, double *E, double *q, double *alpha, double *dwdsb,
                      double *K);
/* RKPlus family (RKPlus.cc): method OC_RK2B, OC_RK3, OC_RK4 or OC_RK4_10. */
enum { OC_RK2B = 0, OC_RK3 = 1, OC_RK4 = 2, OC_RK4_10 = 3 };
long oc_rkplus_step(const oc_config *c, double *cons, double *prims, double *aux, double dt, int method,
                    oc_stats *st);

#ifdef __cplusplus
}
#endif
#endif
