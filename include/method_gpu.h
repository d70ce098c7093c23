/*
 * method_gpu.h -- C ABI of the MI355X time-step engine for METHOD's hot path.
 *
 * The reference drives its time step through abstract C++ classes that take
 * raw host double* arrays in the layout [var][i][j][k] over the ghost-padded
 * grid (AlexJamesWright/METHOD, Project/CPU/Include/simData.h:10). This header
 * is the thin, C-only layer those classes call instead: plain pointers, sizes
 * and int status codes, no C++ types or exceptions across it. Host arrays
 * passed in or out always use the reference layout (padded, var-major, k
 * fastest); on the device the engine keeps its own interior-only layout and
 * fills ghost cells by index remapping inside the kernels.
 *
 * Status codes: 0 = ok, < 0 = error (text from mgpu_last_error()), > 0 = a
 * count of cells whose cons2prim solve did not converge (warning).
 *
 * Threading: one HIP stream per engine; an engine handle must not be used
 * from two host threads at once. Calls that return data to the host
 * synchronise the engine's stream; mgpu_step may still be running on return
 * until the next synchronising call.
 */
#ifndef METHOD_GPU_H
#define METHOD_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGPU_ABI_VERSION 1

enum { MGPU_SRMHD = 0, MGPU_SRRMHD = 1 };
/* MGPU_RKSPLIT: RK2 followed by the operator-split source update and
 * another cons2prim (RKSplit::step, rkSplit.cc:31-53). MGPU_RK2B, MGPU_RK3,
 * MGPU_RK4 and MGPU_RK4_10: the RKPlus family with rhs = source - F
 * (RK2B/RK3/RK4/RK4_10::step, RKPlus.cc:56-726); for RK4_10 quirks = 1 keeps
 * the reference's reset of the stage primitives at its second preparation
 * (RKPlus.cc:612-640), quirks = 0 gives the new stage state its own. */
enum { MGPU_RK2 = 0, MGPU_SSP2 = 1, MGPU_RKSPLIT = 2, MGPU_RK2B = 3, MGPU_RK3 = 4, MGPU_RK4 = 5, MGPU_RK4_10 = 6 };
enum { MGPU_BC_OUTFLOW = 0, MGPU_BC_PERIODIC = 1 };

/* Simulation parameters: the constructor arguments of the reference Data
 * object (Project/CPU/Include/simData.h:205-218) plus the choice of model,
 * WENO order, integrator and per-axis boundary, which the reference makes by
 * constructing SRMHD/SRRMHD, Weno3/Weno5, RK2/SSP2 and Periodic/Outflow/Flow
 * objects (Project/CPU/Src/main.cc:48-60). nz == 0 selects a 2-D grid
 * (simData.cc:74-79). */
typedef struct {
  int nx, ny, nz, Ng;
  double xmin, xmax, ymin, ymax, zmin, zmax;
  double gamma, sigma, cp, cfl;
  int model;       /* MGPU_SRMHD | MGPU_SRRMHD */
  int weno_order;  /* 3 | 5 */
  int integrator;  /* MGPU_RK2 | MGPU_SSP2 | MGPU_RKSPLIT | MGPU_RK2B | MGPU_RK3 | MGPU_RK4 | MGPU_RK4_10 */
  int bc[3];       /* per axis: MGPU_BC_PERIODIC | MGPU_BC_OUTFLOW */
  int quirks;      /* 1: reproduce the reference's quirks (flux indices, C2P retry, RK4_10 reset) */
  int device;      /* HIP device ordinal */
} mgpu_config;

enum { MGPU_GUESS_PLAIN = 0, MGPU_GUESS_LEARNED = 1, MGPU_GUESS_COLD = 2 };

/* Cons2prim work counters, summed over the call. */
typedef struct {
  long long cells;          /* cells solved */
  long long iterations;     /* Newton (SRMHD) or secant (SRRMHD) steps */
  long long failures;       /* cells that did not converge */
  long long max_iterations; /* largest per-cell count */
  long long fallbacks;      /* learned guess only: cells restarted from the previous primitives */
} mgpu_stats;

typedef struct mgpu_engine mgpu_engine;

/* Text of the last error on this thread (never NULL). */
const char *mgpu_last_error(void);

/* Allocates device state for one simulation. Replaces the allocations of
 * Simulation::Simulation (Project/CPU/Src/simulation.cc:22-30) and the
 * integrator's work arrays (RK2.cc:5-18, SSP2.cc:13-34, RKPlus.cc:56-586). */
int mgpu_create(const mgpu_config *cfg, mgpu_engine **out);
int mgpu_destroy(mgpu_engine *e);

/* Number of cons, prims and aux variables of the engine's model
 * (srmhd.cc:25-29, srrmhd.h via SRRMHD::SRRMHD). */
int mgpu_sizes(const mgpu_engine *e, int *ncons, int *nprims, int *naux);

/* Host -> device state in the reference layout (padded [var][i][j][k]);
 * only interior cells are read. Any pointer may be NULL to leave that array. */
int mgpu_upload_state(mgpu_engine *e, const double *cons, const double *prims, const double *aux);

/* Device -> host state in the reference layout with ghost cells filled by the
 * engine's boundary conditions, as Bcs::apply leaves them
 * (boundaryConds.cc:5-145,449-730). Any pointer may be NULL. After SRMHD RK2
 * and RKSplit steps and SRRMHD SSP2 steps, which keep each cell's cons2prim
 * root rather than its aux, the aux is formed from that root first
 * (srmhd.cc:509-557, srrmhd.cc:539-581). */
int mgpu_download_state(mgpu_engine *e, double *cons, double *prims, double *aux);

/* Model::primsToAll over the resident primitives, then the boundary
 * conditions: the body of Simulation::set (simulation.cc:94-96). */
int mgpu_prims_to_all(mgpu_engine *e);

/* One full time step of the configured integrator on the resident state with
 * step size dt: TimeIntegrator::step (RK2.cc:93-100, SSP2.cc:53-173,
 * rkSplit.cc:31-53, RKPlus.cc:127-726), with
 * FVS::F, Model::getPrimitiveVars and Bcs::apply inside. stats may be NULL. */
int mgpu_step(mgpu_engine *e, double dt, mgpu_stats *stats);

/* Simulation::updateTime (simulation.cc:104-133): computes dt from the CFL
 * rule with slow start and end-time clamp, steps, advances t and iters. */
int mgpu_update_time(mgpu_engine *e, double end_time, mgpu_stats *stats);
int mgpu_get_time(const mgpu_engine *e, double *t, double *dt, int *iters);
int mgpu_set_time(mgpu_engine *e, double t, int iters);

/* FVS::F (fluxVectorSplitting.cc:58-131) of the resident state; fnet is
 * written in the reference layout, interior cells only (ghosts zero). */
int mgpu_fnet(mgpu_engine *e, double *fnet);

/* Model::fluxVector (srmhd.cc:72-206, srrmhd.cc:65-192) of the resident
 * state in direction dir, written for every padded cell. */
int mgpu_flux_vector(mgpu_engine *e, int dir, double *f);

/* Model::sourceTerm (srmhd.cc:233-250, srrmhd.cc:209-254) of the resident
 * state, written for every padded cell. */
int mgpu_source(mgpu_engine *e, double *source);

/* FVS::fluxReconstruction (fluxVectorSplitting.cc:4-56, declared at
 * Project/CPU/Include/flux.h:72): the Lax-Friedrichs split
 * f+- = (f +- cons)/2 (alpha = 1) and the WENO face flux at i - 1/2 along
 * axis dir (0, 1, 2 = x, y, z), upwind on f+ plus downwind on f-
 * (weno.cc:23-118). cons, f and frecon are host arrays of the first `vars`
 * variables in the padded reference layout (vars < 0: all of the model's
 * conserved variables); every padded cell is an input, as after Bcs::apply.
 * frecon is written for those variables over every padded cell: the face
 * flux where the index along dir lies in [shift, N - shift), shift =
 * (order + 1)/2, and 0 elsewhere, as the reference's zero-initialised
 * temporaries leave it. With cons and f both NULL the input is the resident
 * state: its cons and its Model::fluxVector along dir (as mgpu_flux_vector
 * forms them, ghost cells holding their boundary images). */
int mgpu_flux_reconstruction(mgpu_engine *e, const double *cons, const double *f, int dir, int vars,
                             double *frecon);

/* Model::getPrimitiveVarsSingleCell (srmhd.cc:297-388, srrmhd.cc:256-312;
 * model.h:85) for n independent cells. Host arrays hold the cells one after
 * the other, each in the single-cell layout the reference passes
 * (cons[Ncons], prims[Nprims], aux[Naux]). prims and aux carry the starting
 * guess in (SRMHD: v^2 from prims, rho h / (1 - v^2) from prims[0] and
 * aux[0]; SRRMHD: aux[10], rho h W^2) and the recovered state out. Returns
 * the number of cells that did not converge (the reference throws for such
 * a cell); stats may be NULL. */
int mgpu_c2p_cells(mgpu_engine *e, long long n, const double *cons, double *prims, double *aux,
                   mgpu_stats *stats);

/* No counterpart in the reference. mgpu_c2p_cells for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): cons, prims and aux are DEVICE arrays
 * in the same single-cell layout, read and written on the engine's stream
 * (order earlier writes to them before the call). The call allocates
 * nothing: its per-block counters and their total live in a static device
 * array of each GPU, sized for at most 4,194,304 cells per call (larger n is
 * an error). Calls on one GPU share that array, so calls that overlap on one
 * GPU (from several threads or streams) race. With stats NULL the call only
 * launches and returns 0 (mgpu_sync before reading the results); otherwise
 * it waits, fills stats and returns the failure count as mgpu_c2p_cells. */
int mgpu_c2p_cells_noalloc(mgpu_engine *e, long long n, const double *cons, double *prims, double *aux,
                           mgpu_stats *stats);

/* Model::sourceTermSingleCell (srmhd.cc:212-223, srrmhd.cc:194-207;
 * model.h:55) for n independent cells in the single-cell layout. SRMHD reads
 * cons only; SRRMHD reads cons and the current J from aux[4..6]. prims is
 * not read and may be NULL. */
int mgpu_source_cells(mgpu_engine *e, long long n, const double *cons, const double *prims, const double *aux,
                      double *source);

/* Model extensions (ModelExtension, modelExtension.h:14-59): MGPU_EXT_REGIME
 * is REGIME (REGIME.h:81, REGIME.cc), the resistive extension of ideal SRMHD,
 * whose source RKSplit::setSource (rkSplit.cc:15-25) and RKPlus::rhs
 * (RKPlus.cc:26-38) add to the model's; the other integrators ignore it, as
 * the reference's do. Replaces constructing REGIME(data, fluxMethod) and
 * handing it to the integrator's constructor. SRMHD engines only. */
enum { MGPU_EXT_NONE = 0, MGPU_EXT_REGIME = 1 };
int mgpu_set_model_extension(mgpu_engine *e, int kind);

/* REGIME::sourceExtension (REGIME.cc:83-138) of the resident state in the
 * reference layout: variables 0..7 over the interior cells; phi and the
 * ghost cells are zero. Needs the extension set. */
int mgpu_source_extension(mgpu_engine *e, double *source);

/* REGIME's public per-cell fields (REGIME.h:83-101: E, q, alpha, dwdsb, K as
 * set by set_vars, set_dwdsb and set_K) of the resident state over the padded
 * grid, each [component][padded cell]: E 3, q 1, alpha 1, dwdsb 12 x 3, K 3.
 * Ghost cells hold their boundary images' primitives. Any pointer may be
 * NULL. Needs the extension set. */
int mgpu_regime_fields(mgpu_engine *e, double *E, double *q, double *alpha, double *dwdsb, double *K);

/* Model::getPrimitiveVars (srmhd.cc:390-562, srrmhd.cc:314-356) on the
 * resident cons with the resident prims as the starting guess. */
int mgpu_c2p(mgpu_engine *e, mgpu_stats *stats);

/* CFL max-wavespeed reduction (no counterpart in the reference, whose dt and
 * flux split take the light speed as the bound: simulation.cc:39-41,
 * fluxVectorSplitting.cc:10). out[d] = the largest |characteristic speed|
 * along axis d over interior cells of the resident state, 0 for axes the
 * grid does not have: the fast magnetosonic bound from the resident prims
 * and aux. SRMHD engines only; an SRRMHD engine gets an error, since its
 * light modes move at 1 in every cell. Synchronous. */
int mgpu_max_wavespeed(mgpu_engine *e, double out[3]);

/* Starting guess of mgpu_c2p's SRMHD Newton solve. The reference starts from
 * the previous primitives (srmhd.cc:438-442): MGPU_GUESS_PLAIN, the default.
 * MGPU_GUESS_LEARNED (no reference counterpart) evaluates an 8-64-64-2 GELU
 * network inside the kernel with BF16 weights on MFMA; params are its 4,866
 * FP32 parameters in the layout of mgpu_mlp_get_params and ymean_ystd its
 * output scales; a cell whose solve does not converge restarts from the
 * previous primitives (counted in mgpu_stats.fallbacks). MGPU_GUESS_COLD
 * starts from v^2 = 0, Z = tau + D, for comparison. */
int mgpu_set_c2p_guess(mgpu_engine *e, int mode, const float *params, const float *ymean_ystd);

/* Newton-step histogram of the SRMHD cons2prim (no reference counterpart;
 * the reference's hybrd keeps no per-cell counts). With on = 1 the 64 bins
 * are cleared and every later mgpu_c2p / mgpu_timed_c2p call adds each cell
 * of its main solve to bin min(steps, 63) (a learned-guess cell that falls
 * back counts both solves); cells that do not converge go to bin 63. on = 0
 * stops counting. mgpu_get_c2p_histogram copies the 64 bins out. */
int mgpu_set_c2p_histogram(mgpu_engine *e, int on);
int mgpu_get_c2p_histogram(mgpu_engine *e, long long *hist);

/* Times the cons2prim kernel alone: reps calls of mgpu_c2p's kernel on the
 * resident cons, each starting from the resident prims as they were on entry
 * (restored outside the timed window); the average kernel time from HIP
 * events on the engine stream goes to *ms_per_call and the last call's
 * counters to *stats. Leaves the last call's prims resident. */
int mgpu_timed_c2p(mgpu_engine *e, int reps, double *ms_per_call, mgpu_stats *stats);

/* ---- training the learned guess (no reference counterpart) ----
 * The network maps the conserved (D, Sx, Sy, Sz, tau, Bx, By, Bz) of a cell
 * to (log W, log(p / (tau + D))). Training runs on the GPU with FP32 master
 * weights, FP32 GEMM / GELU / MSE kernels and Adam. */
typedef struct mgpu_mlp mgpu_mlp;
int mgpu_mlp_create(int device, unsigned long long seed, mgpu_mlp **out);
int mgpu_mlp_destroy(mgpu_mlp *m);
/* n samples: cons8 row-major n x 8 (D, Sx, Sy, Sz, tau, Bx, By, Bz), w_p
 * row-major n x 2 (W, p). Sets the output normalisation from the data. */
int mgpu_mlp_set_data(mgpu_mlp *m, const double *cons8, const double *w_p, long long n);
/* steps Adam steps over consecutive batches of the data; *loss_out gets the
 * last batch's mean squared error (normalised outputs). */
int mgpu_mlp_train(mgpu_mlp *m, int steps, int batch, float lr, float *loss_out);
/* One forward/backward pass on given normalised inputs x (n x 8) and targets
 * t (n x 2): the 4,866 parameter gradients of the mean squared error. */
int mgpu_mlp_gradients(mgpu_mlp *m, const float *x, const float *t, int n, float *grads, float *loss);
int mgpu_mlp_get_params(mgpu_mlp *m, float *params, float *ymean_ystd);
int mgpu_mlp_set_params(mgpu_mlp *m, const float *params, const float *ymean_ystd);
/* Inference through the BF16 MFMA kernel: normalised outputs y (n x 2). */
int mgpu_mlp_predict(mgpu_mlp *m, const double *cons8, long long n, float *y);

/* ---- batched ensemble ----
 * Many independent 2-D ideal SRMHD simulations (RK2, FVS with WENO3/5) of
 * uneven grid size advanced together on one GPU. Each instance behaves as
 * its own Simulation (simulation.cc:8-133): Simulation::set through
 * mgpu_batch_set_prims, updateTime's dt rule (alpha = 1, slow start, clamp
 * to its end_time) per instance, RK2::step (RK2.cc:93-100), and it stops
 * stepping once t reaches end_time, as Simulation::evolve's loop does
 * (simulation.cc:136-160). One persistent kernel per RK stage advances all
 * instances; one step is captured in a hipGraph. Host arrays use the
 * reference layout of each instance (padded, [var][i][j], Ng ghosts). */
typedef struct mgpu_batch mgpu_batch;

typedef struct {
  int Ng;          /* ghost width of the host arrays */
  double gamma;
  int weno_order;  /* 3 or 5 */
  int bc[2];       /* MGPU_BC_* for x and y, shared by every instance */
  int quirks;      /* the reference's cons2prim retry quirk */
  int device;
} mgpu_batch_config;

typedef struct {
  int nx, ny;      /* interior cells */
  double xmin, xmax, ymin, ymax;
  double cfl;
  double end_time; /* the instance stops stepping when t reaches it */
} mgpu_instance;

/* Tile shape (tile_x cells along x times tile_y along y), schedule (0 static,
 * 1 dynamic work counters), workgroups per stage launch, and the measured
 * time per step of the autotuner trial. */
typedef struct {
  int tile_x, tile_y, schedule, blocks;
  double ms_per_step;
} mgpu_batch_tuning;

int mgpu_batch_create(const mgpu_batch_config *cfg, const mgpu_instance *inst, int ninst, mgpu_batch **out);
int mgpu_batch_destroy(mgpu_batch *b);
int mgpu_batch_sizes(const mgpu_batch *b, int *ninst, long long *cells, int *ntiles);
/* Simulation::set for instance i (simulation.cc:80-101): prims in the padded
 * reference layout [8][nx+2Ng][ny+2Ng]; resets the instance's t and iters. */
int mgpu_batch_set_prims(mgpu_batch *b, int i, const double *prims);
int mgpu_batch_upload_state(mgpu_batch *b, int i, const double *cons, const double *prims, const double *aux);
/* Any of cons / prims / aux may be NULL. Ghost cells hold the Bcs values. */
int mgpu_batch_download_state(mgpu_batch *b, int i, double *cons, double *prims, double *aux);
/* nsteps updateTime calls of every instance still short of its end_time.
 * With stats, the cons2prim counters of all instances and stages. */
int mgpu_batch_step(mgpu_batch *b, int nsteps, mgpu_stats *stats);
/* t, the dt of the last step taken and the iteration count of instance i. */
int mgpu_batch_get_time(mgpu_batch *b, int i, double *t, double *dt, int *iters);
/* 1 (default): steps replay the captured graph; 0: direct launches. */
int mgpu_batch_use_graph(mgpu_batch *b, int on);
int mgpu_batch_get_tuning(const mgpu_batch *b, mgpu_batch_tuning *t);
int mgpu_batch_set_tuning(mgpu_batch *b, const mgpu_batch_tuning *t);
/* Times every candidate tuning on trial_steps steps of this batch, keeps
 * the fastest and restores the state and clocks as they were on entry. all
 * (may be NULL) receives up to max_all candidates; *nall the count. */
int mgpu_batch_autotune(mgpu_batch *b, int trial_steps, mgpu_batch_tuning *best, mgpu_batch_tuning *all,
                        int max_all, int *nall);
/* nsteps steps timed with HIP events on the batch stream. */
int mgpu_batch_timed_steps(mgpu_batch *b, int nsteps, double *total_ms);
int mgpu_batch_sync(mgpu_batch *b);

/* No counterpart in the reference. Sweep plan of 3-D grids: 3 (default)
 * sums the x, y and z terms of FVS::F (fluxVectorSplitting.cc:79-112) in the
 * reference's order with an x launch and then one launch of the y and z
 * terms that finishes the stage; 2 runs the same as three launches, the
 * contiguous z launch finishing; 0 runs z, y, x with the x launch finishing;
 * 1 is plan 0 with the z and y launches fused. Plans 2 and 3 give identical
 * bits, and so do plans 0 and 1; the two pairs sum the axes in opposite
 * orders. Only launch count, HBM traffic and register pressure differ. On
 * 2-D grids plan 3 (default) runs the y and x terms as one launch that
 * finishes the stage, and plans 0-2 as a y launch and a finishing x launch,
 * with identical bits; 1-D grids always run their one x launch. */
int mgpu_set_sweep_plan(mgpu_engine *e, int plan);

/* Plan of the REGIME source (no reference counterpart; the reference forms
 * D = M K per axis in arrays, REGIME.cc:514-774, and differences them,
 * REGIME.cc:92-138): 1 (default) forms D in on-chip memory inside the launch
 * that differences it; 0 stores D, K and alpha in the work array and
 * differences them in a second launch. The source is identical bit for bit;
 * mgpu_regime_fields always uses the second form, which leaves its fields. */
int mgpu_set_regime_plan(mgpu_engine *e, int fused);

/* Block until the engine's queued work is done. */
int mgpu_sync(mgpu_engine *e);

/* Runs nsteps steps with the given dt values (length nsteps) and reports the
 * elapsed time of the whole sequence in milliseconds, measured with HIP
 * events recorded on the engine's stream; kernel_ms (may be NULL, length
 * kernel_slots) receives the average duration per launch of each kernel slot
 * of one step, from events recorded between the launches of the same pass,
 * and the slot count
 * is returned in *nslots. */
int mgpu_timed_steps(mgpu_engine *e, int nsteps, const double *dts, double *total_ms,
                     double *kernel_ms, int kernel_slots, int *nslots);

/* Name of kernel slot i in the order reported by mgpu_timed_steps. */
const char *mgpu_kernel_slot_name(const mgpu_engine *e, int slot);

#ifdef __cplusplus
}
#endif
#endif
