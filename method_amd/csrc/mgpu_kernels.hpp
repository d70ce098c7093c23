// Kernel launch interface shared by the engine and the kernel translation unit.
#pragma once

#include <hip/hip_runtime.h>

#include "mgpu_device.hpp"
#include "mgpu_mlp.hpp"
#include "mgpu_srrmhd.hpp"

namespace mgpu {

// Ghost-padded reference layout of a grid (simData.h:10): extents and the
// offset of interior cell (0,0,0).
struct PadGeo {
  int N[3];
  int off[3];
  long long ntot;
};

// Cons2prim counters, spread over kStatBuckets copies so blocks of one launch
// rarely hit the same address; the host sums the buckets.
struct C2PCounters {
  unsigned long long cells, iterations, failures, max_iterations, retry_failures;
};
constexpr int kStatBuckets = 128;

// What the last phase of a sweep does with a cell's net flux: write it, add
// it to the accumulator, finish an RK2 stage (RK2.cc:58-66 / 81-90, with
// cons2prim), finish SSP2 stage two (flux1 complete: store it and solve the
// stage-two implicit equation, SSP2.cc:112-152) or finish the SSP2 step
// (flux2 complete: final combination and cons2prim, SSP2.cc:156-170).
enum SweepMode { kSweepWrite = 0, kSweepAdd = 1, kSweepStage1 = 2, kSweepStage2 = 3, kSweepImex2 = 4,
                 kSweepImexFinal = 5 };

struct SweepArgs {
  Geo geo;
  const double *U;   // cons swept
  const double *V;   // prims swept (vx, vy, vz, p read)
  double *fnet;      // net flux accumulator
  // stage modes
  const double *Un;  // stage 2: cons at t^n
  const double *Vg;  // prims giving the cons2prim starting guess
  double *Uout, *Vout, *Aout;
  double *root;      // cons2prim root, kept when Aout is null: SRMHD (v^2, Z) [2][cell], SRRMHD Z [cell]
  int vout_n;        // prims written to Vout (0: all; 5: rho .. p, for a stage whose prims are
                     // read only by the next stage's sweeps, guess and retry)
  double dt, gamma, sigma, cp;
  double *Pf;        // SSP2: U^n - dt/2 flux1 + dt/2 (source1 + source2), the part of the final
                     // combination (SSP2.cc:156-166) known after stage two
  const double *S1;  // SSP2: source1 (components 8..12)
  double *Sout;      // SSP2 stage one: its source, source1 (components 8..12)
  C2PCounters *stats;
  int *fail_count;
  long long *fail_list;
  unsigned char *fail_flag;
  const double *Fin;  // stored flux vectors [dir][var][cell] (model ids 2, 3)
};

// model: 0 SRMHD, 1 SRRMHD, or 2 / 3 for the same models with the flux vector
// of every cell read from a.Fin instead of formed from (U, V) (write and add
// modes only; the per-cell split f+- still uses U).
void launch_sweep(const SweepArgs &a, int model, int order, int dir, int mode, int quirk, hipStream_t s);
// Flux vectors of every interior cell and active direction into
// F[dir][var][cell], formed as Model::fluxVector forms them from the stored
// cons, prims and aux (srmhd.cc:72-206, srrmhd.cc:65-192) even when those
// arrays do not belong to one state.
void launch_flux_stored(const Geo &g, int model, const double *U, const double *V, const double *A, double *F,
                        int quirk, hipStream_t s);
// Fused z and y sweeps of a 3-D grid: writes their summed net flux to a.fnet.
void launch_sweep_yz(const SweepArgs &a, int model, int order, int quirk, hipStream_t s);
// The y and z sweeps of a 3-D grid in one launch after the x sweep wrote
// a.fnet; mode is the z sweep's (kSweepAdd leaves the net flux in a.fnet).
void launch_sweep_yzf(const SweepArgs &a, int model, int order, int mode, int quirk, hipStream_t s);
// The y and x sweeps of a 2-D grid in one launch; mode is the x sweep's
// (kSweepAdd writes the net flux to a.fnet).
void launch_sweep_xyf(const SweepArgs &a, int model, int order, int mode, hipStream_t s);
void launch_imex1_srrmhd(const SweepArgs &a, hipStream_t s);
void launch_prims_to_all_srrmhd(const Geo &g, const double *V, double *U, double *A, double gam, double sigma,
                                hipStream_t s);
void launch_c2p_srrmhd(const Geo &g, const double *U, double *V, double *A, double gam, double sigma,
                       C2PCounters *st, hipStream_t s);
void launch_source_srrmhd(const Geo &g, const PadGeo &pg, const double *U, const double *V, double *dst,
                          double gam, double sigma, double cp, hipStream_t s);
void launch_flux_vector_srrmhd(const Geo &g, const PadGeo &pg, const double *U, const double *V, double *dst,
                               int dir, int quirk, double gam, double sigma, hipStream_t s);
void launch_prims_to_all_srmhd(const Geo &g, const double *V, double *U, double *A, double gam, hipStream_t s);
// SRMHD aux from the resident cons and the cons2prim root kept by a stage,
// through the same recovery formulas the stage would have applied.
void launch_aux_srmhd(const Geo &g, const double *U, const double *root, double *A, double gam, hipStream_t s);
// SRRMHD aux from the cons and the kept root Z [cell].
void launch_aux_srrmhd(const Geo &g, const double *U, const double *root, double *A, double gam, double sigma,
                       hipStream_t s);
// hist: null, or kHistBins Newton-step counters the cells are added to.
void launch_c2p_srmhd(const Geo &g, const double *U, double *V, double *A, double gam, C2PCounters *st,
                      int *fail_count, long long *fail_list, unsigned char *fail_flag, int cold,
                      unsigned long long *hist, hipStream_t s);
void launch_c2p_retry_srmhd(const Geo &g, const double *U, double *V, double *A, double *root, double gam,
                            const int *fail_count, const long long *fail_list, unsigned char *fail_flag,
                            long long last_cell, C2PCounters *st, hipStream_t s);
void launch_zero_int(int *p, hipStream_t s);
// RKSplit's source half-step and cons2prim over interior cells.
// A null: the cons2prim root (v^2, Z) goes to `root` in place of the aux.
void launch_split_srmhd(const Geo &g, double *U, double *V, double *A, double *root, double dt, double gam,
                        double cp, C2PCounters *st, int *fail_count, long long *fail_list, unsigned char *fail_flag,
                        const double *ext, hipStream_t s);
void launch_split_srrmhd(const Geo &g, double *U, double *V, double *A, double dt, double gam, double sigma,
                         double cp, C2PCounters *st, hipStream_t s);
// Per-axis maximum fast magnetosonic speed over interior cells into out[3]
// (bit patterns of non-negative doubles; out must be zero on entry).
void launch_max_wavespeed_srmhd(const Geo &g, const double *V, const double *A, unsigned long long *out,
                                hipStream_t s);
// One RKPlus stage over interior cells (RKPlus.cc:56-726): with
// rhs = source(Ub, Ab) - fnet (RKPlus::rhs, RKPlus.cc:15-50; no rhs when fnet
// is null), the stage state is
//   half_sum: 0.5 (X0 + X1 + dt rhs)                  (RK2B's corrector)
//   else:     sum_m w[m] Xm + (wR dt) R + (wr dt) rhs
// optionally keeping rhs in Rout, then cons2prim from the primitives Vg into
// (Uout, Vout, Aout); Aout may be null. Uout may alias X0 and Vout Vg.
struct RkStageArgs {
  Geo geo;
  const double *X[4];
  double w[4];
  int nx, half_sum;
  const double *Ub, *Ab, *fnet;
  double wr;
  const double *R;
  double wR;
  double *Rout;
  const double *ext;  // SRMHD model-extension source of the rhs state (variables 0..7), or null
  const double *Vg;
  double *Uout, *Vout, *Aout;
  double dt, gamma, sigma, cp;
  C2PCounters *stats;
  int *fail_count;
  long long *fail_list;
  unsigned char *fail_flag;
};
void launch_rkplus_stage(const RkStageArgs &a, int model, hipStream_t s);
// RK4_10::prepare2's cons update (RKPlus.cc:628-633) over interior cells:
// U2 = U2/25 + 9 U1/25, then U1 = 15 U2 - 5 U1.
void launch_rk410_prepare2(const Geo &g, int ncons, double *U1, double *U2, hipStream_t s);
// REGIME::sourceExtension (REGIME.cc:83-138) of the prims V into src
// (variables 0..7, interior cells), formed on the padded grid with the
// reference's loop ranges; work holds regime_work_doubles(pg) doubles; dx are
// the grid spacings.
size_t regime_work_doubles(const PadGeo &pg);
// Where REGIME's per-cell fields sit in the work array ([..][padded cell]):
// E [3], q, face fluxes [3 axes][3], D [3 axes][8], K [3], alpha.
struct RegimeFields {
  double *E, *q, *F, *D, *K, *alpha;
};
RegimeFields regime_fields_of(const PadGeo &pg, double *work);
// dw/ds_b [12*3][padded cell] from the fields of the last source evaluation.
void launch_regime_dwdsb(const Geo &g, const PadGeo &pg, const double *V, const double *work, double sigma,
                         double *out, hipStream_t s);
// REGIME's source of the primitives V into src (8 x interior cells). fused:
// D and its differences in one launch (nothing but E, q and the face fluxes
// in the work array); otherwise two launches that also leave D, K and alpha
// there. The two give the same source bit for bit.
void launch_regime_source(const Geo &g, const PadGeo &pg, const double dx[3], const double *V, double *work,
                          double *src, double gam, double sigma, int order, bool fused, hipStream_t s);
void launch_compact(const Geo &g, const PadGeo &pg, int nvar, const double *src, double *dst, hipStream_t s);
void launch_expand(const Geo &g, const PadGeo &pg, int nvar, const double *src, double *dst, int ghosts_zero,
                   hipStream_t s);
void launch_source_srmhd(const Geo &g, const PadGeo &pg, const double *U, double *dst, double cp, hipStream_t s);

// ---- reference-API operators outside the time step (mgpu_cells.hip)
// FVS::fluxReconstruction (fluxVectorSplitting.cc:4-56) of variables
// [0, nvars) along axis dir on padded arrays of stride pg.ntot.
void launch_flux_recon(const PadGeo &pg, int order, int dir, int nvars, const double *cons, const double *f,
                       double *frecon, hipStream_t s);
// Counter record of the cell-list cons2prim: one per block, then their sum.
struct CellCounters {
  unsigned long long cells, iterations, failures, max_iterations;
};
constexpr int kCellBlock = 256;  // threads (cells) per block of the cell-list kernels
long long cells_blocks(long long n);
// Model::getPrimitiveVarsSingleCell for n cells in the single-cell layout
// ([cell][var]); prims and aux carry the starting guess in and the result
// out. partial receives cells_blocks(n) records, total their sum.
void launch_c2p_cells(int model, long long n, const double *cons, double *prims, double *aux, double gam,
                      double sigma, CellCounters *partial, CellCounters *total, hipStream_t s);
// The allocation-free cell-list cons2prim keeps its per-block records and
// their total in a static device array sized for this many cells per call.
constexpr long long kNoAllocMaxCells = 1LL << 22;
// noalloc_scratch_setup looks up that array's address on `device` (which
// must be current) once per device and returns nonzero on failure;
// noalloc_scratch returns the address found, or null before set-up.
int noalloc_scratch_setup(int device);
CellCounters *noalloc_scratch(int device);
// Model::sourceTermSingleCell for n cells ([cell][var]).
void launch_source_cells(int model, long long n, const double *cons, const double *aux, double cp, double *src,
                         hipStream_t s);
void launch_flux_vector_srmhd(const Geo &g, const PadGeo &pg, const double *U, const double *V, double *dst, int dir,
                              int quirk, hipStream_t s);

}  // namespace mgpu
