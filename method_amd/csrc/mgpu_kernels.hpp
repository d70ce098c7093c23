// Kernel launch interface shared by the engine and the kernel translation unit.
#pragma once

#include <hip/hip_runtime.h>

#include "mgpu_device.hpp"

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

enum SweepMode { kSweepWrite = 0, kSweepAdd = 1, kSweepStage1 = 2, kSweepStage2 = 3, kSweepOnly = 4 };

struct SweepArgs {
  Geo geo;
  const double *U;   // cons swept
  const double *V;   // prims swept (vx, vy, vz, p read)
  double *fnet;      // net flux accumulator
  // stage modes
  const double *Un;  // stage 2: cons at t^n
  const double *Vg;  // prims giving the cons2prim starting guess
  double *Uout, *Vout, *Aout;
  double dt, gamma;
  C2PCounters *stats;
  int *fail_count;
  long long *fail_list;
  unsigned char *fail_flag;
};

void launch_sweep_srmhd(const SweepArgs &a, int order, int dir, int mode, int quirk, hipStream_t s);
void launch_prims_to_all_srmhd(const Geo &g, const double *V, double *U, double *A, double gam, hipStream_t s);
void launch_c2p_srmhd(const Geo &g, const double *U, double *V, double *A, double gam, C2PCounters *st,
                      int *fail_count, long long *fail_list, unsigned char *fail_flag, hipStream_t s);
void launch_c2p_retry_srmhd(const Geo &g, const double *U, double *V, double *A, double gam, const int *fail_count,
                            const long long *fail_list, unsigned char *fail_flag, long long last_cell,
                            C2PCounters *st, hipStream_t s);
void launch_zero_int(int *p, hipStream_t s);
void launch_compact(const Geo &g, const PadGeo &pg, int nvar, const double *src, double *dst, hipStream_t s);
void launch_expand(const Geo &g, const PadGeo &pg, int nvar, const double *src, double *dst, int ghosts_zero,
                   hipStream_t s);
void launch_source_srmhd(const Geo &g, const PadGeo &pg, const double *U, double *dst, double cp, hipStream_t s);
void launch_flux_vector_srmhd(const Geo &g, const PadGeo &pg, const double *U, const double *V, double *dst, int dir,
                              int quirk, hipStream_t s);

}  // namespace mgpu
