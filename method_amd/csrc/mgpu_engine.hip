// C ABI of the time-step engine (include/method_gpu.h): device state,
// integrator sequencing, host <-> device layout conversion and timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>

#include "method_gpu.h"
#include "mgpu_kernels.hpp"

using namespace mgpu;

namespace {

thread_local std::string g_err = "";

int fail(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}

#define HIP_OK(expr)                                                                       \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) return fail("%s failed: %s", #expr, hipGetErrorString(_e));      \
  } while (0)

// Records a HIP event on the engine stream after every kernel of a step so
// each launch's duration can be read back.
struct Marks {
  std::vector<hipEvent_t> ev;
  std::vector<int> slot;
};

}  // namespace

namespace mgpu {
int engine_fail(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}
}  // namespace mgpu

struct mgpu_engine {
  mgpu_config cfg;
  Geo geo;
  PadGeo pad;
  int ncons, nprims, naux;
  hipStream_t stream = nullptr;
  double *U = nullptr, *V = nullptr, *A = nullptr, *U1 = nullptr, *V1 = nullptr, *fnet = nullptr;
  double *U2 = nullptr, *V2 = nullptr, *Pf = nullptr, *S1 = nullptr;  // SSP2 only
  // RKPlus stage states (aux only for the resistive model, whose source reads
  // the current from aux) and RK4's kept stage-four rhs
  double *rkU[4] = {}, *rkV[4] = {}, *rkA[4] = {}, *rkR = nullptr;
  double *rkF = nullptr;  // RK4_10: stored flux vectors [dir][var][cell]
  int extension = MGPU_EXT_NONE;         // model extension (REGIME) added by RKSplit / RKPlus
  int regime_fused = 1;                  // REGIME's D and source in one launch
  double *rg_work = nullptr, *rg_src = nullptr;  // its padded work arrays and interior source
  double *staging = nullptr;
  size_t staging_elems = 0;
  C2PCounters *dstats = nullptr;
  int *fail_count = nullptr;
  long long *fail_list = nullptr;
  unsigned char *fail_flag = nullptr;
  double t = 0.0, dt = 0.0, dx = 0.0, dy = 0.0, dz = 0.0;
  int iters = 0;
  std::vector<std::string> slot_names;
  Marks *marks = nullptr;
  int guess_mode = MGPU_GUESS_PLAIN;  // starting guess of mgpu_c2p
  MlpDev *mlp = nullptr;              // learned guess network (device)
  unsigned long long *dspeed = nullptr;  // max-wavespeed result, 3 entries
  int fused_yz = 0;                      // 3-D: fused z+y sweep launch (off: slower here)
  int xyz_order = 1;                     // 3-D: sweeps x, y, z with the last (contiguous) one finishing
  int fused_yzf = 1;                     // plan 3: 3-D x sweep, then y + z as one launch; 2-D y + x as one
  // SRMHD RK2 and RKSplit steps keep the cons2prim root (v^2, Z) of each cell
  // instead of writing its 13 aux values, and SRRMHD SSP2 steps the root Z
  // instead of 17: no kernel of those steps reads aux, and what does (state
  // download, wavespeed) forms it from the root first.
  double *root = nullptr;
  bool aux_stale = false;
  // cell-list operators (single-cell model calls): device copies of the
  // cells and their per-block counter records, grown on demand
  unsigned long long *hist = nullptr;  // Newton-step histogram of the standalone cons2prim, when on
  double *cells_buf = nullptr;
  size_t cells_elems = 0;
  CellCounters *cells_part = nullptr;
  long long cells_part_n = 0;
};

namespace {

int slot_of(mgpu_engine *e, const char *name) {
  for (size_t i = 0; i < e->slot_names.size(); ++i)
    if (e->slot_names[i] == name) return (int)i;
  e->slot_names.push_back(name);
  return (int)e->slot_names.size() - 1;
}

int mark(mgpu_engine *e, const char *name) {
  if (!e->marks) return 0;
  hipEvent_t ev;
  HIP_OK(hipEventCreate(&ev));
  HIP_OK(hipEventRecord(ev, e->stream));
  e->marks->ev.push_back(ev);
  e->marks->slot.push_back(slot_of(e, name));
  return 0;
}

int ensure_staging(mgpu_engine *e) {
  const size_t need = (size_t)(e->ncons + e->nprims + e->naux) * e->pad.ntot;
  if (e->staging_elems >= need) return 0;
  if (e->staging) HIP_OK(hipFree(e->staging));
  e->staging = nullptr;
  HIP_OK(hipMalloc(&e->staging, need * sizeof(double)));
  e->staging_elems = need;
  return 0;
}

int zero_stats(mgpu_engine *e) {
  HIP_OK(hipMemsetAsync(e->dstats, 0, sizeof(C2PCounters) * kStatBuckets, e->stream));
  return 0;
}

int read_stats(mgpu_engine *e, mgpu_stats *st) {
  if (!st) return 0;
  C2PCounters h[kStatBuckets];
  HIP_OK(hipMemcpyAsync(h, e->dstats, sizeof h, hipMemcpyDeviceToHost, e->stream));
  HIP_OK(hipStreamSynchronize(e->stream));
  long long cells = 0, iters = 0, fails = 0, maxit = 0, fallbacks = 0;
  for (int b = 0; b < kStatBuckets; ++b) {
    cells += (long long)h[b].cells;
    iters += (long long)h[b].iterations;
    fails += (long long)h[b].failures;
    fallbacks += (long long)h[b].retry_failures;
    maxit = std::max(maxit, (long long)h[b].max_iterations);
  }
  st->cells = cells;
  st->iterations = iters;
  st->failures = fails;
  st->max_iterations = maxit;
  st->fallbacks = fallbacks;
  return 0;
}

SweepArgs sweep_args(mgpu_engine *e, const double *U, const double *V, C2PCounters *st) {
  SweepArgs a;
  memset(&a, 0, sizeof a);
  a.geo = e->geo;
  a.U = U;
  a.V = V;
  a.fnet = e->fnet;
  a.gamma = e->cfg.gamma;
  a.sigma = e->cfg.sigma;
  a.cp = e->cfg.cp;
  a.stats = st;
  a.fail_count = e->fail_count;
  a.fail_list = e->fail_list;
  a.fail_flag = e->fail_flag;
  return a;
}

// The directional sweeps of one stage. 3-D grids sum the axes in the
// reference's order, x, y, z (fluxVectorSplitting.cc:79-112): by default an
// x sweep, then one launch of the y and z sweeps that applies `last_mode`
// (plan 3), or, under plan 2, three launches with the contiguous z sweep
// applying it, with the same bits. 1-D and 2-D grids, and 3-D grids under
// plan 0 or 1, sweep z, y, x with the x sweep applying it (plan 1 fuses the z
// and y sweeps into one launch with the same bits as plan 0); on 2-D grids
// plan 3 runs the y and x sweeps as one launch, with the same bits.
int run_sweeps(mgpu_engine *e, SweepArgs a, int last_mode) {
  const int dims = e->geo.dims;
  const int order = e->cfg.weno_order, quirk = e->cfg.quirks;
  if (dims == 3 && e->fused_yz) {
    launch_sweep_yz(a, e->cfg.model, order, quirk, e->stream);
    if (mark(e, "sweep_yz_write")) return -1;
    const int mode = (last_mode == kSweepWrite || last_mode == kSweepAdd) ? kSweepAdd : last_mode;
    launch_sweep(a, e->cfg.model, order, 0, mode, quirk, e->stream);
    static const char *xnames[6] = {"sweep_x_write", "sweep_x_add", "sweep_x_stage1_c2p", "sweep_x_stage2_c2p",
                                    "sweep_x_imex2", "sweep_x_final"};
    if (mark(e, xnames[mode])) return -1;
    HIP_OK(hipGetLastError());
    return 0;
  }
  if (dims == 1 && last_mode != kSweepWrite && last_mode != kSweepAdd) {
    // the x sweep is the only one: the earlier axes' net flux it adds to is zero
    HIP_OK(hipMemsetAsync(e->fnet, 0, sizeof(double) * e->ncons * e->geo.vstride, e->stream));
  }
  static const char *names[3][6] = {
      {"sweep_x_write", "sweep_x_add", "sweep_x_stage1_c2p", "sweep_x_stage2_c2p", "sweep_x_imex2", "sweep_x_final"},
      {"sweep_y_write", "sweep_y_add", "sweep_y_stage1_c2p", "sweep_y_stage2_c2p", "sweep_y_imex2", "sweep_y_final"},
      {"sweep_z_write", "sweep_z_add", "sweep_z_stage1_c2p", "sweep_z_stage2_c2p", "sweep_z_imex2",
       "sweep_z_final"}};
  if (dims == 3 && e->fused_yzf) {
    launch_sweep(a, e->cfg.model, order, 0, kSweepWrite, quirk, e->stream);
    if (mark(e, names[0][kSweepWrite])) return -1;
    const int mode = (last_mode == kSweepWrite || last_mode == kSweepAdd) ? kSweepAdd : last_mode;
    launch_sweep_yzf(a, e->cfg.model, order, mode, quirk, e->stream);
    static const char *yzf_names[6] = {"sweep_yz_add", "sweep_yz_add", "sweep_yz_stage1_c2p",
                                       "sweep_yz_stage2_c2p", "sweep_yz_imex2", "sweep_yz_final"};
    if (mark(e, yzf_names[mode])) return -1;
    HIP_OK(hipGetLastError());
    return 0;
  }
  if (dims == 2 && e->fused_yzf) {
    const int mode = (last_mode == kSweepWrite || last_mode == kSweepAdd) ? kSweepAdd : last_mode;
    launch_sweep_xyf(a, e->cfg.model, order, mode, e->stream);
    static const char *xyf_names[6] = {"sweep_yx_add", "sweep_yx_add", "sweep_yx_stage1_c2p",
                                       "sweep_yx_stage2_c2p", "sweep_yx_imex2", "sweep_yx_final"};
    if (mark(e, xyf_names[mode])) return -1;
    HIP_OK(hipGetLastError());
    return 0;
  }
  if (e->xyz_order && dims == 3) {
    // x, then y, then z: the reference's summation order (fluxVectorSplitting.cc:79-112);
    // the last axis, the contiguous one, applies last_mode
    for (int dir = 0; dir < dims; ++dir) {
      const bool last = dir == dims - 1;
      int mode = dir == 0 ? kSweepWrite : kSweepAdd;
      if (last && last_mode != kSweepWrite && last_mode != kSweepAdd) mode = last_mode;
      launch_sweep(a, e->cfg.model, order, dir, mode, quirk, e->stream);
      if (mark(e, names[dir][mode])) return -1;
    }
    HIP_OK(hipGetLastError());
    return 0;
  }
  int first = 1;
  for (int dir = dims - 1; dir >= 0; --dir) {
    int mode = dir == 0 ? last_mode : (first ? kSweepWrite : kSweepAdd);
    if (dir == 0 && (last_mode == kSweepWrite || last_mode == kSweepAdd)) mode = first ? kSweepWrite : kSweepAdd;
    launch_sweep(a, e->cfg.model, order, dir, mode, quirk, e->stream);
    if (mark(e, names[dir][mode])) return -1;
    first = 0;
  }
  HIP_OK(hipGetLastError());
  return 0;
}

int retry(mgpu_engine *e, const double *U, double *V, double *A, C2PCounters *st, double *root = nullptr) {
  if (e->cfg.model != MGPU_SRMHD) return 0;  // the resistive recovery has no retry (srrmhd.cc:574-579)
  const long long last = e->cfg.quirks ? e->geo.vstride - 1 : -1;
  launch_c2p_retry_srmhd(e->geo, U, V, A, root, e->cfg.gamma, e->fail_count, e->fail_list, e->fail_flag, last, st,
                         e->stream);
  HIP_OK(hipGetLastError());
  return mark(e, "c2p_retry");
}

// RK2::step (RK2.cc:93-100): predictor, finalise, corrector, finalise.
int rk2_step(mgpu_engine *e, double dt, C2PCounters *st) {
  SweepArgs a = sweep_args(e, e->U, e->V, st);
  a.dt = dt;
  a.Vg = e->V;
  a.Uout = e->U1;
  a.Vout = e->V1;
  a.Aout = nullptr;
  // V1 is read only by stage two's sweeps (SRMHD vx .. p, SRRMHD rho .. p), guess and retry (rho .. p)
  a.vout_n = 5;
  if (run_sweeps(e, a, kSweepStage1)) return -1;
  if (retry(e, e->U1, e->V1, nullptr, st)) return -1;

  SweepArgs b = sweep_args(e, e->U1, e->V1, st);
  b.dt = dt;
  b.Un = e->U;
  b.Vg = e->V;
  b.Uout = e->U;
  b.Vout = e->V;
  const bool lazy_aux = e->root != nullptr;
  b.Aout = lazy_aux ? nullptr : e->A;
  b.root = e->root;
  if (run_sweeps(e, b, kSweepStage2)) return -1;
  if (retry(e, e->U, e->V, lazy_aux ? nullptr : e->A, st, e->root)) return -1;
  e->aux_stale = lazy_aux;
  return 0;
}

// The aux of the resident state, formed now if the last step left it.
int fresh_aux(mgpu_engine *e) {
  if (!e->aux_stale) return 0;
  if (e->cfg.model == MGPU_SRMHD)
    launch_aux_srmhd(e->geo, e->U, e->root, e->A, e->cfg.gamma, e->stream);
  else
    launch_aux_srrmhd(e->geo, e->U, e->root, e->A, e->cfg.gamma, e->cfg.sigma, e->stream);
  HIP_OK(hipGetLastError());
  e->aux_stale = false;
  return 0;
}

// SSP2::step (SSP2.cc:53-173): implicit stage one, flux1 sweeps ending in the
// stage-two solve, flux2 sweeps ending in the final combination.
int ssp2_step(mgpu_engine *e, double dt, C2PCounters *st) {
  SweepArgs a = sweep_args(e, e->U, e->V, st);
  a.dt = dt;
  a.Un = e->U;
  a.Vg = e->V;
  a.Uout = e->U1;
  a.Vout = e->V1;
  a.vout_n = 5;  // V1 and V2 are read only as the next stage's flux prims and guess (rho .. p)
  a.Aout = nullptr;
  a.Sout = e->S1;  // stage one writes source1
  launch_imex1_srrmhd(a, e->stream);
  HIP_OK(hipGetLastError());
  if (mark(e, "imex_stage1_c2p")) return -1;

  SweepArgs b = sweep_args(e, e->U1, e->V1, st);
  b.dt = dt;
  b.Un = e->U;
  b.Vg = e->V1;
  b.Uout = e->U2;
  b.Vout = e->V2;
  b.vout_n = 5;
  b.Aout = nullptr;
  b.Pf = e->Pf;
  b.S1 = e->S1;
  if (run_sweeps(e, b, kSweepImex2)) return -1;

  SweepArgs c = sweep_args(e, e->U2, e->V2, st);
  c.dt = dt;
  c.Un = e->U;
  c.Vg = e->V2;
  c.Uout = e->U;
  c.Vout = e->V;
  // the step keeps each cell's root Z; the aux is formed from it when read
  c.Aout = nullptr;
  c.root = e->root;
  c.Pf = e->Pf;
  if (run_sweeps(e, c, kSweepImexFinal)) return -1;
  e->aux_stale = true;
  return 0;
}

// The model extension's source of the state with prims V into e->rg_src
// (REGIME::sourceExtension, REGIME.cc:83-138), or null without an extension.
const double *extension_source(mgpu_engine *e, const double *V, bool keep_fields = false) {
  if (e->extension != MGPU_EXT_REGIME) return nullptr;
  const double dx[3] = {e->dx, e->dy, e->dz};
  launch_regime_source(e->geo, e->pad, dx, V, e->rg_work, e->rg_src, e->cfg.gamma, e->cfg.sigma,
                       e->cfg.weno_order, e->regime_fused && !keep_fields, e->stream);
  mark(e, "regime_source");
  return e->rg_src;
}

// RKSplit::step (rkSplit.cc:31-53): RK2::step, then cons += dt source (plus
// the model extension's, rkSplit.cc:15-25) and finalise (cons2prim, with the
// SRMHD retry).
int rksplit_step(mgpu_engine *e, double dt, C2PCounters *st) {
  if (rk2_step(e, dt, st)) return -1;
  if (e->cfg.model == MGPU_SRMHD) {
    const double *ext = extension_source(e, e->V);
    HIP_OK(hipGetLastError());
    launch_split_srmhd(e->geo, e->U, e->V, e->root ? nullptr : e->A, e->root, dt, e->cfg.gamma, e->cfg.cp, st,
                       e->fail_count, e->fail_list, e->fail_flag, ext, e->stream);
  }
  else
    launch_split_srrmhd(e->geo, e->U, e->V, e->A, dt, e->cfg.gamma, e->cfg.sigma, e->cfg.cp, st, e->stream);
  HIP_OK(hipGetLastError());
  if (mark(e, "split_source_c2p")) return -1;
  return retry(e, e->U, e->V, e->root ? nullptr : e->A, st, e->root);
}

// ---------------------------------------------------------------- RKPlus

RkStageArgs rk_args(mgpu_engine *e, double dt, C2PCounters *st) {
  RkStageArgs a;
  memset(&a, 0, sizeof a);
  a.geo = e->geo;
  a.dt = dt;
  a.gamma = e->cfg.gamma;
  a.sigma = e->cfg.sigma;
  a.cp = e->cfg.cp;
  a.stats = st;
  a.fail_count = e->fail_count;
  a.fail_list = e->fail_list;
  a.fail_flag = e->fail_flag;
  return a;
}

// RKPlus::rhs of the state (Ub, Vb, Ab) (RKPlus.cc:15-50): its net flux into
// e->fnet by the three sweeps; the stage kernel adds the source.
int rk_rhs(mgpu_engine *e, RkStageArgs &a, const double *Ub, const double *Vb, const double *Ab) {
  SweepArgs s = sweep_args(e, Ub, Vb, nullptr);
  if (run_sweeps(e, s, kSweepAdd)) return -1;
  a.Ub = Ub;
  a.Ab = Ab;
  a.fnet = e->fnet;
  a.ext = extension_source(e, Vb);
  HIP_OK(hipGetLastError());
  return 0;
}

// The same from flux vectors formed from the stored arrays as the reference
// forms them, for a state whose cons and prims/aux disagree (RK4_10 after its
// second preparation).
int rk_rhs_stored(mgpu_engine *e, RkStageArgs &a, const double *Ub, const double *Vb, const double *Ab) {
  launch_flux_stored(e->geo, e->cfg.model, Ub, Vb, Ab, e->rkF, e->cfg.quirks, e->stream);
  HIP_OK(hipGetLastError());
  if (mark(e, "flux_stored")) return -1;
  SweepArgs s = sweep_args(e, Ub, Vb, nullptr);
  s.Fin = e->rkF;
  for (int dir = e->geo.dims - 1; dir >= 0; --dir) {
    launch_sweep(s, e->cfg.model + 2, e->cfg.weno_order, dir, dir == e->geo.dims - 1 ? kSweepWrite : kSweepAdd, 0,
                 e->stream);
    HIP_OK(hipGetLastError());
    if (mark(e, "sweep_stored")) return -1;
  }
  a.Ub = Ub;
  a.Ab = Ab;
  a.fnet = e->fnet;
  a.ext = extension_source(e, Vb);
  HIP_OK(hipGetLastError());
  return 0;
}

// Stage combination and finalise of its state (TimeIntegrator::finalise,
// timeInt.h:117-132): cons2prim from Vg, the SRMHD retry.
int rk_finish(mgpu_engine *e, const RkStageArgs &a, C2PCounters *st) {
  launch_rkplus_stage(a, e->cfg.model, e->stream);
  HIP_OK(hipGetLastError());
  if (mark(e, "rkplus_stage_c2p")) return -1;
  return retry(e, a.Uout, a.Vout, a.Aout, st);
}

void rk_terms(RkStageArgs &a, std::initializer_list<const double *> x, std::initializer_list<double> w) {
  a.nx = 0;
  auto wi = w.begin();
  for (const double *p : x) {
    a.X[a.nx] = p;
    a.w[a.nx++] = *wi++;
  }
}

void rk_out(RkStageArgs &a, double *U, double *V, double *A, const double *Vg) {
  a.Uout = U;
  a.Vout = V;
  a.Aout = A;
  a.Vg = Vg;
}

// RK2B, RK3, RK4 and RK4_10::step (RKPlus.cc:127-726). A stage's cons2prim
// starts from the previous stage's prims, the last stage's from the step's.
int rkplus_step(mgpu_engine *e, double dt, C2PCounters *st) {
  double *U = e->U, *V = e->V, *A = e->A;
  double **Us = e->rkU, **Vs = e->rkV, **As = e->rkA;
  switch (e->cfg.integrator) {
    case MGPU_RK2B: {
      RkStageArgs a = rk_args(e, dt, st);
      if (rk_rhs(e, a, U, V, A)) return -1;
      rk_terms(a, {U}, {1.0});
      a.wr = 1.0;
      rk_out(a, Us[0], Vs[0], As[0], V);
      if (rk_finish(e, a, st)) return -1;
      RkStageArgs b = rk_args(e, dt, st);
      if (rk_rhs(e, b, Us[0], Vs[0], As[0])) return -1;
      rk_terms(b, {U, Us[0]}, {1.0, 1.0});
      b.half_sum = 1;
      rk_out(b, U, V, A, V);
      return rk_finish(e, b, st);
    }
    case MGPU_RK3: {
      RkStageArgs a = rk_args(e, dt, st);
      if (rk_rhs(e, a, U, V, A)) return -1;
      rk_terms(a, {U}, {1.0});
      a.wr = 1.0;
      rk_out(a, Us[0], Vs[0], As[0], V);
      if (rk_finish(e, a, st)) return -1;
      RkStageArgs b = rk_args(e, dt, st);
      if (rk_rhs(e, b, Us[0], Vs[0], As[0])) return -1;
      rk_terms(b, {U, Us[0]}, {3.0 / 4.0, 1.0 / 4.0});
      b.wr = 1.0 / 4.0;
      rk_out(b, Us[1], Vs[1], As[1], Vs[0]);
      if (rk_finish(e, b, st)) return -1;
      RkStageArgs c = rk_args(e, dt, st);
      if (rk_rhs(e, c, Us[1], Vs[1], As[1])) return -1;
      rk_terms(c, {U, Us[1]}, {1.0 / 3.0, 2.0 / 3.0});
      c.wr = 2.0 / 3.0;
      rk_out(c, U, V, A, V);
      return rk_finish(e, c, st);
    }
    case MGPU_RK4: {
      // RK4::stage1..stage5 (RKPlus.cc:330-470)
      static const double wn[4] = {1.0, 0.444370493651235, 0.620101851488403, 0.178079954393132};
      static const double wp[4] = {0.0, 0.555629506348765, 0.379898148511597, 0.821920045606868};
      static const double wr[4] = {0.391752226571890, 0.368410593050371, 0.251891774271694, 0.544974750228521};
      for (int k = 0; k < 4; ++k) {
        const double *Ub = k ? Us[k - 1] : U, *Vb = k ? Vs[k - 1] : V, *Ab = k ? As[k - 1] : A;
        RkStageArgs a = rk_args(e, dt, st);
        if (rk_rhs(e, a, Ub, Vb, Ab)) return -1;
        if (k)
          rk_terms(a, {U, Us[k - 1]}, {wn[k], wp[k]});
        else
          rk_terms(a, {U}, {wn[0]});
        a.wr = wr[k];
        if (k == 3) a.Rout = e->rkR;
        rk_out(a, Us[k], Vs[k], As[k], Vb);
        if (rk_finish(e, a, st)) return -1;
      }
      RkStageArgs a = rk_args(e, dt, st);
      if (rk_rhs(e, a, Us[3], Vs[3], As[3])) return -1;
      rk_terms(a, {Us[1], Us[2], Us[3]}, {0.517231671970585, 0.096059710526147, 0.386708617503269});
      a.R = e->rkR;
      a.wR = 0.063692468666290;
      a.wr = 0.226007483236906;
      rk_out(a, U, V, A, V);
      return rk_finish(e, a, st);
    }
    default: {
      // RK4_10::prepare1, stageRepeat x 5, prepare2, stageRepeat x 4,
      // stageFinal (RKPlus.cc:586-726)
      const size_t cb = sizeof(double) * e->ncons * e->geo.vstride;
      const size_t pb = sizeof(double) * e->nprims * e->geo.vstride;
      const size_t ab = sizeof(double) * e->naux * e->geo.vstride;
      HIP_OK(hipMemcpyAsync(Us[0], U, cb, hipMemcpyDeviceToDevice, e->stream));
      HIP_OK(hipMemcpyAsync(Us[1], U, cb, hipMemcpyDeviceToDevice, e->stream));
      HIP_OK(hipMemcpyAsync(Vs[0], V, pb, hipMemcpyDeviceToDevice, e->stream));
      if (As[0]) HIP_OK(hipMemcpyAsync(As[0], A, ab, hipMemcpyDeviceToDevice, e->stream));
      for (int s = 0; s < 9; ++s) {
        if (s == 5) {
          HIP_OK(hipMemcpyAsync(Vs[0], V, pb, hipMemcpyDeviceToDevice, e->stream));
          if (As[0]) HIP_OK(hipMemcpyAsync(As[0], A, ab, hipMemcpyDeviceToDevice, e->stream));
          launch_rk410_prepare2(e->geo, e->ncons, Us[0], Us[1], e->stream);
          HIP_OK(hipGetLastError());
          if (!e->cfg.quirks) {
            // the new stage state gets its own primitives before its rhs
            RkStageArgs c = rk_args(e, dt, st);
            rk_terms(c, {Us[0]}, {1.0});
            rk_out(c, Us[0], Vs[0], As[0], Vs[0]);
            if (rk_finish(e, c, st)) return -1;
          }
        }
        RkStageArgs a = rk_args(e, dt, st);
        // with the quirk the first stage after prepare2 pairs new cons with
        // the step's prims and aux: its fluxes come from those stored arrays
        if (s == 5 && e->cfg.quirks) {
          if (rk_rhs_stored(e, a, Us[0], Vs[0], As[0])) return -1;
        } else if (rk_rhs(e, a, Us[0], Vs[0], As[0])) {
          return -1;
        }
        rk_terms(a, {Us[0]}, {1.0});
        a.wr = 1.0 / 6.0;
        rk_out(a, Us[0], Vs[0], As[0], Vs[0]);
        if (rk_finish(e, a, st)) return -1;
      }
      RkStageArgs a = rk_args(e, dt, st);
      if (rk_rhs(e, a, Us[0], Vs[0], As[0])) return -1;
      rk_terms(a, {Us[1], Us[0]}, {1.0, 3.0 / 5.0});
      a.wr = 1.0 / 10.0;
      rk_out(a, U, V, A, V);
      return rk_finish(e, a, st);
    }
  }
}

int do_step(mgpu_engine *e, double dt, mgpu_stats *stats) {
  C2PCounters *st = stats ? e->dstats : nullptr;
  if (stats && zero_stats(e)) return -1;
  if (e->cfg.integrator == MGPU_SSP2) {
    if (ssp2_step(e, dt, st)) return -1;
  } else if (e->cfg.integrator == MGPU_RKSPLIT) {
    if (rksplit_step(e, dt, st)) return -1;
  } else if (e->cfg.integrator >= MGPU_RK2B) {
    if (rkplus_step(e, dt, st)) return -1;
  } else if (rk2_step(e, dt, st)) {
    return -1;
  }
  if (stats) {
    if (read_stats(e, stats)) return -1;
    return stats->failures > 0 ? (int)std::min<long long>(stats->failures, 1 << 30) : 0;
  }
  return 0;
}

double cfl_dt(const mgpu_engine *e, double end_time) {
  // Simulation::updateTime (simulation.cc:113-124), alpha = 1.
  const double sd = std::sqrt((double)e->geo.dims);
  const double dtX = e->cfg.cfl * e->dx / (1.0 * sd);
  const double dtY = e->cfg.cfl * e->dy / (1.0 * sd);
  const double dtZ = e->cfg.cfl * e->dz / (1.0 * sd);
  double dt = (dtX <= dtY && dtX <= dtZ) ? dtX : ((dtY < dtZ) ? dtY : dtZ);
  if (e->iters < 15) dt *= 0.1;
  if (e->iters >= 15 && e->iters < 30) dt *= 0.25;
  if (e->t + dt > end_time) dt = end_time - e->t;
  return dt;
}

}  // namespace

extern "C" {

const char *mgpu_last_error(void) { return g_err.c_str(); }

int mgpu_create(const mgpu_config *cfg, mgpu_engine **out) {
  if (!cfg || !out) return fail("mgpu_create: null argument");
  *out = nullptr;
  const mgpu_config &c = *cfg;
  if (c.model != MGPU_SRMHD && c.model != MGPU_SRRMHD) return fail("mgpu_create: unknown model %d", c.model);
  if (c.integrator < MGPU_RK2 || c.integrator > MGPU_RK4_10)
    return fail("mgpu_create: unknown integrator %d", c.integrator);
  if (c.model == MGPU_SRMHD && c.integrator == MGPU_SSP2)
    return fail("mgpu_create: SSP2 is provided with the SRRMHD model only");
  if (c.weno_order != 3 && c.weno_order != 5) return fail("mgpu_create: weno_order must be 3 or 5");
  const int shift = (c.weno_order + 1) / 2;
  if (c.Ng < shift + 1)  // WenoBase::checkSufficientGhostZones (weno.cc:13-20)
    return fail("You must increase number of boundary cells. (order %d needs Ng >= %d, have %d)", c.weno_order,
                shift + 1, c.Ng);
  if (c.nx < 1 || c.ny < 0 || c.nz < 0) return fail("mgpu_create: need nx >= 1, ny >= 0, nz >= 0");
  if (c.sigma < 0.0) return fail("Conductivity must be non-negative, sigma >= 0.");  // simData.cc:100-102
  for (int ax = 0; ax < 3; ++ax)
    if (c.bc[ax] != MGPU_BC_PERIODIC && c.bc[ax] != MGPU_BC_OUTFLOW) return fail("mgpu_create: bad bc on axis %d", ax);
  // 1-D grids have ny = nz = 0, 2-D grids nz = 0 (Data::initData, simData.cc:74-97)
  if (c.ny == 0 && c.nz > 0) return fail("mgpu_create: a grid with nz > 0 needs ny > 0");
  const int n1 = c.ny > 0 ? c.ny : 1;
  const int n2 = c.nz > 0 ? c.nz : 1;
  if (c.nx < c.Ng || (c.ny > 0 && c.ny < c.Ng) || (c.nz > 0 && c.nz < c.Ng))
    return fail("mgpu_create: every active axis needs at least Ng interior cells");
  HIP_OK(hipSetDevice(c.device));
  // the allocation-free cell-list cons2prim's scratch address on this GPU
  if (noalloc_scratch_setup(c.device))
    return fail("mgpu_create: hipGetSymbolAddress of the cell-list scratch failed on device %d", c.device);

  mgpu_engine *e = new mgpu_engine();
  e->cfg = c;
  const bool res = c.model == MGPU_SRRMHD;
  e->ncons = res ? kSrrmhdCons : kSrmhdCons;
  e->nprims = res ? kSrrmhdPrims : kSrmhdPrims;
  e->naux = res ? kSrrmhdAux : kSrmhdAux;
  Geo &g = e->geo;
  g.n[0] = c.nx;
  g.n[1] = n1;
  g.n[2] = n2;
  g.stride[2] = 1;
  g.stride[1] = n2;
  g.stride[0] = (long long)n1 * n2;
  g.vstride = (long long)c.nx * n1 * n2;
  g.dims = c.nz > 0 ? 3 : (c.ny > 0 ? 2 : 1);
  for (int ax = 0; ax < 3; ++ax) g.periodic[ax] = c.bc[ax] == MGPU_BC_PERIODIC;
  // Data::initData / Simulation::Simulation (simData.cc:74-79, simulation.cc:34-36)
  e->dx = (c.xmax - c.xmin) / c.nx;
  e->dy = c.ny > 0 ? (c.ymax - c.ymin) / c.ny : INFINITY;
  e->dz = c.nz > 0 ? (c.zmax - c.zmin) / c.nz : INFINITY;
  g.inv_d[0] = 1.0 / e->dx;
  g.inv_d[1] = c.ny > 0 ? 1.0 / e->dy : 0.0;
  g.inv_d[2] = c.nz > 0 ? 1.0 / e->dz : 0.0;
  PadGeo &p = e->pad;
  p.N[0] = c.nx + 2 * c.Ng;
  p.N[1] = c.ny > 0 ? c.ny + 2 * c.Ng : 1;
  p.N[2] = c.nz > 0 ? c.nz + 2 * c.Ng : 1;
  p.off[0] = c.Ng;
  p.off[1] = c.ny > 0 ? c.Ng : 0;
  p.off[2] = c.nz > 0 ? c.Ng : 0;
  p.ntot = (long long)p.N[0] * p.N[1] * p.N[2];

  auto cleanup = [&](int rc) {
    mgpu_destroy(e);
    return rc;
  };
  const size_t n = (size_t)g.vstride;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) return cleanup(fail("stream create"));
  struct {
    double **ptr;
    size_t elems;
  } bufs[] = {{&e->U, n * e->ncons},  {&e->V, n * e->nprims}, {&e->A, n * e->naux},
              {&e->U1, n * e->ncons}, {&e->V1, n * e->nprims}, {&e->fnet, n * e->ncons},
              {&e->U2, n * e->ncons}, {&e->V2, n * e->nprims}, {&e->Pf, n * e->ncons},
              {&e->S1, n * 5}};
  const int nbufs = c.integrator == MGPU_SSP2 ? 10 : 6;
  for (int bi = 0; bi < nbufs; ++bi) {
    auto &b = bufs[bi];
    if (hipMalloc(b.ptr, b.elems * sizeof(double)) != hipSuccess)
      return cleanup(fail("hipMalloc of %zu bytes failed", b.elems * sizeof(double)));
    if (hipMemset(*b.ptr, 0, b.elems * sizeof(double)) != hipSuccess) return cleanup(fail("hipMemset failed"));
  }
  if (c.model == MGPU_SRMHD && (c.integrator == MGPU_RK2 || c.integrator == MGPU_RKSPLIT)) {
    if (hipMalloc(&e->root, 2 * n * sizeof(double)) != hipSuccess)
      return cleanup(fail("hipMalloc of the cons2prim root store failed"));
  }
  if (c.model == MGPU_SRRMHD && c.integrator == MGPU_SSP2) {
    if (hipMalloc(&e->root, n * sizeof(double)) != hipSuccess)
      return cleanup(fail("hipMalloc of the cons2prim root store failed"));
  }
  if (c.integrator >= MGPU_RK2B) {
    // stage states of RK2B, RK3, RK4, RK4_10 (RKPlus.cc:56-586)
    static const int kStages[4] = {1, 2, 4, 2};
    const int ns = kStages[c.integrator - MGPU_RK2B];
    for (int k = 0; k < ns; ++k) {
      struct {
        double **ptr;
        size_t elems;
      } sb[] = {{&e->rkU[k], n * e->ncons}, {&e->rkV[k], n * e->nprims}, {&e->rkA[k], res ? n * e->naux : 0}};
      for (auto &b : sb) {
        if (!b.elems) continue;
        if (hipMalloc(b.ptr, b.elems * sizeof(double)) != hipSuccess)
          return cleanup(fail("hipMalloc of %zu bytes failed", b.elems * sizeof(double)));
        if (hipMemset(*b.ptr, 0, b.elems * sizeof(double)) != hipSuccess) return cleanup(fail("hipMemset failed"));
      }
    }
    if (c.integrator == MGPU_RK4 && hipMalloc(&e->rkR, n * e->ncons * sizeof(double)) != hipSuccess)
      return cleanup(fail("hipMalloc of the RK4 rhs store failed"));
    if (c.integrator == MGPU_RK4_10 && hipMalloc(&e->rkF, n * e->ncons * g.dims * sizeof(double)) != hipSuccess)
      return cleanup(fail("hipMalloc of the RK4_10 flux store failed"));
  }
  if (hipMalloc(&e->dstats, sizeof(C2PCounters) * kStatBuckets) != hipSuccess || hipMalloc(&e->fail_count, sizeof(int)) != hipSuccess ||
      hipMalloc(&e->fail_list, n * sizeof(long long)) != hipSuccess || hipMalloc(&e->fail_flag, n) != hipSuccess)
    return cleanup(fail("hipMalloc of C2P work arrays failed"));
  if (hipMemset(e->fail_count, 0, sizeof(int)) != hipSuccess || hipMemset(e->fail_flag, 0, n) != hipSuccess ||
      hipMemset(e->dstats, 0, sizeof(C2PCounters) * kStatBuckets) != hipSuccess)
    return cleanup(fail("hipMemset failed"));
  *out = e;
  return 0;
}

int mgpu_destroy(mgpu_engine *e) {
  if (!e) return 0;
  if (e->stream) hipStreamSynchronize(e->stream);
  double *bufs[] = {e->U,  e->V,  e->A,  e->U1, e->V1, e->fnet,     e->staging,
                    e->U2, e->V2, e->Pf, e->S1, e->root};
  for (double *b : bufs)
    if (b) hipFree(b);
  for (int k = 0; k < 4; ++k) {
    if (e->rkU[k]) hipFree(e->rkU[k]);
    if (e->rkV[k]) hipFree(e->rkV[k]);
    if (e->rkA[k]) hipFree(e->rkA[k]);
  }
  if (e->rkR) hipFree(e->rkR);
  if (e->rkF) hipFree(e->rkF);
  if (e->rg_work) hipFree(e->rg_work);
  if (e->rg_src) hipFree(e->rg_src);
  if (e->dstats) hipFree(e->dstats);
  if (e->dspeed) hipFree(e->dspeed);
  if (e->fail_count) hipFree(e->fail_count);
  if (e->fail_list) hipFree(e->fail_list);
  if (e->fail_flag) hipFree(e->fail_flag);
  if (e->mlp) hipFree(e->mlp);
  if (e->hist) hipFree(e->hist);
  if (e->cells_buf) hipFree(e->cells_buf);
  if (e->cells_part) hipFree(e->cells_part);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
  return 0;
}

int mgpu_sizes(const mgpu_engine *e, int *ncons, int *nprims, int *naux) {
  if (!e) return fail("null engine");
  if (ncons) *ncons = e->ncons;
  if (nprims) *nprims = e->nprims;
  if (naux) *naux = e->naux;
  return 0;
}

int mgpu_upload_state(mgpu_engine *e, const double *cons, const double *prims, const double *aux) {
  if (!e) return fail("null engine");
  if (ensure_staging(e)) return -1;
  // an aux left as a root belongs to the state about to be replaced: form it
  // now, so a partial upload keeps it as the arrays it does not replace
  if (!aux && fresh_aux(e)) return -1;
  const double *src[3] = {cons, prims, aux};
  double *dst[3] = {e->U, e->V, e->A};
  const int nv[3] = {e->ncons, e->nprims, e->naux};
  for (int a = 0; a < 3; ++a) {
    if (!src[a]) continue;
    HIP_OK(hipMemcpyAsync(e->staging, src[a], sizeof(double) * nv[a] * e->pad.ntot, hipMemcpyHostToDevice, e->stream));
    launch_compact(e->geo, e->pad, nv[a], e->staging, dst[a], e->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(e->stream));
  }
  if (aux) e->aux_stale = false;
  return 0;
}

int mgpu_download_state(mgpu_engine *e, double *cons, double *prims, double *aux) {
  if (!e) return fail("null engine");
  if (ensure_staging(e)) return -1;
  if (aux && fresh_aux(e)) return -1;
  double *dst[3] = {cons, prims, aux};
  const double *src[3] = {e->U, e->V, e->A};
  const int nv[3] = {e->ncons, e->nprims, e->naux};
  for (int a = 0; a < 3; ++a) {
    if (!dst[a]) continue;
    launch_expand(e->geo, e->pad, nv[a], src[a], e->staging, 0, e->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(dst[a], e->staging, sizeof(double) * nv[a] * e->pad.ntot, hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipStreamSynchronize(e->stream));
  }
  return 0;
}

int mgpu_prims_to_all(mgpu_engine *e) {
  if (!e) return fail("null engine");
  if (e->cfg.model == MGPU_SRMHD)
    launch_prims_to_all_srmhd(e->geo, e->V, e->U, e->A, e->cfg.gamma, e->stream);
  else
    launch_prims_to_all_srrmhd(e->geo, e->V, e->U, e->A, e->cfg.gamma, e->cfg.sigma, e->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(e->stream));
  e->aux_stale = false;
  return 0;
}

int mgpu_step(mgpu_engine *e, double dt, mgpu_stats *stats) {
  if (!e) return fail("null engine");
  if (!(dt > 0.0)) return fail("mgpu_step: dt must be positive");
  return do_step(e, dt, stats);
}

int mgpu_update_time(mgpu_engine *e, double end_time, mgpu_stats *stats) {
  if (!e) return fail("null engine");
  const double dt = cfl_dt(e, end_time);
  if (!(dt > 0.0)) return fail("mgpu_update_time: end time reached");
  e->dt = dt;
  const int rc = do_step(e, dt, stats);
  if (rc < 0) return rc;
  e->t += dt;
  e->iters++;
  return rc;
}

int mgpu_get_time(const mgpu_engine *e, double *t, double *dt, int *iters) {
  if (!e) return fail("null engine");
  if (t) *t = e->t;
  if (dt) *dt = e->dt;
  if (iters) *iters = e->iters;
  return 0;
}

int mgpu_set_time(mgpu_engine *e, double t, int iters) {
  if (!e) return fail("null engine");
  e->t = t;
  e->iters = iters;
  return 0;
}

int mgpu_fnet(mgpu_engine *e, double *fnet) {
  if (!e || !fnet) return fail("null argument");
  if (ensure_staging(e)) return -1;
  SweepArgs a = sweep_args(e, e->U, e->V, nullptr);
  if (run_sweeps(e, a, kSweepAdd)) return -1;
  launch_expand(e->geo, e->pad, e->ncons, e->fnet, e->staging, 1, e->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(fnet, e->staging, sizeof(double) * e->ncons * e->pad.ntot, hipMemcpyDeviceToHost, e->stream));
  HIP_OK(hipStreamSynchronize(e->stream));
  return 0;
}

int mgpu_flux_vector(mgpu_engine *e, int dir, double *f) {
  if (!e || !f) return fail("null argument");
  if (dir < 0 || dir >= e->geo.dims) return fail("mgpu_flux_vector: dir %d out of range", dir);
  if (ensure_staging(e)) return -1;
  if (e->cfg.model == MGPU_SRMHD)
    launch_flux_vector_srmhd(e->geo, e->pad, e->U, e->V, e->staging, dir, e->cfg.quirks, e->stream);
  else
    launch_flux_vector_srrmhd(e->geo, e->pad, e->U, e->V, e->staging, dir, e->cfg.quirks, e->cfg.gamma,
                              e->cfg.sigma, e->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(f, e->staging, sizeof(double) * e->ncons * e->pad.ntot, hipMemcpyDeviceToHost, e->stream));
  HIP_OK(hipStreamSynchronize(e->stream));
  return 0;
}

int mgpu_set_model_extension(mgpu_engine *e, int kind) {
  if (!e) return fail("null engine");
  if (kind != MGPU_EXT_NONE && kind != MGPU_EXT_REGIME) return fail("mgpu_set_model_extension: unknown kind %d", kind);
  if (kind == MGPU_EXT_REGIME && e->cfg.model != MGPU_SRMHD)
    return fail("mgpu_set_model_extension: REGIME extends the ideal SRMHD model only");
  // the REGIME kernels index the padded grid with 32-bit cell numbers
  if (kind == MGPU_EXT_REGIME && e->pad.ntot >= (1LL << 31))
    return fail("mgpu_set_model_extension: REGIME needs fewer than 2^31 padded cells (have %lld)", e->pad.ntot);
  if (kind == MGPU_EXT_REGIME && !e->rg_work) {
    const size_t nw = regime_work_doubles(e->pad), ns = (size_t)8 * e->geo.vstride;
    if (hipMalloc(&e->rg_work, nw * sizeof(double)) != hipSuccess ||
        hipMalloc(&e->rg_src, ns * sizeof(double)) != hipSuccess)
      return fail("hipMalloc of the REGIME work arrays failed");
  }
  e->extension = kind;
  return 0;
}

int mgpu_source_extension(mgpu_engine *e, double *source) {
  if (!e || !source) return fail("null argument");
  if (e->extension != MGPU_EXT_REGIME) return fail("mgpu_source_extension: no model extension is set");
  if (ensure_staging(e)) return -1;
  extension_source(e, e->V);
  HIP_OK(hipGetLastError());
  // variables 0..7 of the reference's 9-variable source array; phi's stays 0
  HIP_OK(hipMemsetAsync(e->staging, 0, sizeof(double) * e->ncons * e->pad.ntot, e->stream));
  launch_expand(e->geo, e->pad, 8, e->rg_src, e->staging, 1, e->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(source, e->staging, sizeof(double) * e->ncons * e->pad.ntot, hipMemcpyDeviceToHost,
                        e->stream));
  HIP_OK(hipStreamSynchronize(e->stream));
  return 0;
}

int mgpu_regime_fields(mgpu_engine *e, double *E, double *q, double *alpha, double *dwdsb, double *K) {
  if (!e) return fail("null engine");
  if (e->extension != MGPU_EXT_REGIME) return fail("mgpu_regime_fields: no model extension is set");
  extension_source(e, e->V, true);
  HIP_OK(hipGetLastError());
  const RegimeFields f = regime_fields_of(e->pad, e->rg_work);
  const size_t nb = sizeof(double) * e->pad.ntot;
  if (E) HIP_OK(hipMemcpyAsync(E, f.E, 3 * nb, hipMemcpyDeviceToHost, e->stream));
  if (q) HIP_OK(hipMemcpyAsync(q, f.q, nb, hipMemcpyDeviceToHost, e->stream));
  if (alpha) HIP_OK(hipMemcpyAsync(alpha, f.alpha, nb, hipMemcpyDeviceToHost, e->stream));
  if (K) HIP_OK(hipMemcpyAsync(K, f.K, 3 * nb, hipMemcpyDeviceToHost, e->stream));
  if (dwdsb) {
    double *tmp = nullptr;
    HIP_OK(hipMalloc(&tmp, 36 * nb));
    launch_regime_dwdsb(e->geo, e->pad, e->V, e->rg_work, e->cfg.sigma, tmp, e->stream);
    const hipError_t le = hipGetLastError();
    const hipError_t ce = le == hipSuccess ? hipMemcpyAsync(dwdsb, tmp, 36 * nb, hipMemcpyDeviceToHost, e->stream) : le;
    const hipError_t se = hipStreamSynchronize(e->stream);
    hipFree(tmp);
    if (ce != hipSuccess || se != hipSuccess) return fail("mgpu_regime_fields: dwdsb readback failed");
  }
  HIP_OK(hipStreamSynchronize(e->stream));
  return 0;
}

int mgpu_source(mgpu_engine *e, double *source) {
  if (!e || !source) return fail("null argument");
  if (ensure_staging(e)) return -1;
  if (e->cfg.model == MGPU_SRMHD)
    launch_source_srmhd(e->geo, e->pad, e->U, e->staging, e->cfg.cp, e->stream);
  else
    launch_source_srrmhd(e->geo, e->pad, e->U, e->V, e->staging, e->cfg.gamma, e->cfg.sigma, e->cfg.cp, e->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(source, e->staging, sizeof(double) * e->ncons * e->pad.ntot, hipMemcpyDeviceToHost,
                        e->stream));
  HIP_OK(hipStreamSynchronize(e->stream));
  return 0;
}

int mgpu_flux_reconstruction(mgpu_engine *e, const double *cons, const double *f, int dir, int vars,
                             double *frecon) {
  if (!e || !frecon) return fail("mgpu_flux_reconstruction: null argument");
  if (dir < 0 || dir > 2) return fail("mgpu_flux_reconstruction: dir %d out of range", dir);
  if ((cons == nullptr) != (f == nullptr))
    return fail("mgpu_flux_reconstruction: give both cons and f, or neither for the resident state");
  const int nv = vars < 0 ? e->ncons : vars;
  if (nv > e->ncons) return fail("mgpu_flux_reconstruction: vars %d exceeds the model's %d", vars, e->ncons);
  if (ensure_staging(e)) return -1;
  const size_t plane = (size_t)e->pad.ntot;
  // staging: cons [ncons][padded], f [ncons][padded], frecon [ncons][padded]
  double *sc = e->staging, *sf = sc + (size_t)e->ncons * plane, *sr = sf + (size_t)e->ncons * plane;
  if (nv > 0) {
    if (cons) {
      HIP_OK(hipMemcpyAsync(sc, cons, sizeof(double) * nv * plane, hipMemcpyHostToDevice, e->stream));
      HIP_OK(hipMemcpyAsync(sf, f, sizeof(double) * nv * plane, hipMemcpyHostToDevice, e->stream));
    } else {
      launch_expand(e->geo, e->pad, nv, e->U, sc, 0, e->stream);
      if (dir < e->geo.dims) {
        if (e->cfg.model == MGPU_SRMHD)
          launch_flux_vector_srmhd(e->geo, e->pad, e->U, e->V, sf, dir, e->cfg.quirks, e->stream);
        else
          launch_flux_vector_srrmhd(e->geo, e->pad, e->U, e->V, sf, dir, e->cfg.quirks, e->cfg.gamma,
                                    e->cfg.sigma, e->stream);
      } else {
        // an axis the grid does not have: every face lies outside the range
        HIP_OK(hipMemsetAsync(sf, 0, sizeof(double) * nv * plane, e->stream));
      }
      HIP_OK(hipGetLastError());
    }
    launch_flux_recon(e->pad, e->cfg.weno_order, dir, nv, sc, sf, sr, e->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(frecon, sr, sizeof(double) * nv * plane, hipMemcpyDeviceToHost, e->stream));
  }
  HIP_OK(hipStreamSynchronize(e->stream));
  return 0;
}

namespace {
// Device room for n cells of the single-cell layout and their counter records.
int ensure_cells(mgpu_engine *e, long long n) {
  const size_t need = (size_t)n * (e->ncons + e->nprims + e->naux + e->ncons);
  if (e->cells_elems < need) {
    if (e->cells_buf) HIP_OK(hipFree(e->cells_buf));
    e->cells_buf = nullptr;
    e->cells_elems = 0;
    HIP_OK(hipMalloc(&e->cells_buf, need * sizeof(double)));
    e->cells_elems = need;
  }
  const long long nb = cells_blocks(n) + 1;
  if (e->cells_part_n < nb) {
    if (e->cells_part) HIP_OK(hipFree(e->cells_part));
    e->cells_part = nullptr;
    e->cells_part_n = 0;
    HIP_OK(hipMalloc(&e->cells_part, nb * sizeof(CellCounters)));
    e->cells_part_n = nb;
  }
  return 0;
}
}  // namespace

int mgpu_c2p_cells(mgpu_engine *e, long long n, const double *cons, double *prims, double *aux,
                   mgpu_stats *stats) {
  if (!e) return fail("null engine");
  if (n < 0) return fail("mgpu_c2p_cells: negative cell count");
  if (n > 0 && (!cons || !prims || !aux)) return fail("mgpu_c2p_cells: null array");
  if (ensure_cells(e, std::max(n, 1LL))) return -1;
  double *dc = e->cells_buf, *dp = dc + (size_t)n * e->ncons, *da = dp + (size_t)n * e->nprims;
  const size_t bc = sizeof(double) * n * e->ncons, bp = sizeof(double) * n * e->nprims,
               ba = sizeof(double) * n * e->naux;
  if (n > 0) {
    HIP_OK(hipMemcpyAsync(dc, cons, bc, hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipMemcpyAsync(dp, prims, bp, hipMemcpyHostToDevice, e->stream));
    HIP_OK(hipMemcpyAsync(da, aux, ba, hipMemcpyHostToDevice, e->stream));
  }
  CellCounters *total = e->cells_part + cells_blocks(n);
  launch_c2p_cells(e->cfg.model, n, dc, dp, da, e->cfg.gamma, e->cfg.sigma, e->cells_part, total, e->stream);
  HIP_OK(hipGetLastError());
  CellCounters t;
  if (n > 0) {
    HIP_OK(hipMemcpyAsync(prims, dp, bp, hipMemcpyDeviceToHost, e->stream));
    HIP_OK(hipMemcpyAsync(aux, da, ba, hipMemcpyDeviceToHost, e->stream));
  }
  HIP_OK(hipMemcpyAsync(&t, total, sizeof t, hipMemcpyDeviceToHost, e->stream));
  HIP_OK(hipStreamSynchronize(e->stream));
  if (stats) {
    memset(stats, 0, sizeof *stats);
    stats->cells = (long long)t.cells;
    stats->iterations = (long long)t.iterations;
    stats->failures = (long long)t.failures;
    stats->max_iterations = (long long)t.max_iterations;
  }
  return t.failures > 0 ? (int)std::min<unsigned long long>(t.failures, 1ULL << 30) : 0;
}

int mgpu_c2p_cells_noalloc(mgpu_engine *e, long long n, const double *cons, double *prims, double *aux,
                           mgpu_stats *stats) {
  if (!e) return fail("null engine");
  if (n < 0) return fail("mgpu_c2p_cells_noalloc: negative cell count");
  if (n > kNoAllocMaxCells)
    return fail("mgpu_c2p_cells_noalloc: %lld cells exceed the maximum of %lld per call", n, kNoAllocMaxCells);
  if (n > 0 && (!cons || !prims || !aux)) return fail("mgpu_c2p_cells_noalloc: null array");
  CellCounters *part = noalloc_scratch(e->cfg.device);
  if (!part) return fail("mgpu_c2p_cells_noalloc: no scratch address for device %d", e->cfg.device);
  // every record the launch reads is written by it first: nothing to clear
  CellCounters *total = part + cells_blocks(n);
  launch_c2p_cells(e->cfg.model, n, cons, prims, aux, e->cfg.gamma, e->cfg.sigma, part, total, e->stream);
  HIP_OK(hipGetLastError());
  if (!stats) return 0;
  CellCounters t;
  HIP_OK(hipMemcpyAsync(&t, total, sizeof t, hipMemcpyDeviceToHost, e->stream));
  HIP_OK(hipStreamSynchronize(e->stream));
  memset(stats, 0, sizeof *stats);
  stats->cells = (long long)t.cells;
  stats->iterations = (long long)t.iterations;
  stats->failures = (long long)t.failures;
  stats->max_iterations = (long long)t.max_iterations;
  return t.failures > 0 ? (int)std::min<unsigned long long>(t.failures, 1ULL << 30) : 0;
}

int mgpu_source_cells(mgpu_engine *e, long long n, const double *cons, const double *prims, const double *aux,
                      double *source) {
  (void)prims;
  if (!e) return fail("null engine");
  if (n < 0) return fail("mgpu_source_cells: negative cell count");
  const bool needs_aux = e->cfg.model != MGPU_SRMHD;
  if (n > 0 && (!cons || !source || (needs_aux && !aux))) return fail("mgpu_source_cells: null array");
  if (n == 0) return 0;
  if (ensure_cells(e, n)) return -1;
  double *dc = e->cells_buf, *da = dc + (size_t)n * e->ncons, *ds = da + (size_t)n * e->naux;
  HIP_OK(hipMemcpyAsync(dc, cons, sizeof(double) * n * e->ncons, hipMemcpyHostToDevice, e->stream));
  if (needs_aux) HIP_OK(hipMemcpyAsync(da, aux, sizeof(double) * n * e->naux, hipMemcpyHostToDevice, e->stream));
  launch_source_cells(e->cfg.model, n, dc, da, e->cfg.cp, ds, e->stream);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(source, ds, sizeof(double) * n * e->ncons, hipMemcpyDeviceToHost, e->stream));
  HIP_OK(hipStreamSynchronize(e->stream));
  return 0;
}

namespace {
int launch_c2p(mgpu_engine *e) {
  if (e->cfg.model == MGPU_SRMHD && e->guess_mode == MGPU_GUESS_LEARNED)
    launch_c2p_learned_srmhd(e->geo, e->U, e->V, e->A, e->mlp, e->cfg.gamma, e->dstats, e->hist, e->stream);
  else if (e->cfg.model == MGPU_SRMHD)
    launch_c2p_srmhd(e->geo, e->U, e->V, e->A, e->cfg.gamma, e->dstats, e->fail_count, e->fail_list, e->fail_flag,
                     e->guess_mode == MGPU_GUESS_COLD, e->hist, e->stream);
  else
    launch_c2p_srrmhd(e->geo, e->U, e->V, e->A, e->cfg.gamma, e->cfg.sigma, e->dstats, e->stream);
  HIP_OK(hipGetLastError());
  e->aux_stale = false;  // every cell's aux written
  return 0;
}
}  // namespace

int mgpu_c2p(mgpu_engine *e, mgpu_stats *stats) {
  if (!e) return fail("null engine");
  if (zero_stats(e)) return -1;
  if (launch_c2p(e)) return -1;
  if (e->guess_mode != MGPU_GUESS_LEARNED && retry(e, e->U, e->V, e->A, e->dstats)) return -1;
  mgpu_stats local;
  if (read_stats(e, &local)) return -1;
  if (stats) *stats = local;
  return local.failures > 0 ? (int)std::min<long long>(local.failures, 1 << 30) : 0;
}

int mgpu_timed_c2p(mgpu_engine *e, int reps, double *ms_per_call, mgpu_stats *stats) {
  if (!e || reps <= 0) return fail("mgpu_timed_c2p: bad arguments");
  if (ensure_staging(e)) return -1;
  const size_t bytes = sizeof(double) * e->nprims * e->geo.vstride;
  HIP_OK(hipMemcpyAsync(e->staging, e->V, bytes, hipMemcpyDeviceToDevice, e->stream));
  hipEvent_t a, b;
  HIP_OK(hipEventCreate(&a));
  HIP_OK(hipEventCreate(&b));
  double total = 0.0;
  int rc = 0;
  for (int r = 0; r < reps && rc == 0; ++r) {
    HIP_OK(hipMemcpyAsync(e->V, e->staging, bytes, hipMemcpyDeviceToDevice, e->stream));
    if (zero_stats(e)) rc = -1;
    HIP_OK(hipEventRecord(a, e->stream));
    if (rc == 0 && launch_c2p(e)) rc = -1;
    HIP_OK(hipEventRecord(b, e->stream));
    HIP_OK(hipEventSynchronize(b));
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, a, b));
    total += ms;
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (rc) return rc;
  if (ms_per_call) *ms_per_call = total / reps;
  mgpu_stats local;
  if (read_stats(e, &local)) return -1;
  if (stats) *stats = local;
  return local.failures > 0 ? (int)std::min<long long>(local.failures, 1 << 30) : 0;
}

int mgpu_set_c2p_guess(mgpu_engine *e, int mode, const float *params, const float *ymean_ystd) {
  if (!e) return fail("null engine");
  if (mode != MGPU_GUESS_PLAIN && mode != MGPU_GUESS_LEARNED && mode != MGPU_GUESS_COLD)
    return fail("mgpu_set_c2p_guess: unknown mode %d", mode);
  if (mode != MGPU_GUESS_PLAIN && e->cfg.model != MGPU_SRMHD)
    return fail("mgpu_set_c2p_guess: only the SRMHD cons2prim takes a learned or cold guess");
  if (mode == MGPU_GUESS_LEARNED) {
    if (!params || !ymean_ystd) return fail("mgpu_set_c2p_guess: learned mode needs parameters");
    MlpDev pack;
    mlp_pack(params, ymean_ystd, ymean_ystd + 2, &pack);
    if (!e->mlp) HIP_OK(hipMalloc(&e->mlp, sizeof(MlpDev)));
    HIP_OK(hipMemcpy(e->mlp, &pack, sizeof pack, hipMemcpyHostToDevice));
  }
  e->guess_mode = mode;
  return 0;
}

int mgpu_set_c2p_histogram(mgpu_engine *e, int on) {
  if (!e) return fail("null engine");
  if (on && e->cfg.model != MGPU_SRMHD) return fail("mgpu_set_c2p_histogram: SRMHD engines only");
  if (!on) {
    if (e->hist) HIP_OK(hipFree(e->hist));
    e->hist = nullptr;
    return 0;
  }
  if (!e->hist) HIP_OK(hipMalloc(&e->hist, kHistBins * sizeof(unsigned long long)));
  HIP_OK(hipMemsetAsync(e->hist, 0, kHistBins * sizeof(unsigned long long), e->stream));
  HIP_OK(hipStreamSynchronize(e->stream));
  return 0;
}

int mgpu_get_c2p_histogram(mgpu_engine *e, long long *hist) {
  if (!e || !hist) return fail("null argument");
  if (!e->hist) return fail("mgpu_get_c2p_histogram: the histogram is off");
  unsigned long long h[kHistBins];
  HIP_OK(hipMemcpyAsync(h, e->hist, sizeof h, hipMemcpyDeviceToHost, e->stream));
  HIP_OK(hipStreamSynchronize(e->stream));
  for (int b = 0; b < kHistBins; ++b) hist[b] = (long long)h[b];
  return 0;
}

int mgpu_max_wavespeed(mgpu_engine *e, double out[3]) {
  if (!e || !out) return fail("mgpu_max_wavespeed: null argument");
  if (e->cfg.model != MGPU_SRMHD)
    return fail("mgpu_max_wavespeed: only the SRMHD fast magnetosonic bound is reduced; the resistive "
                "model's light modes travel at 1 in every cell");
  if (!e->dspeed) HIP_OK(hipMalloc(&e->dspeed, 3 * sizeof(unsigned long long)));
  if (fresh_aux(e)) return -1;
  HIP_OK(hipMemsetAsync(e->dspeed, 0, 3 * sizeof(unsigned long long), e->stream));
  launch_max_wavespeed_srmhd(e->geo, e->V, e->A, e->dspeed, e->stream);
  HIP_OK(hipGetLastError());
  unsigned long long bits[3];
  HIP_OK(hipMemcpyAsync(bits, e->dspeed, sizeof bits, hipMemcpyDeviceToHost, e->stream));
  HIP_OK(hipStreamSynchronize(e->stream));
  for (int d = 0; d < 3; ++d) memcpy(&out[d], &bits[d], sizeof(double));
  return 0;
}

int mgpu_set_sweep_plan(mgpu_engine *e, int fused_yz) {
  if (!e) return fail("null engine");
  if (fused_yz < 0 || fused_yz > 3) return fail("mgpu_set_sweep_plan: plan %d out of range", fused_yz);
  e->fused_yz = fused_yz == 1;
  e->xyz_order = fused_yz == 2;
  e->fused_yzf = fused_yz == 3;
  return 0;
}

int mgpu_set_regime_plan(mgpu_engine *e, int fused) {
  if (!e) return fail("null engine");
  e->regime_fused = fused ? 1 : 0;
  return 0;
}

int mgpu_sync(mgpu_engine *e) {
  if (!e) return fail("null engine");
  HIP_OK(hipStreamSynchronize(e->stream));
  return 0;
}

int mgpu_timed_steps(mgpu_engine *e, int nsteps, const double *dts, double *total_ms, double *kernel_ms,
                     int kernel_slots, int *nslots) {
  if (!e || !dts || nsteps <= 0) return fail("mgpu_timed_steps: bad arguments");
  HIP_OK(hipStreamSynchronize(e->stream));
  Marks marks;
  e->marks = &marks;
  int rc = mark(e, "start");
  for (int s = 0; s < nsteps && rc == 0; ++s) rc = do_step(e, dts[s], nullptr);
  e->marks = nullptr;
  if (rc == 0) {
    if (hipStreamSynchronize(e->stream) != hipSuccess) rc = fail("stream sync failed");
  }
  if (rc == 0) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, marks.ev.front(), marks.ev.back());
    if (total_ms) *total_ms = ms;
    std::vector<double> sum(e->slot_names.size(), 0.0);
    std::vector<int> cnt(e->slot_names.size(), 0);
    for (size_t i = 1; i < marks.ev.size(); ++i) {
      float d = 0.f;
      hipEventElapsedTime(&d, marks.ev[i - 1], marks.ev[i]);
      sum[marks.slot[i]] += d;
      cnt[marks.slot[i]] += 1;
    }
    if (nslots) *nslots = (int)e->slot_names.size();
    if (kernel_ms)
      for (int i = 0; i < kernel_slots && i < (int)sum.size(); ++i) kernel_ms[i] = cnt[i] ? sum[i] / cnt[i] : 0.0;
  }
  for (hipEvent_t ev : marks.ev) hipEventDestroy(ev);
  return rc;
}

const char *mgpu_kernel_slot_name(const mgpu_engine *e, int slot) {
  if (!e || slot < 0 || slot >= (int)e->slot_names.size()) return "";
  return e->slot_names[slot].c_str();
}

}  // extern "C"
