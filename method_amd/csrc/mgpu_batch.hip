// Batched ensemble of independent 2-D ideal SRMHD simulations on one GPU.
//
// Every instance behaves as its own Simulation::updateTime loop
// (simulation.cc:104-133) with RK2::step (RK2.cc:93-100) and FVS::F with
// WENO3/5 (fluxVectorSplitting.cc:58-131); the instances differ in grid size,
// initial state and time step. One persistent kernel per RK stage advances
// the tiles of all instances: for a TX x TY tile it reconstructs the y faces
// and then the x faces through LDS, keeps the y net flux in registers, and
// applies the stage update and the Newton cons2prim in place, so the net flux
// never goes to HBM. A clock kernel advances each instance's t, iteration
// count and dt on the device, and one step (stage one, stage two, clock) is
// captured in a hipGraph and replayed. Tile shape, schedule and workgroup
// count come from an autotuner that times the candidates on the batch itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "method_gpu.h"
#include "mgpu_kernels.hpp"

using namespace mgpu;

namespace mgpu {
int engine_fail(const char *fmt, ...);
}

namespace {

constexpr int kThreads = 256;
constexpr int kXcds = 8;  // MI355X: workgroups go round robin over 8 XCDs, each with its own L2

enum { kSchedStatic = 0, kSchedDynamic = 1 };

// Per-instance constants on the device.
struct InstDev {
  int nx, ny;          // interior cells
  int tile_base;       // first tile of this instance in the batch's tile list
  int tiles_y;         // tiles along y
  long long off;       // index of cell (0, 0) in the concatenated cell range
  double inv_dx, inv_dy;
  double dt_base;      // cfl min(dx, dy) / sqrt(2) before the slow-start factors
  double end_time;
};

// Per-instance time state, advanced on the device: t, the dt of the last
// step taken (what the reference's d->dt holds after updateTime), the dt of
// the next step, the iteration count and whether t is short of end_time.
struct Clock {
  double t, dt, next_dt;
  int iters, active;
};

struct BatchArgs {
  const InstDev *inst;
  const Clock *clock;
  const int *tile_inst;  // instance of each tile
  int ntiles, ninst;
  long long vs;          // cells of all instances: the stride between variables
  int px, py;            // periodic in x / y (else outflow)
  const double *U, *V;   // stage input
  const double *Un;      // cons at t^n (stage two)
  const double *Vg;      // prims at t^n, the cons2prim starting guess
  double *Uout, *Vout;
  double *root;          // stage two: cons2prim root (v^2, Z) [2][cell], from which the aux is formed
  double gamma;
  C2PCounters *stats;
  int *fail_count;
  long long *fail_list;
  unsigned char *fail_flag;
  int *sync;             // [0, 8): per-XCD tile counters; [8]: finished workgroups
  int quirks;
  int sched;
};

template <int ORDER, int TX, int TY>
struct BatchShape {
  static constexpr int R = Weno<ORDER>::R;
  static constexpr int NV = kSrmhdCons;
  static constexpr int WY = TY + 2 * R;  // cells per row in the y phase
  static constexpr int HX = TX + 2 * R;  // rows in the x phase
  // one half (f+, then f-) of the split fluxes at a time
  static constexpr int kLdsY = NV * TX * WY;
  static constexpr int kLdsX = NV * HX * TY;
  static constexpr int kLds = kLdsY > kLdsX ? kLdsY : kLdsX;
  static_assert(TX * TY == kThreads, "one cell per thread");
};

// Lax-Friedrichs split flux (alpha = 1) of one cell along DIR
// (fluxVectorSplitting.cc:4-28).
template <int DIR>
__device__ __forceinline__ void split_flux(const BatchArgs &a, long long q, double fp[9], double fm[9]) {
  const long long vs = a.vs;
  double u[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) u[v] = a.U[v * vs + q];
  const double vx = a.V[1 * vs + q], vy = a.V[2 * vs + q], vz = a.V[3 * vs + q], p = a.V[4 * vs + q];
  double f[9];
  srmhd_flux<DIR, false>(u, vx, vy, vz, p, f);
#pragma unroll
  for (int v = 0; v < 9; ++v) {
    fp[v] = 0.5 * (f[v] + u[v]);
    fm[v] = 0.5 * (f[v] - u[v]);
  }
}

// The WENO face values of one tile phase along DIR, left in s_f[item] for
// the NITEMS (variable, face) items. Phase 1 forms the split fluxes of the
// NCELLS tile-and-halo cells (cell_q gives a cell's index, or -1 outside the
// instance, where the fluxes are zero), puts f+ in LDS at s_f[v NCELLS +
// idx] and keeps f- in registers; the upwind half of every item is formed
// from f+ (item_at gives the LDS offset of its first stencil cell, STEP the
// offset between stencil cells), f- takes f+'s place and the downwind halves
// follow. Holding one half at a time halves the LDS of a workgroup, so more
// of them share a CU; the items are dealt out one per thread per pass.
template <int ORDER, int DIR, int NCELLS, int NITEMS, int STEP, class CellQ, class ItemAt>
__device__ __forceinline__ void tile_faces(const BatchArgs &a, double *s_f, int tid, CellQ cell_q, ItemAt item_at) {
  constexpr int NV = kSrmhdCons;
  constexpr int kCellPasses = (NCELLS + kThreads - 1) / kThreads;
  constexpr int kPasses = (NITEMS + kThreads - 1) / kThreads;
  double fm_keep[kCellPasses][NV];
#pragma unroll
  for (int cp = 0; cp < kCellPasses; ++cp) {
    const int idx = tid + cp * kThreads;
    if (idx < NCELLS) {
      double fp[NV];
      const long long q = cell_q(idx);
      if (q >= 0) {
        split_flux<DIR>(a, q, fp, fm_keep[cp]);
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) fp[v] = fm_keep[cp][v] = 0.0;
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) s_f[v * NCELLS + idx] = fp[v];
    }
  }
  // the kept f- values are formed before the barrier
#pragma unroll
  for (int cp = 0; cp < kCellPasses; ++cp)
#pragma unroll
    for (int v = 0; v < NV; ++v) asm volatile("" : "+v"(fm_keep[cp][v]));
  __syncthreads();
  double nu[kPasses], du[kPasses];
#pragma unroll
  for (int k = 0; k < kPasses; ++k) {
    const int item = tid + k * kThreads;
    if (item < NITEMS) {
      const double *fpv = s_f + item_at(item);
      Weno<ORDER>::up([&](int m) { return fpv[m * STEP]; }, nu[k], du[k]);
    }
  }
  // the upwind halves are formed before the barrier after which f- overwrites f+
#pragma unroll
  for (int k = 0; k < kPasses; ++k) asm volatile("" : "+v"(nu[k]), "+v"(du[k]));
  __syncthreads();
#pragma unroll
  for (int cp = 0; cp < kCellPasses; ++cp) {
    const int idx = tid + cp * kThreads;
    if (idx < NCELLS) {
#pragma unroll
      for (int v = 0; v < NV; ++v) s_f[v * NCELLS + idx] = fm_keep[cp][v];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPasses; ++k) {
    const int item = tid + k * kThreads;
    if (item < NITEMS) {
      const double *fmv = s_f + item_at(item);
      double nd, dd;
      Weno<ORDER>::down([&](int m) { return fmv[m * STEP]; }, nd, dd);
      nu[k] = weno_combine(nu[k], du[k], nd, dd);
    }
  }
#pragma unroll
  for (int k = 0; k < kPasses; ++k) asm volatile("" : "+v"(nu[k]));
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPasses; ++k) {
    const int item = tid + k * kThreads;
    if (item < NITEMS) s_f[item] = nu[k];
  }
  __syncthreads();
}

struct Acc {
  int cells = 0, iters = 0, fails = 0, maxit = 0;
};

// One tile of one instance: net flux, stage update, cons2prim, store.
template <int ORDER, int TX, int TY, int STAGE>
__device__ __forceinline__ void batch_tile(const BatchArgs &a, double *s_f, double *s_acc, int t, Acc &acc) {
  using S = BatchShape<ORDER, TX, TY>;
  constexpr int R = S::R, NV = S::NV, WY = S::WY, HX = S::HX;
  const int b = a.tile_inst[t];
  if (!a.clock[b].active) return;  // the instance has reached its end time
  const InstDev I = a.inst[b];
  const double dt = a.clock[b].next_dt;
  const int local = t - I.tile_base;
  const int ty = local % I.tiles_y, tx = local / I.tiles_y;
  const int x0 = tx * TX, y0 = ty * TY;
  const long long vs = a.vs;
  // The thread index as a value the compiler cannot see through: otherwise
  // it hoists every tile-independent LDS offset of the face phases out of the
  // tile loop and keeps them in registers across the cons2prim.
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int r = tid / TY, c = tid % TY;  // the thread's cell in the tile
  __syncthreads();                       // the previous tile is done with LDS

  // ---- y faces: rows x0 .. x0+TX-1, cells y0-R .. y0+TY+R-1; cell idx = rr WY + cy.
  // The stencil of face lf of row rr (the left face of tile cell lf) starts
  // at cell lf of the row.
  constexpr int kFacesY = TX * (TY + 1);
  tile_faces<ORDER, 1, TX * WY, NV * kFacesY, 1>(
      a, s_f, tid,
      [&](int idx) -> long long {
        const int rr = idx / WY, cy = idx - rr * WY;
        const int gx = x0 + rr;
        if (gx >= I.nx) return -1;
        const int gy = remap(min(y0 + cy - R, I.ny - 1 + R), I.ny, a.py);
        return I.off + (long long)gx * I.ny + gy;
      },
      [&](int item) {
        const int v = item / kFacesY, idx = item - v * kFacesY;
        const int rr = idx / (TY + 1), lf = idx - rr * (TY + 1);
        return (v * TX + rr) * WY + lf;
      });
  // the y net flux waits in LDS for the x term (registers are the scarcer)
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double *g = s_f + (v * TX + r) * (TY + 1) + c;
    s_acc[v * kThreads + tid] = (g[1] - g[0]) * I.inv_dy;
  }
  __syncthreads();

  // ---- x faces: rows x0-R .. x0+TX+R-1, cells y0 .. y0+TY-1; cell idx = rx TY + cc
  constexpr int kFacesX = (TX + 1) * TY;
  tile_faces<ORDER, 0, HX * TY, NV * kFacesX, TY>(
      a, s_f, tid,
      [&](int idx) -> long long {
        const int rx = idx / TY, cc = idx - rx * TY;
        const int gy = y0 + cc;
        if (gy >= I.ny) return -1;
        const int gx = remap(min(x0 + rx - R, I.nx - 1 + R), I.nx, a.px);
        return I.off + (long long)gx * I.ny + gy;
      },
      [&](int item) {
        const int v = item / kFacesX, idx = item - v * kFacesX;
        const int lf = idx / TY, cc = idx - lf * TY;
        return (v * HX + lf) * TY + cc;
      });

  const int gx = x0 + r, gy = y0 + c;
  if (gx >= I.nx || gy >= I.ny) return;
  const long long q = I.off + (long long)gx * I.ny + gy;
  double un[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double *g = s_f + (v * (TX + 1) + r) * TY + c;
    const double fn = fma(g[TY] - g[0], I.inv_dx, s_acc[v * kThreads + tid]);
    un[v] = STAGE == 1 ? rk2_stage1(a.U[v * vs + q], fn, dt) : rk2_stage2(a.Un[v * vs + q], a.U[v * vs + q], fn, dt);
  }
  // TimeIntegrator::finalise's cons2prim (timeInt.h:117-132, srmhd.cc:390-562)
  double v2, Z;
  srmhd_plain_guess(a.Vg[0 * vs + q], a.Vg[1 * vs + q], a.Vg[2 * vs + q], a.Vg[3 * vs + q], a.Vg[4 * vs + q],
                    a.gamma, v2, Z);
  const double BS = un[5] * un[1] + un[6] * un[2] + un[7] * un[3];
  const double Bsq = un[5] * un[5] + un[6] * un[6] + un[7] * un[7];
  const double Ssq = un[1] * un[1] + un[2] * un[2] + un[3] * un[3];
  const int its = srmhd_newton(un[0], un[4], Bsq, Ssq, BS, a.gamma, v2, Z);
  double P[kSrmhdPrims], A[kSrmhdAux];  // A unused: the aux is formed from the root when read
  srmhd_recover(un, a.gamma, v2, Z, P, A);
  // streamed out with non-temporal stores: they do not fit in the caches and
  // would only evict the tiles' halo cells from the L2
#pragma unroll
  for (int v = 0; v < kSrmhdCons; ++v) __builtin_nontemporal_store(un[v], &a.Uout[v * vs + q]);
  // stage one: rho .. p of V1, all that stage two's sweeps, guess and retry read
#pragma unroll
  for (int v = 0; v < (STAGE == 1 ? 5 : kSrmhdPrims); ++v) __builtin_nontemporal_store(P[v], &a.Vout[v * vs + q]);
  if (STAGE == 2) {
    __builtin_nontemporal_store(v2, &a.root[q]);
    __builtin_nontemporal_store(Z, &a.root[vs + q]);
  }
  acc.cells += 1;
  if (its < 0) {
    acc.fails += 1;
    const int slot = atomicAdd(a.fail_count, 1);
    a.fail_list[slot] = q;
    a.fail_flag[q] = 1;
  } else {
    acc.iters += its;
    acc.maxit = max(acc.maxit, its);
  }
}

__device__ __forceinline__ void reduce_acc(C2PCounters *st, Acc acc) {
  __shared__ int s_red[kThreads / 64][4];
  for (int off = 32; off > 0; off >>= 1) {
    acc.cells += __shfl_xor(acc.cells, off);
    acc.iters += __shfl_xor(acc.iters, off);
    acc.fails += __shfl_xor(acc.fails, off);
    acc.maxit = max(acc.maxit, __shfl_xor(acc.maxit, off));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_red[w][0] = acc.cells;
    s_red[w][1] = acc.iters;
    s_red[w][2] = acc.fails;
    s_red[w][3] = acc.maxit;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    long long c = 0, it = 0, f = 0;
    int m = 0;
    for (int k = 0; k < kThreads / 64; ++k) {
      c += s_red[k][0];
      it += s_red[k][1];
      f += s_red[k][2];
      m = max(m, s_red[k][3]);
    }
    if (c > 0) {
      C2PCounters *bk = st + (blockIdx.x & (kStatBuckets - 1));
      atomicAdd(&bk->cells, (unsigned long long)c);
      atomicAdd(&bk->iterations, (unsigned long long)it);
      if (f) atomicAdd(&bk->failures, (unsigned long long)f);
      atomicMax(&bk->max_iterations, (unsigned long long)m);
    }
  }
}

// Retry of the stage's failed cells from the mean root of converged face
// neighbours (srmhd.cc:465-507), run by the last workgroup to finish, then the
// fail list is cleared. With quirks, the system solved is the instance's last
// cell's, as the reference's retry reuses the last cell's arguments.
__device__ void batch_retry(const BatchArgs &a, bool stage2) {
  const int nfail = __hip_atomic_load(a.fail_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const long long vs = a.vs;
  const double gam = a.gamma;
  for (int f = threadIdx.x; f < nfail; f += kThreads) {
    const long long q = a.fail_list[f];
    int lo = 0, hi = a.ninst - 1;  // the instance holding q: last with off <= q
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.inst[mid].off <= q) lo = mid;
      else hi = mid - 1;
    }
    const InstDev I = a.inst[lo];
    const long long loc = q - I.off;
    const int pos[2] = {(int)(loc / I.ny), (int)(loc % I.ny)};
    const int ext[2] = {I.nx, I.ny};
    const long long stride[2] = {I.ny, 1};
    double sv2 = 0.0, sZ = 0.0;
    int cnt = 0;
    for (int ax = 0; ax < 2; ++ax) {
      for (int s = -1; s <= 1; s += 2) {
        const int np = pos[ax] + s;
        if (np < 0 || np >= ext[ax] || ext[ax] == 1) continue;
        const long long m = q + s * stride[ax];
        if (a.fail_flag[m]) continue;
        double v2, Z;
        srmhd_plain_guess(a.Vout[m], a.Vout[vs + m], a.Vout[2 * vs + m], a.Vout[3 * vs + m], a.Vout[4 * vs + m], gam,
                          v2, Z);
        sv2 += v2;
        sZ += Z;
        ++cnt;
      }
    }
    double v2 = cnt ? sv2 / cnt : 0.0, Z = cnt ? sZ / cnt : 0.0;
    const long long src = a.quirks ? I.off + (long long)I.nx * I.ny - 1 : q;
    double u[9];
    for (int v = 0; v < 9; ++v) u[v] = a.Uout[v * vs + src];
    const double BS = u[5] * u[1] + u[6] * u[2] + u[7] * u[3];
    const double Bsq = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
    const double Ssq = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
    const int its = srmhd_newton(u[0], u[4], Bsq, Ssq, BS, gam, v2, Z);
    for (int v = 0; v < 9; ++v) u[v] = a.Uout[v * vs + q];
    double P[8], aux[13];
    srmhd_recover(u, gam, v2, Z, P, aux);
    for (int v = 0; v < 8; ++v) a.Vout[v * vs + q] = P[v];
    if (stage2) {
      a.root[q] = v2;
      a.root[vs + q] = Z;
    }
    if (its < 0) atomicAdd(&a.stats->retry_failures, 1ull);
  }
  __syncthreads();
  for (int f = threadIdx.x; f < nfail; f += kThreads) a.fail_flag[a.fail_list[f]] = 0;
}

__device__ __forceinline__ void xcd_range(int ntiles, int x, int &lo, int &hi) {
  lo = (int)((long long)ntiles * x / kXcds);
  hi = (int)((long long)ntiles * (x + 1) / kXcds);
}

// One RK2 stage of every active instance. Workgroup w runs on XCD w % 8; the
// tiles are cut into 8 contiguous ranges, one per XCD, so neighbouring tiles
// (which share halo cells) meet in the same L2. Static schedule: the XCD's
// workgroups stride through its range. Dynamic: they take tiles from the
// XCD's counter, then from the other XCDs' counters.
template <int ORDER, int TX, int TY, int STAGE>
__global__ __launch_bounds__(kThreads, 3) void k_batch_stage(BatchArgs a) {
  __shared__ double s_f[BatchShape<ORDER, TX, TY>::kLds];
  __shared__ double s_acc[kSrmhdCons * kThreads];
  __shared__ int s_tile, s_last;
  Acc acc;
  const int x = blockIdx.x % kXcds;
  if (a.sched == kSchedStatic) {
    int lo, hi;
    xcd_range(a.ntiles, x, lo, hi);
    const int nbx = ((int)gridDim.x - x + kXcds - 1) / kXcds;
    for (int t = lo + (int)blockIdx.x / kXcds; t < hi; t += nbx) batch_tile<ORDER, TX, TY, STAGE>(a, s_f, s_acc, t, acc);
  } else {
    for (int k = 0; k < kXcds; ++k) {
      const int xx = (x + k) % kXcds;
      int lo, hi;
      xcd_range(a.ntiles, xx, lo, hi);
      for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_tile = lo + atomicAdd(&a.sync[xx], 1);
        __syncthreads();
        const int t = s_tile;
        if (t >= hi) break;
        batch_tile<ORDER, TX, TY, STAGE>(a, s_f, s_acc, t, acc);
      }
    }
  }
  reduce_acc(a.stats, acc);

  // the last workgroup to finish sees every other one's stores (release
  // fence before its count, acquire fence after), retries failed cells and
  // resets the counters for the next launch
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(&a.sync[kXcds], 1) == (int)gridDim.x - 1;
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  batch_retry(a, STAGE == 2);
  __syncthreads();
  if (threadIdx.x == 0) {
    *a.fail_count = 0;
    for (int k = 0; k <= kXcds; ++k) a.sync[k] = 0;
  }
}

// Simulation::updateTime's bookkeeping (simulation.cc:113-132) per instance:
// with advance, the step just taken moves t by its dt and counts one
// iteration; then the instance is active while t < end_time, and its next dt
// is the CFL value with the slow-start factors, clamped to end at end_time.
__global__ void k_batch_clock(const InstDev *inst, Clock *clock, int ninst, int advance) {
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < ninst; b += gridDim.x * blockDim.x) {
    Clock c = clock[b];
    const InstDev I = inst[b];
    if (advance && c.active) {
      c.t = c.t + c.next_dt;
      c.dt = c.next_dt;
      c.iters += 1;
    }
    c.active = c.t < I.end_time;
    double dt = I.dt_base;
    if (c.iters < 15) dt *= 0.1;
    if (c.iters >= 15 && c.iters < 30) dt *= 0.25;
    if (c.t + dt > I.end_time) dt = I.end_time - c.t;
    c.next_dt = dt;
    clock[b] = c;
  }
}

// ------------------------------------------------------------- dispatch
struct TileShape {
  int tx, ty;
};
constexpr TileShape kShapes[] = {{16, 16}, {8, 32}, {32, 8}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

template <int ORDER, int TX, int TY, int STAGE>
const void *kernel_ptr() {
  return reinterpret_cast<const void *>(&k_batch_stage<ORDER, TX, TY, STAGE>);
}

template <int ORDER, int STAGE>
const void *stage_kernel_o(int shape) {
  switch (shape) {
    case 0: return kernel_ptr<ORDER, 16, 16, STAGE>();
    case 1: return kernel_ptr<ORDER, 8, 32, STAGE>();
    default: return kernel_ptr<ORDER, 32, 8, STAGE>();
  }
}

const void *stage_kernel(int order, int shape, int stage) {
  if (order == 3) return stage == 1 ? stage_kernel_o<3, 1>(shape) : stage_kernel_o<3, 2>(shape);
  return stage == 1 ? stage_kernel_o<5, 1>(shape) : stage_kernel_o<5, 2>(shape);
}

#define BATCH_OK(expr)                                                                               \
  do {                                                                                               \
    hipError_t _e = (expr);                                                                          \
    if (_e != hipSuccess) return engine_fail("%s failed: %s", #expr, hipGetErrorString(_e));          \
  } while (0)

}  // namespace

struct mgpu_batch {
  mgpu_batch_config cfg;
  int ninst = 0;
  std::vector<mgpu_instance> inst;
  std::vector<long long> off;  // first cell of each instance
  long long vs = 0;            // total cells
  long long max_cells = 0, max_pad = 0;
  int ncu = 0;
  hipStream_t stream = nullptr;
  double *U = nullptr, *V = nullptr, *A = nullptr, *U1 = nullptr, *V1 = nullptr;
  // Steps keep each cell's cons2prim root instead of writing its 13 aux
  // values; the aux is formed from it before anything reads A.
  double *root = nullptr;
  bool aux_stale = false;
  double *stage_pad = nullptr, *stage_cmp = nullptr;  // per-instance host transfer staging
  double *backup = nullptr;                           // autotuner's copy of U, V, A
  InstDev *dinst = nullptr;
  Clock *dclock = nullptr;
  Clock *dclock_backup = nullptr;
  int *dtile_inst = nullptr;
  int ntiles = 0;
  C2PCounters *dstats = nullptr;
  int *fail_count = nullptr;
  long long *fail_list = nullptr;
  unsigned char *fail_flag = nullptr;
  int *dsync = nullptr;
  mgpu_batch_tuning tune;
  int use_graph = 1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

int shape_index(int tx, int ty) {
  for (int s = 0; s < kNumShapes; ++s)
    if (kShapes[s].tx == tx && kShapes[s].ty == ty) return s;
  return -1;
}

void drop_graph(mgpu_batch *b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Tile list and per-instance descriptors for the current tile shape.
int build_tiles(mgpu_batch *b) {
  const int tx = b->tune.tile_x, ty = b->tune.tile_y;
  std::vector<InstDev> h(b->ninst);
  std::vector<int> tiles;
  for (int i = 0; i < b->ninst; ++i) {
    const mgpu_instance &in = b->inst[i];
    InstDev &d = h[i];
    d.nx = in.nx;
    d.ny = in.ny;
    d.off = b->off[i];
    // Data::initData / Simulation::Simulation (simData.cc:74-79, simulation.cc:34-36)
    const double dx = (in.xmax - in.xmin) / in.nx;
    const double dy = (in.ymax - in.ymin) / in.ny;
    d.inv_dx = 1.0 / dx;
    d.inv_dy = 1.0 / dy;
    // Simulation::updateTime (simulation.cc:113-118): 2-D, dz infinite, alpha = 1
    const double sd = std::sqrt(2.0);
    const double dtX = in.cfl * dx / (1.0 * sd);
    const double dtY = in.cfl * dy / (1.0 * sd);
    const double dtZ = in.cfl * INFINITY / (1.0 * sd);
    d.dt_base = (dtX <= dtY && dtX <= dtZ) ? dtX : ((dtY < dtZ) ? dtY : dtZ);
    d.end_time = in.end_time;
    const int tnx = (in.nx + tx - 1) / tx, tny = (in.ny + ty - 1) / ty;
    d.tile_base = (int)tiles.size();
    d.tiles_y = tny;
    for (int k = 0; k < tnx * tny; ++k) tiles.push_back(i);
  }
  if (b->dtile_inst) hipFree(b->dtile_inst);
  b->dtile_inst = nullptr;
  b->ntiles = (int)tiles.size();
  BATCH_OK(hipMalloc(&b->dtile_inst, tiles.size() * sizeof(int)));
  BATCH_OK(hipMemcpy(b->dtile_inst, tiles.data(), tiles.size() * sizeof(int), hipMemcpyHostToDevice));
  BATCH_OK(hipMemcpy(b->dinst, h.data(), h.size() * sizeof(InstDev), hipMemcpyHostToDevice));
  drop_graph(b);
  return 0;
}

BatchArgs stage_args(mgpu_batch *b, int stage) {
  BatchArgs a;
  memset(&a, 0, sizeof a);
  a.inst = b->dinst;
  a.clock = b->dclock;
  a.tile_inst = b->dtile_inst;
  a.ntiles = b->ntiles;
  a.ninst = b->ninst;
  a.vs = b->vs;
  a.px = b->cfg.bc[0] == MGPU_BC_PERIODIC;
  a.py = b->cfg.bc[1] == MGPU_BC_PERIODIC;
  a.gamma = b->cfg.gamma;
  a.stats = b->dstats;
  a.fail_count = b->fail_count;
  a.fail_list = b->fail_list;
  a.fail_flag = b->fail_flag;
  a.sync = b->dsync;
  a.quirks = b->cfg.quirks;
  a.sched = b->tune.schedule;
  a.Vg = b->V;
  if (stage == 1) {  // U1 = U - dt F(U)
    a.U = b->U;
    a.V = b->V;
    a.Uout = b->U1;
    a.Vout = b->V1;
  } else {  // U = (U + U1 - dt F(U1)) / 2
    a.U = b->U1;
    a.V = b->V1;
    a.Un = b->U;
    a.Uout = b->U;
    a.Vout = b->V;
    a.root = b->root;
  }
  return a;
}

int launch_clock(mgpu_batch *b, int advance) {
  hipLaunchKernelGGL(k_batch_clock, dim3((b->ninst + 63) / 64), dim3(64), 0, b->stream, b->dinst, b->dclock, b->ninst,
                     advance);
  BATCH_OK(hipGetLastError());
  return 0;
}

// RK2::step of every active instance followed by the clock update.
int enqueue_step(mgpu_batch *b) {
  for (int stage = 1; stage <= 2; ++stage) {
    BatchArgs a = stage_args(b, stage);
    void *args[] = {&a};
    const void *k = stage_kernel(b->cfg.weno_order, shape_index(b->tune.tile_x, b->tune.tile_y), stage);
    BATCH_OK(hipLaunchKernel(k, dim3(b->tune.blocks), dim3(kThreads), args, 0, b->stream));
  }
  return launch_clock(b, 1);
}

int ensure_graph(mgpu_batch *b) {
  if (b->exec) return 0;
  BATCH_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_step(b);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  BATCH_OK(e);
  b->graph = g;
  BATCH_OK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return 0;
}

int run_steps(mgpu_batch *b, int nsteps) {
  if (b->use_graph && ensure_graph(b)) return -1;
  for (int s = 0; s < nsteps; ++s) {
    if (b->use_graph) {
      BATCH_OK(hipGraphLaunch(b->exec, b->stream));
    } else if (enqueue_step(b)) {
      return -1;
    }
  }
  if (nsteps > 0) b->aux_stale = true;
  return 0;
}

// The aux of every instance, formed from the kept roots if steps left it.
int fresh_aux(mgpu_batch *b) {
  if (!b->aux_stale) return 0;
  Geo g;
  memset(&g, 0, sizeof g);
  g.vstride = b->vs;
  launch_aux_srmhd(g, b->U, b->root, b->A, b->cfg.gamma, b->stream);
  BATCH_OK(hipGetLastError());
  b->aux_stale = false;
  return 0;
}

int max_blocks_per_cu(mgpu_batch *b, int shape) {
  int n = 0;
  const void *k = stage_kernel(b->cfg.weno_order, shape, 2);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, kThreads, 0) != hipSuccess) return 1;
  return std::max(n, 1);
}

Geo instance_geo(const mgpu_batch *b, int i) {
  const mgpu_instance &in = b->inst[i];
  Geo g;
  memset(&g, 0, sizeof g);
  g.n[0] = in.nx;
  g.n[1] = in.ny;
  g.n[2] = 1;
  g.stride[2] = 1;
  g.stride[1] = 1;
  g.stride[0] = in.ny;
  g.vstride = (long long)in.nx * in.ny;
  g.dims = 2;
  g.periodic[0] = b->cfg.bc[0] == MGPU_BC_PERIODIC;
  g.periodic[1] = b->cfg.bc[1] == MGPU_BC_PERIODIC;
  g.periodic[2] = 1;
  g.inv_d[0] = 1.0 / ((in.xmax - in.xmin) / in.nx);
  g.inv_d[1] = 1.0 / ((in.ymax - in.ymin) / in.ny);
  return g;
}

PadGeo instance_pad(const mgpu_batch *b, int i) {
  const mgpu_instance &in = b->inst[i];
  PadGeo p;
  p.N[0] = in.nx + 2 * b->cfg.Ng;
  p.N[1] = in.ny + 2 * b->cfg.Ng;
  p.N[2] = 1;
  p.off[0] = b->cfg.Ng;
  p.off[1] = b->cfg.Ng;
  p.off[2] = 0;
  p.ntot = (long long)p.N[0] * p.N[1];
  return p;
}

// Copies nvar variable planes of instance i between the compact staging
// buffer (stride = the instance's cells) and a batch field (stride = all cells).
int move_planes(mgpu_batch *b, int i, int nvar, double *field, bool to_field) {
  const long long n = (long long)b->inst[i].nx * b->inst[i].ny;
  for (int v = 0; v < nvar; ++v) {
    double *f = field + v * b->vs + b->off[i];
    double *s = b->stage_cmp + v * n;
    BATCH_OK(hipMemcpyAsync(to_field ? f : s, to_field ? s : f, n * sizeof(double), hipMemcpyDeviceToDevice,
                            b->stream));
  }
  return 0;
}

int reset_clock(mgpu_batch *b, int i) {
  Clock c;
  memset(&c, 0, sizeof c);
  BATCH_OK(hipMemcpyAsync(b->dclock + i, &c, sizeof c, hipMemcpyHostToDevice, b->stream));
  return launch_clock(b, 0);
}

int check_index(const mgpu_batch *b, int i) {
  if (!b) return engine_fail("null batch");
  if (i < 0 || i >= b->ninst) return engine_fail("instance %d out of range [0, %d)", i, b->ninst);
  return 0;
}

int read_stats_batch(mgpu_batch *b, mgpu_stats *st) {
  C2PCounters h[kStatBuckets];
  BATCH_OK(hipMemcpyAsync(h, b->dstats, sizeof h, hipMemcpyDeviceToHost, b->stream));
  BATCH_OK(hipStreamSynchronize(b->stream));
  memset(st, 0, sizeof *st);
  for (int k = 0; k < kStatBuckets; ++k) {
    st->cells += (long long)h[k].cells;
    st->iterations += (long long)h[k].iterations;
    st->failures += (long long)h[k].failures;
    st->fallbacks += (long long)h[k].retry_failures;
    st->max_iterations = std::max(st->max_iterations, (long long)h[k].max_iterations);
  }
  return 0;
}

}  // namespace

extern "C" {

int mgpu_batch_create(const mgpu_batch_config *cfg, const mgpu_instance *inst, int ninst, mgpu_batch **out) {
  if (!cfg || !inst || !out || ninst <= 0) return engine_fail("mgpu_batch_create: bad arguments");
  *out = nullptr;
  const mgpu_batch_config &c = *cfg;
  if (c.weno_order != 3 && c.weno_order != 5) return engine_fail("mgpu_batch_create: weno_order must be 3 or 5");
  if (c.Ng < (c.weno_order + 1) / 2) return engine_fail("mgpu_batch_create: Ng too small for the WENO order");
  for (int ax = 0; ax < 2; ++ax)
    if (c.bc[ax] != MGPU_BC_PERIODIC && c.bc[ax] != MGPU_BC_OUTFLOW)
      return engine_fail("mgpu_batch_create: bad bc on axis %d", ax);
  for (int i = 0; i < ninst; ++i) {
    const mgpu_instance &in = inst[i];
    if (in.nx < c.Ng || in.ny < c.Ng) return engine_fail("mgpu_batch_create: instance %d has fewer than Ng cells", i);
    if (!(in.xmax > in.xmin) || !(in.ymax > in.ymin) || !(in.cfl > 0.0))
      return engine_fail("mgpu_batch_create: instance %d has a bad domain or cfl", i);
    // An instance that could never step would keep no cons2prim root, and the
    // aux formed from the roots after any batch step would be meaningless.
    if (!(in.end_time > 0.0) || !std::isfinite(in.end_time))
      return engine_fail("mgpu_batch_create: instance %d needs a finite end_time > 0", i);
  }
  BATCH_OK(hipSetDevice(c.device));
  mgpu_batch *b = new mgpu_batch();
  b->cfg = c;
  b->ninst = ninst;
  b->inst.assign(inst, inst + ninst);
  for (int i = 0; i < ninst; ++i) {
    const long long cells = (long long)inst[i].nx * inst[i].ny;
    b->off.push_back(b->vs);
    b->vs += cells;
    b->max_cells = std::max(b->max_cells, cells);
    b->max_pad = std::max(b->max_pad, (long long)(inst[i].nx + 2 * c.Ng) * (inst[i].ny + 2 * c.Ng));
  }
  auto cleanup = [&](int rc) {
    mgpu_batch_destroy(b);
    return rc;
  };
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c.device) != hipSuccess) return cleanup(engine_fail("device query failed"));
  b->ncu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(engine_fail("stream create failed"));
  const size_t n = (size_t)b->vs;
  struct {
    double **p;
    size_t elems;
  } bufs[] = {{&b->U, 9 * n},
              {&b->V, 8 * n},
              {&b->A, 13 * n},
              {&b->root, 2 * n},
              {&b->U1, 9 * n},
              {&b->V1, 8 * n},
              {&b->stage_pad, 30 * (size_t)b->max_pad},
              {&b->stage_cmp, 30 * (size_t)b->max_cells}};
  for (auto &bf : bufs) {
    if (hipMalloc(bf.p, bf.elems * sizeof(double)) != hipSuccess)
      return cleanup(engine_fail("hipMalloc of %zu bytes failed", bf.elems * sizeof(double)));
    if (hipMemset(*bf.p, 0, bf.elems * sizeof(double)) != hipSuccess) return cleanup(engine_fail("hipMemset failed"));
  }
  if (hipMalloc(&b->dinst, ninst * sizeof(InstDev)) != hipSuccess ||
      hipMalloc(&b->dclock, ninst * sizeof(Clock)) != hipSuccess ||
      hipMalloc(&b->dstats, kStatBuckets * sizeof(C2PCounters)) != hipSuccess ||
      hipMalloc(&b->fail_count, sizeof(int)) != hipSuccess || hipMalloc(&b->fail_list, n * sizeof(long long)) != hipSuccess ||
      hipMalloc(&b->fail_flag, n) != hipSuccess || hipMalloc(&b->dsync, 16 * sizeof(int)) != hipSuccess)
    return cleanup(engine_fail("hipMalloc of batch work arrays failed"));
  if (hipMemset(b->dclock, 0, ninst * sizeof(Clock)) != hipSuccess || hipMemset(b->fail_count, 0, sizeof(int)) != hipSuccess ||
      hipMemset(b->fail_flag, 0, n) != hipSuccess || hipMemset(b->dsync, 0, 16 * sizeof(int)) != hipSuccess ||
      hipMemset(b->dstats, 0, kStatBuckets * sizeof(C2PCounters)) != hipSuccess)
    return cleanup(engine_fail("hipMemset failed"));
  // default tuning until the autotuner runs: 16 x 16 tiles, static schedule,
  // as many workgroups as fit on the device at once
  b->tune.tile_x = 16;
  b->tune.tile_y = 16;
  b->tune.schedule = kSchedStatic;
  b->tune.blocks = b->ncu * max_blocks_per_cu(b, 0);
  b->tune.ms_per_step = 0.0;
  if (build_tiles(b)) return cleanup(-1);
  if (launch_clock(b, 0) || hipStreamSynchronize(b->stream) != hipSuccess) return cleanup(engine_fail("clock init failed"));
  *out = b;
  return 0;
}

int mgpu_batch_destroy(mgpu_batch *b) {
  if (!b) return 0;
  if (b->stream) hipStreamSynchronize(b->stream);
  drop_graph(b);
  double *bufs[] = {b->U, b->V, b->A, b->root, b->U1, b->V1, b->stage_pad, b->stage_cmp, b->backup};
  for (double *p : bufs)
    if (p) hipFree(p);
  void *other[] = {b->dinst, b->dclock, b->dclock_backup, b->dtile_inst, b->dstats, b->fail_count, b->fail_list,
                   b->fail_flag, b->dsync};
  for (void *p : other)
    if (p) hipFree(p);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int mgpu_batch_set_prims(mgpu_batch *b, int i, const double *prims) {
  if (check_index(b, i)) return -1;
  if (!prims) return engine_fail("mgpu_batch_set_prims: null prims");
  if (fresh_aux(b)) return -1;  // the other instances' aux first
  const Geo g = instance_geo(b, i);
  const PadGeo p = instance_pad(b, i);
  const long long n = g.vstride;
  BATCH_OK(hipMemcpyAsync(b->stage_pad, prims, 8 * p.ntot * sizeof(double), hipMemcpyHostToDevice, b->stream));
  // compact prims into planes 9..16 of the staging buffer, then primsToAll
  // writes cons into planes 0..8 and aux into 17..29 (Simulation::set,
  // simulation.cc:80-101)
  double *sc = b->stage_cmp, *sv = sc + 9 * n, *sa = sc + 17 * n;
  launch_compact(g, p, 8, b->stage_pad, sv, b->stream);
  launch_prims_to_all_srmhd(g, sv, sc, sa, b->cfg.gamma, b->stream);
  BATCH_OK(hipGetLastError());
  for (int v = 0; v < 9; ++v)
    BATCH_OK(hipMemcpyAsync(b->U + v * b->vs + b->off[i], sc + v * n, n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
  for (int v = 0; v < 8; ++v)
    BATCH_OK(hipMemcpyAsync(b->V + v * b->vs + b->off[i], sv + v * n, n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
  for (int v = 0; v < 13; ++v)
    BATCH_OK(hipMemcpyAsync(b->A + v * b->vs + b->off[i], sa + v * n, n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
  if (reset_clock(b, i)) return -1;
  BATCH_OK(hipStreamSynchronize(b->stream));
  return 0;
}

int mgpu_batch_upload_state(mgpu_batch *b, int i, const double *cons, const double *prims, const double *aux) {
  if (check_index(b, i)) return -1;
  if (!cons || !prims || !aux) return engine_fail("mgpu_batch_upload_state: null array");
  if (fresh_aux(b)) return -1;  // the other instances' aux first
  const Geo g = instance_geo(b, i);
  const PadGeo p = instance_pad(b, i);
  const double *src[3] = {cons, prims, aux};
  double *dst[3] = {b->U, b->V, b->A};
  const int nv[3] = {9, 8, 13};
  for (int f = 0; f < 3; ++f) {
    BATCH_OK(hipMemcpyAsync(b->stage_pad, src[f], nv[f] * p.ntot * sizeof(double), hipMemcpyHostToDevice, b->stream));
    launch_compact(g, p, nv[f], b->stage_pad, b->stage_cmp, b->stream);
    BATCH_OK(hipGetLastError());
    if (move_planes(b, i, nv[f], dst[f], true)) return -1;
  }
  if (reset_clock(b, i)) return -1;
  BATCH_OK(hipStreamSynchronize(b->stream));
  return 0;
}

int mgpu_batch_download_state(mgpu_batch *b, int i, double *cons, double *prims, double *aux) {
  if (check_index(b, i)) return -1;
  const Geo g = instance_geo(b, i);
  const PadGeo p = instance_pad(b, i);
  if (aux && fresh_aux(b)) return -1;
  double *dst[3] = {cons, prims, aux};
  double *src[3] = {b->U, b->V, b->A};
  const int nv[3] = {9, 8, 13};
  for (int f = 0; f < 3; ++f) {
    if (!dst[f]) continue;
    if (move_planes(b, i, nv[f], src[f], false)) return -1;
    launch_expand(g, p, nv[f], b->stage_cmp, b->stage_pad, 0, b->stream);
    BATCH_OK(hipGetLastError());
    BATCH_OK(hipMemcpyAsync(dst[f], b->stage_pad, nv[f] * p.ntot * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    BATCH_OK(hipStreamSynchronize(b->stream));
  }
  return 0;
}

int mgpu_batch_step(mgpu_batch *b, int nsteps, mgpu_stats *stats) {
  if (!b || nsteps < 0) return engine_fail("mgpu_batch_step: bad arguments");
  if (stats) BATCH_OK(hipMemsetAsync(b->dstats, 0, kStatBuckets * sizeof(C2PCounters), b->stream));
  if (run_steps(b, nsteps)) return -1;
  if (!stats) return 0;
  if (read_stats_batch(b, stats)) return -1;
  return stats->failures > 0 ? (int)std::min<long long>(stats->failures, 1 << 30) : 0;
}

int mgpu_batch_get_time(mgpu_batch *b, int i, double *t, double *dt, int *iters) {
  if (check_index(b, i)) return -1;
  Clock c;
  BATCH_OK(hipMemcpyAsync(&c, b->dclock + i, sizeof c, hipMemcpyDeviceToHost, b->stream));
  BATCH_OK(hipStreamSynchronize(b->stream));
  if (t) *t = c.t;
  if (dt) *dt = c.dt;
  if (iters) *iters = c.iters;
  return 0;
}

int mgpu_batch_use_graph(mgpu_batch *b, int on) {
  if (!b) return engine_fail("null batch");
  b->use_graph = on ? 1 : 0;
  return 0;
}

int mgpu_batch_get_tuning(const mgpu_batch *b, mgpu_batch_tuning *t) {
  if (!b || !t) return engine_fail("mgpu_batch_get_tuning: null argument");
  *t = b->tune;
  return 0;
}

int mgpu_batch_set_tuning(mgpu_batch *b, const mgpu_batch_tuning *t) {
  if (!b || !t) return engine_fail("mgpu_batch_set_tuning: null argument");
  const int s = shape_index(t->tile_x, t->tile_y);
  if (s < 0) return engine_fail("mgpu_batch_set_tuning: no %dx%d tile kernel", t->tile_x, t->tile_y);
  if (t->schedule != kSchedStatic && t->schedule != kSchedDynamic)
    return engine_fail("mgpu_batch_set_tuning: unknown schedule %d", t->schedule);
  if (t->blocks < kXcds || t->blocks > 64 * b->ncu)
    return engine_fail("mgpu_batch_set_tuning: blocks must be in [%d, %d]", kXcds, 64 * b->ncu);
  BATCH_OK(hipStreamSynchronize(b->stream));
  const bool shape_changed = t->tile_x != b->tune.tile_x || t->tile_y != b->tune.tile_y;
  b->tune = *t;
  if (shape_changed) return build_tiles(b);
  drop_graph(b);
  return 0;
}

int mgpu_batch_timed_steps(mgpu_batch *b, int nsteps, double *total_ms) {
  if (!b || nsteps <= 0 || !total_ms) return engine_fail("mgpu_batch_timed_steps: bad arguments");
  if (b->use_graph && ensure_graph(b)) return -1;
  hipEvent_t e0, e1;
  BATCH_OK(hipEventCreate(&e0));
  BATCH_OK(hipEventCreate(&e1));
  int rc = 0;
  if (hipEventRecord(e0, b->stream) != hipSuccess) rc = engine_fail("event record failed");
  if (rc == 0) rc = run_steps(b, nsteps);
  if (rc == 0 && hipEventRecord(e1, b->stream) != hipSuccess) rc = engine_fail("event record failed");
  if (rc == 0 && hipEventSynchronize(e1) != hipSuccess) rc = engine_fail("event sync failed");
  float ms = 0.f;
  if (rc == 0) hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *total_ms = ms;
  return rc;
}

int mgpu_batch_autotune(mgpu_batch *b, int trial_steps, mgpu_batch_tuning *best, mgpu_batch_tuning *all, int max_all,
                        int *nall) {
  if (!b || trial_steps <= 0) return engine_fail("mgpu_batch_autotune: bad arguments");
  // the trials step the batch: keep its state and clocks and put them back
  const size_t n = (size_t)b->vs;
  if (!b->backup) BATCH_OK(hipMalloc(&b->backup, 30 * n * sizeof(double)));
  if (!b->dclock_backup) BATCH_OK(hipMalloc(&b->dclock_backup, b->ninst * sizeof(Clock)));
  if (fresh_aux(b)) return -1;
  BATCH_OK(hipMemcpyAsync(b->backup, b->U, 9 * n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
  BATCH_OK(hipMemcpyAsync(b->backup + 9 * n, b->V, 8 * n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
  BATCH_OK(hipMemcpyAsync(b->backup + 17 * n, b->A, 13 * n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
  BATCH_OK(hipMemcpyAsync(b->dclock_backup, b->dclock, b->ninst * sizeof(Clock), hipMemcpyDeviceToDevice, b->stream));
  const mgpu_batch_tuning start = b->tune;
  mgpu_batch_tuning win = start;
  win.ms_per_step = 1e300;
  int count = 0;
  int rc = 0;
  for (int s = 0; s < kNumShapes && rc == 0; ++s) {
    const int maxb = max_blocks_per_cu(b, s);
    for (int sched = kSchedStatic; sched <= kSchedDynamic && rc == 0; ++sched) {
      for (int bpc = 1; bpc <= maxb && rc == 0; ++bpc) {
        mgpu_batch_tuning cand;
        cand.tile_x = kShapes[s].tx;
        cand.tile_y = kShapes[s].ty;
        cand.schedule = sched;
        cand.blocks = b->ncu * bpc;
        cand.ms_per_step = 0.0;
        double ms = 0.0;
        rc = mgpu_batch_set_tuning(b, &cand);
        if (rc == 0) rc = run_steps(b, 1);  // warm-up (and graph build)
        if (rc == 0) rc = mgpu_batch_timed_steps(b, trial_steps, &ms);
        cand.ms_per_step = ms / trial_steps;
        if (rc == 0 && all && count < max_all) all[count] = cand;
        ++count;
        if (rc == 0 && cand.ms_per_step < win.ms_per_step) win = cand;
      }
    }
  }
  if (rc == 0) rc = mgpu_batch_set_tuning(b, &win);
  else mgpu_batch_set_tuning(b, &start);
  BATCH_OK(hipMemcpyAsync(b->U, b->backup, 9 * n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
  BATCH_OK(hipMemcpyAsync(b->V, b->backup + 9 * n, 8 * n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
  BATCH_OK(hipMemcpyAsync(b->A, b->backup + 17 * n, 13 * n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
  BATCH_OK(hipMemcpyAsync(b->dclock, b->dclock_backup, b->ninst * sizeof(Clock), hipMemcpyDeviceToDevice, b->stream));
  BATCH_OK(hipStreamSynchronize(b->stream));
  b->aux_stale = false;  // the restored aux belongs to the restored state
  if (nall) *nall = count;
  if (best) *best = b->tune;
  return rc;
}

int mgpu_batch_sync(mgpu_batch *b) {
  if (!b) return engine_fail("null batch");
  BATCH_OK(hipStreamSynchronize(b->stream));
  return 0;
}

int mgpu_batch_sizes(const mgpu_batch *b, int *ninst, long long *cells, int *ntiles) {
  if (!b) return engine_fail("null batch");
  if (ninst) *ninst = b->ninst;
  if (cells) *cells = b->vs;
  if (ntiles) *ntiles = b->ntiles;
  return 0;
}

}  // extern "C"
