// HIP kernels of the SRMHD time step on gfx950 (FP64).
//
// One RK2 stage (RK2.cc:31-91) is two or three directional sweep launches.
// Each sweep stages a ghost-padded tile of Lax-Friedrichs split fluxes in LDS
// (FVS::fluxReconstruction, fluxVectorSplitting.cc:4-56), reconstructs every
// face of the tile with WENO (weno.cc:123-229) and differences neighbouring
// faces into the cell's net flux (FVS::F, fluxVectorSplitting.cc:79-112). The
// first sweep writes the net flux, later ones add to it, and the last one also
// applies the stage update and the per-cell Newton cons2prim, so the stage's
// new cons and prims leave the kernel without another pass over HBM.
#include "mgpu_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace mgpu {

namespace {

constexpr int kThreads = 256;

// Block-level reduction of C2P counters into one of the stats buckets.
__device__ __forceinline__ void reduce_stats(C2PCounters *st, int cells, int iters, int fails, int maxit) {
  if (!st) return;
  __shared__ int s_red[kThreads / 64][4];
  for (int off = 32; off > 0; off >>= 1) {
    cells += __shfl_xor(cells, off);
    iters += __shfl_xor(iters, off);
    fails += __shfl_xor(fails, off);
    maxit = max(maxit, __shfl_xor(maxit, off));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_red[w][0] = cells;
    s_red[w][1] = iters;
    s_red[w][2] = fails;
    s_red[w][3] = maxit;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int c = 0, it = 0, f = 0, m = 0;
    for (int k = 0; k < kThreads / 64; ++k) {
      c += s_red[k][0];
      it += s_red[k][1];
      f += s_red[k][2];
      m = max(m, s_red[k][3]);
    }
    if (c > 0) {
      C2PCounters *b = st + (blockIdx.x & (kStatBuckets - 1));
      atomicAdd(&b->cells, (unsigned long long)c);
      atomicAdd(&b->iterations, (unsigned long long)it);
      if (f) atomicAdd(&b->failures, (unsigned long long)f);
      atomicMax(&b->max_iterations, (unsigned long long)m);
    }
  }
}

// Model traits used by the sweep: conserved count, which primitives the flux
// needs and the flux itself.
struct SrmhdM {
  static constexpr int NC = kSrmhdCons, NP = kSrmhdPrims, NA = kSrmhdAux;
  static constexpr int P0 = 1, NPL = 4;  // vx, vy, vz, p
  static constexpr bool kStored = false;
  template <int DIR, bool Q>
  __device__ __forceinline__ static void flux(const double *u, const double *pl, const SweepArgs &, long long,
                                              double *f) {
    srmhd_flux<DIR, Q>(u, pl[0], pl[1], pl[2], pl[3], f);
  }
};

struct SrrmhdM {
  static constexpr int NC = kSrrmhdCons, NP = kSrrmhdPrims, NA = kSrrmhdAux;
  static constexpr int P0 = 0, NPL = 5;  // rho, vx, vy, vz, p
  static constexpr bool kStored = false;
  template <int DIR, bool Q>
  __device__ __forceinline__ static void flux(const double *u, const double *pl, const SweepArgs &a, long long,
                                              double *f) {
    srrmhd_flux<DIR, Q>(u, pl[0], pl[1], pl[2], pl[3], pl[4], a.gamma, a.sigma, f);
  }
};

// Model M with the flux vectors read from a.Fin (written by k_flux_stored).
template <class M>
struct StoredM {
  static constexpr int NC = M::NC, NP = M::NP, NA = M::NA;
  static constexpr int P0 = 0, NPL = 1;
  static constexpr bool kStored = true;
  template <int DIR, bool Q>
  __device__ __forceinline__ static void flux(const double *, const double *, const SweepArgs &a, long long q,
                                              double *f) {
    const long long vs = a.geo.vstride;
    const double *F = a.Fin + (long long)DIR * NC * vs;
#pragma unroll
    for (int v = 0; v < NC; ++v) f[v] = F[v * vs + q];
  }
};

// SRMHD::fluxVector from stored arrays (srmhd.cc:72-206): B, W and b^mu come
// from the prims (the aux the reference reads is their function), D, S, tau
// and phi from the cons.
template <int DIR, bool QUIRK>
__device__ __forceinline__ void srmhd_flux_stored(const double u[9], const double P[8], double f[9]) {
  double uu[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) uu[v] = u[v];
  uu[5] = P[5];
  uu[6] = P[6];
  uu[7] = P[7];
  srmhd_flux<DIR, QUIRK>(uu, P[1], P[2], P[3], P[4], f);
}

template <int DIR, bool QUIRK>
__device__ __forceinline__ void flux_stored_cell(int model, const double *U, const double *V, const double *A,
                                                 long long vs, long long q, double *F) {
  double u[kSrrmhdCons], P[kSrrmhdPrims], aux[kSrrmhdAux], f[kSrrmhdCons];
  if (model == 0) {
    for (int v = 0; v < kSrmhdCons; ++v) u[v] = U[v * vs + q];
    for (int v = 0; v < kSrmhdPrims; ++v) P[v] = V[v * vs + q];
    srmhd_flux_stored<DIR, QUIRK>(u, P, f);
    for (int v = 0; v < kSrmhdCons; ++v) F[((long long)DIR * kSrmhdCons + v) * vs + q] = f[v];
  } else {
    for (int v = 0; v < kSrrmhdCons; ++v) u[v] = U[v * vs + q];
    for (int v = 0; v < kSrrmhdPrims; ++v) P[v] = V[v * vs + q];
    for (int v = 0; v < kSrrmhdAux; ++v) aux[v] = A[v * vs + q];
    srrmhd_flux_stored<DIR, QUIRK>(u, P, aux, f);
    for (int v = 0; v < kSrrmhdCons; ++v) F[((long long)DIR * kSrrmhdCons + v) * vs + q] = f[v];
  }
}

__global__ __launch_bounds__(kThreads) void k_flux_stored(Geo g, int model, const double *U, const double *V,
                                                          const double *A, double *F, int quirk) {
  const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (q >= g.vstride) return;
  flux_stored_cell<0, false>(model, U, V, A, g.vstride, q, F);
  if (g.dims > 1) flux_stored_cell<1, false>(model, U, V, A, g.vstride, q, F);
  if (g.dims > 2) {
    if (quirk)
      flux_stored_cell<2, true>(model, U, V, A, g.vstride, q, F);
    else
      flux_stored_cell<2, false>(model, U, V, A, g.vstride, q, F);
  }
}

template <class M, int ORDER, int DAX, int CAX, bool DFAST, int TL_ = (DFAST ? 32 : 16), int TC_ = (DFAST ? 8 : 16),
          bool SPLIT_ALL = false>
struct SweepShape {
  static constexpr int R = Weno<ORDER>::R;
  static constexpr int TL = TL_;  // cells along the sweep axis
  static constexpr int TC = TC_;  // pencils along the column axis
  static constexpr int TLH = TL + 2 * R;      // with halo
  static constexpr int EAX = 3 - DAX - CAX;
  static constexpr int kFaces = (TL + 1) * TC;
  // SRMHD stages f+ and then f- through LDS (kSplit), which halves the LDS.
  // The resistive model's 14 variables keep both unless SPLIT_ALL asks for
  // the split: where the threads take the tile's cells in two passes the
  // kept f- needs 28 more registers, and the fused y+z sweep with 8-pencil
  // groups measured 1-9 % slower with it.
  static constexpr bool kSplit = M::NC <= 9 || SPLIT_ALL;
  static constexpr int kLds = (kSplit ? 1 : 2) * M::NC * TLH * TC;
  // LDS position of (variable slot v, cell l along the sweep, pencil c): the
  // unit-stride index follows the threads' fast index, which is l when the
  // sweep axis is the contiguous one (DFAST) and the pencil index otherwise.
  __device__ static constexpr int fidx(int v, int l, int c) {
    return DFAST ? (v * TC + c) * TLH + l : (v * TLH + l) * TC + c;
  }
  // Same for the face fluxes (TL + 1 faces per pencil).
  __device__ static constexpr int gidx(int v, int lf, int c) {
    return DFAST ? (v * TC + c) * (TL + 1) + lf : (v * (TL + 1) + lf) * TC + c;
  }
  static constexpr int kStrideL = DFAST ? 1 : TC;  // LDS step between neighbouring cells of a pencil
  // fidx(v, lf, cc) of work item v * kFaces + idx, idx enumerating (lf, cc) in
  // gidx order: for the strided layout idx = lf * TC + cc already is the
  // offset within the variable; for the contiguous one each pencil skips the
  // 2R - 1 halo slots past its TL + 1 faces.
  __device__ static unsigned item_fidx(unsigned item) {
    const unsigned v = item / kFaces;
    unsigned off = item + v * (unsigned)(TLH * TC - kFaces);
    if (DFAST) off += (item - v * kFaces) / (TL + 1) * (unsigned)(TLH - TL - 1);
    return off;
  }
};

// Counters of one cell's cons2prim, accumulated over a block.
struct CellC2P {
  int cells = 0, iters = 0, fails = 0, maxit = 0;
  __device__ void add(int its) {
    cells = 1;
    if (its < 0) {
      fails = 1;
    } else {
      iters = its;
      maxit = its;
    }
  }
};

// RK2 stage end for SRMHD: update, Newton C2P, write.
__device__ __forceinline__ void finish_rk2(const SweepArgs &a, long long q, const double *un, SrmhdM,
                                           CellC2P &cs) {
  const long long vs = a.geo.vstride;
  double v2, Z;
  // the guess is read once per cell: non-temporal loads keep it from evicting
  // the tiles' halo cells from the L2
  srmhd_plain_guess(__builtin_nontemporal_load(&a.Vg[0 * vs + q]), __builtin_nontemporal_load(&a.Vg[1 * vs + q]),
                    __builtin_nontemporal_load(&a.Vg[2 * vs + q]), __builtin_nontemporal_load(&a.Vg[3 * vs + q]),
                    __builtin_nontemporal_load(&a.Vg[4 * vs + q]), a.gamma, v2, Z);
  const double BS = un[5] * un[1] + un[6] * un[2] + un[7] * un[3];
  const double Bsq = un[5] * un[5] + un[6] * un[6] + un[7] * un[7];
  const double Ssq = un[1] * un[1] + un[2] * un[2] + un[3] * un[3];
  const int its = srmhd_newton(un[0], un[4], Bsq, Ssq, BS, a.gamma, v2, Z);
  double P[kSrmhdPrims], A[kSrmhdAux];
  srmhd_recover(un, a.gamma, v2, Z, P, A);
  // the stage's outputs stream out (non-temporal stores: the next launch
  // reads them, but they do not fit in the caches and would only evict the
  // tiles' halo cells from the L2)
#pragma unroll
  for (int v = 0; v < kSrmhdCons; ++v) __builtin_nontemporal_store(un[v], &a.Uout[v * vs + q]);
  const int nvout = a.vout_n ? a.vout_n : kSrmhdPrims;
#pragma unroll
  for (int v = 0; v < kSrmhdPrims; ++v)
    if (v < nvout) __builtin_nontemporal_store(P[v], &a.Vout[v * vs + q]);
  if (a.Aout) {
#pragma unroll
    for (int v = 0; v < kSrmhdAux; ++v) a.Aout[v * vs + q] = A[v];
  } else if (a.root) {
    __builtin_nontemporal_store(v2, &a.root[q]);
    __builtin_nontemporal_store(Z, &a.root[vs + q]);
  }
  cs.add(its);
  if (its < 0) {
    const int slot = atomicAdd(a.fail_count, 1);
    a.fail_list[slot] = q;
    a.fail_flag[q] = 1;
  }
}

// Cons2prim of a resistive stage value with the secant rule, writing cons,
// prims and (if requested) aux; returns the recovered aux in A.
__device__ __forceinline__ void c2p_store_srrmhd(const SweepArgs &a, long long q, const double *un, double *A,
                                                 CellC2P &cs) {
  const long long vs = a.geo.vstride;
  const double Zg = srrmhd_guess_Z(a.Vg[0 * vs + q], a.Vg[1 * vs + q], a.Vg[2 * vs + q], a.Vg[3 * vs + q],
                                   a.Vg[4 * vs + q], a.gamma);
  double P[kSrrmhdPrims];
  const int its = srrmhd_c2p_cell(un, a.gamma, a.sigma, Zg, P, A);
  // streamed out, as finish_rk2 does for SRMHD
#pragma unroll
  for (int v = 0; v < kSrrmhdCons; ++v) __builtin_nontemporal_store(un[v], &a.Uout[v * vs + q]);
  const int nvout = a.vout_n ? a.vout_n : kSrrmhdPrims;
#pragma unroll
  for (int v = 0; v < kSrrmhdPrims; ++v)
    if (v < nvout) __builtin_nontemporal_store(P[v], &a.Vout[v * vs + q]);
  if (a.Aout) {
#pragma unroll
    for (int v = 0; v < kSrrmhdAux; ++v) a.Aout[v * vs + q] = A[v];
  } else if (a.root) {
    __builtin_nontemporal_store(A[10], &a.root[q]);  // Z: the aux follows from it and the cons (srrmhd_recover)
  }
  cs.add(its);
}

__device__ __forceinline__ void finish_rk2(const SweepArgs &a, long long q, const double *un, SrrmhdM,
                                           CellC2P &cs) {
  double A[kSrrmhdAux];
  c2p_store_srrmhd(a, q, un, A, cs);
}

// SSP2 stage two at one cell once flux1 is complete (SSP2.cc:112-148): solve
// from c = cons - dt (flux1 - om2gam source1), cons2prim, and store the part
// of the final combination known now, U^n - dt/2 flux1 + dt/2 (source1 +
// source2), so the final sweep reads one array instead of U^n, flux1 and
// both sources.
__device__ __forceinline__ void imex2_apply(const SweepArgs &a, long long q, const double *fn, CellC2P &cs) {
  const long long vs = a.geo.vstride;
  double c[kSrrmhdCons], pf[kSrrmhdCons], x[kSrrmhdCons], A[kSrrmhdAux], s1[5];
#pragma unroll
  for (int v = 0; v < kSrrmhdCons; ++v) {
    const double un = a.Un[v * vs + q];
    c[v] = un - a.dt * fn[v];
    pf[v] = un - 0.5 * a.dt * fn[v];
  }
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    s1[m] = a.S1[m * vs + q];
    c[8 + m] += a.dt * kSsp2Om2Gam * s1[m];
  }
  const double Zg = srrmhd_guess_Z(a.Vg[0 * vs + q], a.Vg[1 * vs + q], a.Vg[2 * vs + q], a.Vg[3 * vs + q],
                                   a.Vg[4 * vs + q], a.gamma);
  const int its = imex_stage(c, a.dt, a.gamma, a.sigma, a.cp, Zg, x);
  c2p_store_srrmhd(a, q, x, A, cs);
  if (its < 0) cs.fails = 1;
  double s5[5];
  srrmhd_source5(x, A, a.cp, s5);
#pragma unroll
  for (int m = 0; m < 5; ++m) pf[8 + m] += 0.5 * a.dt * (s1[m] + s5[m]);
#pragma unroll
  for (int v = 0; v < kSrrmhdCons; ++v) __builtin_nontemporal_store(pf[v], &a.Pf[v * vs + q]);
}

// SSP2's final combination at one cell, cons += -dt/2 (flux1 + flux2 -
// source1 - source2) (SSP2.cc:156-166), from the part stage two stored, then
// cons2prim.
__device__ __forceinline__ void imex_final_apply(const SweepArgs &a, long long q, const double *fn, CellC2P &cs) {
  const long long vs = a.geo.vstride;
  double un[kSrrmhdCons], A[kSrrmhdAux];
#pragma unroll
  for (int v = 0; v < kSrrmhdCons; ++v) un[v] = a.Pf[v * vs + q] - 0.5 * a.dt * fn[v];
  c2p_store_srrmhd(a, q, un, A, cs);
}

// Stored-flux sweeps only accumulate the net flux (write and add modes).
template <class M>
__device__ __forceinline__ void finish_rk2(const SweepArgs &, long long, const double *, StoredM<M>, CellC2P &) {}

// What a sweep does with a cell's net flux fn once its own axis is added
// (the sweep's mode): keep it for the next sweep, or finish the stage there.
template <class M, int MODE>
__device__ __forceinline__ void sweep_apply(const SweepArgs &a, long long q, const double *fn, CellC2P &cs) {
  constexpr int NV = M::NC;
  const long long vs = a.geo.vstride;
  if (MODE == kSweepWrite || MODE == kSweepAdd) {
#pragma unroll
    for (int v = 0; v < NV; ++v) a.fnet[v * vs + q] = fn[v];
  } else if (MODE == kSweepStage1 || MODE == kSweepStage2) {
    double un[NV];
    if (MODE == kSweepStage1) {
#pragma unroll
      for (int v = 0; v < NV; ++v) un[v] = rk2_stage1(a.U[v * vs + q], fn[v], a.dt);
    } else {
#pragma unroll
      for (int v = 0; v < NV; ++v)
        un[v] = rk2_stage2(__builtin_nontemporal_load(&a.Un[v * vs + q]), a.U[v * vs + q], fn[v], a.dt);
    }
    finish_rk2(a, q, un, M(), cs);
  } else if (MODE == kSweepImex2) {
    imex2_apply(a, q, fn, cs);
  } else if (MODE == kSweepImexFinal) {
    imex_final_apply(a, q, fn, cs);
  }
}

// Phases 1 and 2 of a directional sweep on one tile: the split fluxes of the
// tile's cells and their halo along DAX into LDS, then the WENO value of
// every face of the tile. On return (after a barrier) s_f holds the face
// values in S::gidx order: face lf of pencil cc is the left face of tile
// cell lf. d0 and c0 are the tile's first cell along DAX and CAX, base_e the
// offset of its plane along the third axis, nd and nc the grid's extents,
// tid the thread's index in the workgroup.
template <class M, int ORDER, int DAX, int CAX, bool DFAST, int TL_, int TC_, bool QUIRK, bool SPLIT_ALL = false,
          int NLDS>
__device__ __forceinline__ void sweep_tile_faces(const SweepArgs &a, double (&s_f)[NLDS], int d0, int c0,
                                                 long long base_e, int nd, int nc, unsigned tid) {
  using S = SweepShape<M, ORDER, DAX, CAX, DFAST, TL_, TC_, SPLIT_ALL>;
  constexpr int R = S::R, TC = S::TC, TLH = S::TLH;
  constexpr int NV = M::NC;
  const Geo &g = a.geo;
  const long long vs = g.vstride;

  if constexpr (S::kSplit) {
    // Phase 1: split fluxes of every tile cell and its halo along DAX. f+ goes
    // to LDS now; each thread keeps its cells' f- in registers until the upwind
    // half of every face is done, then f- takes f+'s place. LDS holds one of
    // the two at a time, which halves it and lets more blocks share a CU.
    constexpr int kCells = TLH * TC;
    constexpr int kCellPasses = (kCells + kThreads - 1) / kThreads;
    double fm_keep[kCellPasses][NV];
  #pragma unroll
    for (int cp = 0; cp < kCellPasses; ++cp) {
      const unsigned idx = tid + cp * kThreads;
      if (idx < kCells) {
        const int l = DFAST ? idx % TLH : idx / TC;
        const int cc = DFAST ? idx / TLH : idx % TC;
        const int gc = c0 + cc;
        double fp[NV];
        if (gc < nc) {
          // cells past nd + R - 1 only feed faces outside the grid; clamping
          // them keeps the periodic image inside [0, nd) for any tile overhang
          const int gd = remap(min(d0 + l - R, nd - 1 + R), nd, g.periodic[DAX]);
          const long long q = base_e + (long long)gd * g.stride[DAX] + (long long)gc * g.stride[CAX];
          double u[NV], pl[M::NPL];
  #pragma unroll
          for (int v = 0; v < NV; ++v) u[v] = a.U[v * vs + q];
  #pragma unroll
          for (int v = 0; v < M::NPL; ++v) pl[v] = a.V[(M::P0 + v) * vs + q];
          double f[NV];
          M::template flux<DAX, QUIRK>(u, pl, a, q, f);
  #pragma unroll
          for (int v = 0; v < NV; ++v) {
            fp[v] = 0.5 * (f[v] + u[v]);
            fm_keep[cp][v] = 0.5 * (f[v] - u[v]);
          }
        } else {
  #pragma unroll
          for (int v = 0; v < NV; ++v) fp[v] = fm_keep[cp][v] = 0.0;
        }
  #pragma unroll
        for (int v = 0; v < NV; ++v) s_f[S::fidx(v, l, cc)] = fp[v];
      }
    }
    // likewise the kept f- values are formed before the barrier
  #pragma unroll
    for (int cp = 0; cp < kCellPasses; ++cp)
  #pragma unroll
      for (int v = 0; v < NV; ++v) asm volatile("" : "+v"(fm_keep[cp][v]));
    __syncthreads();

    // Phase 2: WENO face values, face lf being the left face of tile cell lf.
    // The (variable, face) pairs are dealt out one per thread per pass, so the
    // faces beyond one per thread cost a fraction of a pass spread over the
    // waves instead of a full pass of all variables on one wave. Upwind halves
    // from f+ first, then downwind halves from f-.
    constexpr int kItems = NV * S::kFaces;
    constexpr int kPasses = (kItems + kThreads - 1) / kThreads;
    double nu[kPasses], du[kPasses];
  #pragma unroll
    for (int k = 0; k < kPasses; ++k) {
      const unsigned item = tid + k * kThreads;
      if (item < kItems) {
        const double *fpv = s_f + S::item_fidx(item);
        Weno<ORDER>::up([&](int m) { return fpv[m * S::kStrideL]; }, nu[k], du[k]);
      }
    }
    // the upwind halves are formed here, before the barrier after which f-
    // overwrites f+: left to itself the compiler keeps the loaded f+ stencils
    // alive across it and forms them later, which doubles the registers
#pragma unroll
    for (int k = 0; k < kPasses; ++k) asm volatile("" : "+v"(nu[k]), "+v"(du[k]));
    __syncthreads();
  #pragma unroll
    for (int cp = 0; cp < kCellPasses; ++cp) {
      const unsigned idx = tid + cp * kThreads;
      if (idx < kCells) {
        const int l = DFAST ? idx % TLH : idx / TC;
        const int cc = DFAST ? idx / TLH : idx % TC;
  #pragma unroll
        for (int v = 0; v < NV; ++v) s_f[S::fidx(v, l, cc)] = fm_keep[cp][v];
      }
    }
    __syncthreads();
  #pragma unroll
    for (int k = 0; k < kPasses; ++k) {
      const unsigned item = tid + k * kThreads;
      if (item < kItems) {
        const double *fmv = s_f + S::item_fidx(item);
        double nd, dd;
        Weno<ORDER>::down([&](int m) { return fmv[m * S::kStrideL]; }, nd, dd);
        nu[k] = weno_combine(nu[k], du[k], nd, dd);
      }
    }
#pragma unroll
    for (int k = 0; k < kPasses; ++k) asm volatile("" : "+v"(nu[k]));
    __syncthreads();
    // item = v * kFaces + idx, idx enumerating (lf, cc) in gidx order for both
    // layouts, so the face of item goes to s_f[item]
  #pragma unroll
    for (int k = 0; k < kPasses; ++k) {
      const unsigned item = tid + k * kThreads;
      if (item < kItems) s_f[item] = nu[k];
    }
    __syncthreads();
  } else {
    // Phase 1: f+ and f- of every tile cell and its halo side by side in LDS.
    for (unsigned idx = tid; idx < TLH * TC; idx += kThreads) {
      const int l = DFAST ? idx % TLH : idx / TC;
      const int cc = DFAST ? idx / TLH : idx % TC;
      const int gc = c0 + cc;
      double fp[NV], fm[NV];
      if (gc < nc) {
        const int gd = remap(min(d0 + l - R, nd - 1 + R), nd, g.periodic[DAX]);
        const long long q = base_e + (long long)gd * g.stride[DAX] + (long long)gc * g.stride[CAX];
        double u[NV], pl[M::NPL];
#pragma unroll
        for (int v = 0; v < NV; ++v) u[v] = a.U[v * vs + q];
#pragma unroll
        for (int v = 0; v < M::NPL; ++v) pl[v] = a.V[(M::P0 + v) * vs + q];
        double f[NV];
        M::template flux<DAX, QUIRK>(u, pl, a, q, f);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          fp[v] = 0.5 * (f[v] + u[v]);
          fm[v] = 0.5 * (f[v] - u[v]);
        }
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) fp[v] = fm[v] = 0.0;
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        s_f[S::fidx(v, l, cc)] = fp[v];
        s_f[S::fidx(NV + v, l, cc)] = fm[v];
      }
    }
    __syncthreads();
    constexpr int kItems = NV * S::kFaces;
    constexpr int kPasses = (kItems + kThreads - 1) / kThreads;
    double face[kPasses];
#pragma unroll
    for (int k = 0; k < kPasses; ++k) {
      const unsigned item = tid + k * kThreads;
      if (item < kItems) {
        const double *fpv = s_f + S::item_fidx(item);
        const double *fmv = fpv + NV * TLH * TC;
        face[k] = Weno<ORDER>::face([&](int m) { return fpv[m * S::kStrideL]; },
                                    [&](int m) { return fmv[m * S::kStrideL]; });
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPasses; ++k) {
      const unsigned item = tid + k * kThreads;
      if (item < kItems) s_f[item] = face[k];
    }
    __syncthreads();
  }

}

// One directional sweep of the flux divergence of model M.
template <class M, int ORDER, int DAX, int CAX, bool DFAST, int MODE, bool QUIRK>
__global__ __launch_bounds__(kThreads) void k_sweep(SweepArgs a) {
  using S = SweepShape<M, ORDER, DAX, CAX, DFAST>;
  constexpr int R = S::R, TL = S::TL, TC = S::TC, TLH = S::TLH, EAX = S::EAX;
  static_assert(TL * TC == kThreads, "one cell per thread in the difference phase");
  constexpr int NV = M::NC;
  // f+ (then f-) of each variable [NV][TLH][TC]; reused for face fluxes.
  __shared__ double s_f[S::kLds];

  const Geo &g = a.geo;
  const int nd = g.n[DAX], nc = g.n[CAX];
  const int tiles_d = (nd + TL - 1) / TL;
  const int tiles_c = (nc + TC - 1) / TC;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int td = lb % tiles_d;
  const int rest = lb / tiles_d;
  const int tc = rest % tiles_c;
  const int e = rest / tiles_c;
  const int d0 = td * TL, c0 = tc * TC;
  const long long base_e = (long long)e * g.stride[EAX];
  const long long vs = g.vstride;
  const unsigned tid = threadIdx.x;

  if constexpr (S::kSplit) {
    // Phase 1: split fluxes of every tile cell and its halo along DAX. f+ goes
    // to LDS now; each thread keeps its cells' f- in registers until the upwind
    // half of every face is done, then f- takes f+'s place. LDS holds one of
    // the two at a time, which halves it and lets more blocks share a CU.
    constexpr int kCells = TLH * TC;
    constexpr int kCellPasses = (kCells + kThreads - 1) / kThreads;
    double fm_keep[kCellPasses][NV];
  #pragma unroll
    for (int cp = 0; cp < kCellPasses; ++cp) {
      const unsigned idx = tid + cp * kThreads;
      if (idx < kCells) {
        const int l = DFAST ? idx % TLH : idx / TC;
        const int cc = DFAST ? idx / TLH : idx % TC;
        const int gc = c0 + cc;
        double fp[NV];
        if (gc < nc) {
          // cells past nd + R - 1 only feed faces outside the grid; clamping
          // them keeps the periodic image inside [0, nd) for any tile overhang
          const int gd = remap(min(d0 + l - R, nd - 1 + R), nd, g.periodic[DAX]);
          const long long q = base_e + (long long)gd * g.stride[DAX] + (long long)gc * g.stride[CAX];
          double u[NV], pl[M::NPL];
  #pragma unroll
          for (int v = 0; v < NV; ++v) u[v] = a.U[v * vs + q];
  #pragma unroll
          for (int v = 0; v < M::NPL; ++v) pl[v] = a.V[(M::P0 + v) * vs + q];
          double f[NV];
          M::template flux<DAX, QUIRK>(u, pl, a, q, f);
  #pragma unroll
          for (int v = 0; v < NV; ++v) {
            fp[v] = 0.5 * (f[v] + u[v]);
            fm_keep[cp][v] = 0.5 * (f[v] - u[v]);
          }
        } else {
  #pragma unroll
          for (int v = 0; v < NV; ++v) fp[v] = fm_keep[cp][v] = 0.0;
        }
  #pragma unroll
        for (int v = 0; v < NV; ++v) s_f[S::fidx(v, l, cc)] = fp[v];
      }
    }
    __syncthreads();

    // Phase 2: WENO face values, face lf being the left face of tile cell lf.
    // The (variable, face) pairs are dealt out one per thread per pass, so the
    // faces beyond one per thread cost a fraction of a pass spread over the
    // waves instead of a full pass of all variables on one wave. Upwind halves
    // from f+ first, then downwind halves from f-.
    constexpr int kItems = NV * S::kFaces;
    constexpr int kPasses = (kItems + kThreads - 1) / kThreads;
    double nu[kPasses], du[kPasses];
  #pragma unroll
    for (int k = 0; k < kPasses; ++k) {
      const unsigned item = tid + k * kThreads;
      if (item < kItems) {
        const double *fpv = s_f + S::item_fidx(item);
        Weno<ORDER>::up([&](int m) { return fpv[m * S::kStrideL]; }, nu[k], du[k]);
      }
    }
    __syncthreads();
  #pragma unroll
    for (int cp = 0; cp < kCellPasses; ++cp) {
      const unsigned idx = tid + cp * kThreads;
      if (idx < kCells) {
        const int l = DFAST ? idx % TLH : idx / TC;
        const int cc = DFAST ? idx / TLH : idx % TC;
  #pragma unroll
        for (int v = 0; v < NV; ++v) s_f[S::fidx(v, l, cc)] = fm_keep[cp][v];
      }
    }
    __syncthreads();
  #pragma unroll
    for (int k = 0; k < kPasses; ++k) {
      const unsigned item = tid + k * kThreads;
      if (item < kItems) {
        const double *fmv = s_f + S::item_fidx(item);
        double nd, dd;
        Weno<ORDER>::down([&](int m) { return fmv[m * S::kStrideL]; }, nd, dd);
        nu[k] = weno_combine(nu[k], du[k], nd, dd);
      }
    }
    __syncthreads();
    // item = v * kFaces + idx, idx enumerating (lf, cc) in gidx order for both
    // layouts, so the face of item goes to s_f[item]
  #pragma unroll
    for (int k = 0; k < kPasses; ++k) {
      const unsigned item = tid + k * kThreads;
      if (item < kItems) s_f[item] = nu[k];
    }
    __syncthreads();
  } else {
    // Phase 1: f+ and f- of every tile cell and its halo side by side in LDS.
    for (unsigned idx = tid; idx < TLH * TC; idx += kThreads) {
      const int l = DFAST ? idx % TLH : idx / TC;
      const int cc = DFAST ? idx / TLH : idx % TC;
      const int gc = c0 + cc;
      double fp[NV], fm[NV];
      if (gc < nc) {
        const int gd = remap(min(d0 + l - R, nd - 1 + R), nd, g.periodic[DAX]);
        const long long q = base_e + (long long)gd * g.stride[DAX] + (long long)gc * g.stride[CAX];
        double u[NV], pl[M::NPL];
#pragma unroll
        for (int v = 0; v < NV; ++v) u[v] = a.U[v * vs + q];
#pragma unroll
        for (int v = 0; v < M::NPL; ++v) pl[v] = a.V[(M::P0 + v) * vs + q];
        double f[NV];
        M::template flux<DAX, QUIRK>(u, pl, a, q, f);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          fp[v] = 0.5 * (f[v] + u[v]);
          fm[v] = 0.5 * (f[v] - u[v]);
        }
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) fp[v] = fm[v] = 0.0;
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        s_f[S::fidx(v, l, cc)] = fp[v];
        s_f[S::fidx(NV + v, l, cc)] = fm[v];
      }
    }
    __syncthreads();
    constexpr int kItems = NV * S::kFaces;
    constexpr int kPasses = (kItems + kThreads - 1) / kThreads;
    double face[kPasses];
#pragma unroll
    for (int k = 0; k < kPasses; ++k) {
      const unsigned item = tid + k * kThreads;
      if (item < kItems) {
        const double *fpv = s_f + S::item_fidx(item);
        const double *fmv = fpv + NV * TLH * TC;
        face[k] = Weno<ORDER>::face([&](int m) { return fpv[m * S::kStrideL]; },
                                    [&](int m) { return fmv[m * S::kStrideL]; });
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPasses; ++k) {
      const unsigned item = tid + k * kThreads;
      if (item < kItems) s_f[item] = face[k];
    }
    __syncthreads();
  }

  // Phase 3: net flux of each tile cell, then the sweep's mode.
  const int l = DFAST ? tid % TL : tid / TC;
  const int cc = DFAST ? tid / TL : tid % TC;
  const int gd = d0 + l, gc = c0 + cc;
  const bool live = gd < nd && gc < nc;
  CellC2P cs;
  if (live) {
    const long long q = base_e + (long long)gd * g.stride[DAX] + (long long)gc * g.stride[CAX];
    const double inv_d = g.inv_d[DAX];
    // net flux of this axis, added to the earlier axes' as one fused
    // multiply-add so every kernel that accumulates it rounds alike
    double fn[NV];
    if (MODE == kSweepWrite) {
      // streamed out, as the stage outputs are (finish_rk2)
#pragma unroll
      for (int v = 0; v < NV; ++v)
        __builtin_nontemporal_store((s_f[S::gidx(v, l + 1, cc)] - s_f[S::gidx(v, l, cc)]) * inv_d, &a.fnet[v * vs + q]);
    } else {
#pragma unroll
      for (int v = 0; v < NV; ++v)
        fn[v] = fma(s_f[S::gidx(v, l + 1, cc)] - s_f[S::gidx(v, l, cc)], inv_d, a.fnet[v * vs + q]);
      if (MODE == kSweepAdd) {
#pragma unroll
        for (int v = 0; v < NV; ++v) a.fnet[v * vs + q] = fn[v];
      } else if (MODE == kSweepStage1 || MODE == kSweepStage2) {
        double un[NV];
        if (MODE == kSweepStage1) {
#pragma unroll
          for (int v = 0; v < NV; ++v) un[v] = rk2_stage1(a.U[v * vs + q], fn[v], a.dt);
        } else {
#pragma unroll
          for (int v = 0; v < NV; ++v) un[v] = rk2_stage2(a.Un[v * vs + q], a.U[v * vs + q], fn[v], a.dt);
        }
        finish_rk2(a, q, un, M(), cs);
      } else if (MODE == kSweepImex2) {
        imex2_apply(a, q, fn, cs);
      } else if (MODE == kSweepImexFinal) {
        imex_final_apply(a, q, fn, cs);
      }
    }
  }
  if (MODE >= kSweepStage1) reduce_stats(a.stats, cs.cells, cs.iters, cs.fails, cs.maxit);
}

// The y and z terms of a 3-D stage in one launch, after an x sweep has
// written the x term (sweep plan 3): a 16 x 16 (y, z) tile of one x plane
// reconstructs its y faces, then its z faces, each phase exactly as a
// separate sweep does it. Each cell's net flux with the y term added waits
// (in registers or LDS) for the z term, and the z phase then applies the
// sweep's mode. The sums
// are formed in the separate sweeps' order, so plan 3 gives plan 2's bits,
// and the y-added net flux never makes an HBM round trip. SRRMHD's 14
// variables take each phase over the tile's pencils in two groups of 8, so
// that a thread keeps the split fluxes of one tile-or-halo cell and the face
// values of about five (variable, face) items at a time (174-203 VGPRs where
// the whole tile takes 184-256); SRMHD's 9 variables run faster as one group
// of 16, with each cell's y-added net flux kept in registers instead of LDS:
// 116-128 VGPRs and 23-25 KB of LDS, 4 waves per SIMD (in groups of 8 they
// fit 4 waves per SIMD too but took 1 % longer).
template <class M, int ORDER, int MODE, bool QUIRK>
__global__ __launch_bounds__(kThreads, M::NC > 9 ? 1 : 4) void k_sweep_yzf(SweepArgs a) {
  constexpr int kGroup = M::NC > 9 ? 8 : 16;  // pencils per group
  using SY = SweepShape<M, ORDER, 1, 2, false, 16, kGroup>;
  using SZ = SweepShape<M, ORDER, 2, 1, true, 16, kGroup>;
  constexpr int NV = M::NC;
  constexpr int kLds = SY::kLds > SZ::kLds ? SY::kLds : SZ::kLds;
  __shared__ double s_f[kLds];
  const Geo &g = a.geo;
  const int ny = g.n[1], nz = g.n[2];
  const int tiles_y = (ny + 15) / 16, tiles_z = (nz + 15) / 16;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int tz = lb % tiles_z;
  const int rest = lb / tiles_z;
  const int ty = rest % tiles_y;
  const int i = rest / tiles_y;
  const int y0 = ty * 16, z0 = tz * 16;
  const long long base_e = (long long)i * g.stride[0];
  const long long vs = g.vstride;
  const unsigned tid = threadIdx.x;
  // the thread's cell is (y0 + r, z0 + c)
  const int r = tid / 16, c = tid % 16;
  const bool live = y0 + r < ny && z0 + c < nz;
  const long long q = base_e + (long long)(y0 + r) * g.stride[1] + (z0 + c);

  CellC2P cs;
  if constexpr (kGroup == 16) {
    // y phase over the whole tile, pencils along z
    // each cell's net flux with the y term added waits in registers
    unsigned ty = tid;
    asm volatile("" : "+v"(ty));
    sweep_tile_faces<M, ORDER, 1, 2, false, 16, 16, false>(a, s_f, y0, z0, base_e, ny, nz, ty);
    double acc[NV];
    if (live) {
#pragma unroll
      for (int v = 0; v < NV; ++v)
        acc[v] = fma(s_f[SY::gidx(v, r + 1, c)] - s_f[SY::gidx(v, r, c)], g.inv_d[1],
                     __builtin_nontemporal_load(&a.fnet[v * vs + q]));  // read once
    }
    __syncthreads();  // every y face is read before the z phase refills LDS
    // the z phase's loads stay after the y phase, so its registers are free
    __builtin_amdgcn_sched_barrier(0);
    // Each phase takes the thread index as a value the compiler cannot see
    // through, so that it neither forms the z phase's LDS offsets early nor
    // shares them with the y phase's, which kept them in registers across
    // both phases (156 VGPRs instead of 128).
    unsigned tz = tid;
    asm volatile("" : "+v"(tz));
    // z phase, pencils along y: the thread's cell is tile cell c of pencil r,
    // and it finishes straight from the faces in LDS
    sweep_tile_faces<M, ORDER, 2, 1, true, 16, 16, QUIRK>(a, s_f, z0, y0, base_e, nz, ny, tz);
    if (live) {
      double fn[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v)
        fn[v] = fma(s_f[SZ::gidx(v, c + 1, r)] - s_f[SZ::gidx(v, c, r)], g.inv_d[2], acc[v]);
      sweep_apply<M, MODE>(a, q, fn, cs);
    }
  } else {
    __shared__ double s_acc[NV][kThreads];  // each cell's net flux, y term added, then z
    // y phase: faces along y; group h holds the kGroup pencils from z0 + kGroup h
#pragma unroll 1
    for (int h = 0; h < 16 / kGroup; ++h) {
      sweep_tile_faces<M, ORDER, 1, 2, false, 16, kGroup, false>(a, s_f, y0, z0 + h * kGroup, base_e, ny, nz, tid);
      const int cg = c - h * kGroup;
      if (live && cg >= 0 && cg < kGroup) {
#pragma unroll
        for (int v = 0; v < NV; ++v)
          s_acc[v][tid] = fma(s_f[SY::gidx(v, r + 1, cg)] - s_f[SY::gidx(v, r, cg)], g.inv_d[1], a.fnet[v * vs + q]);
      }
      __syncthreads();  // every face is read before LDS is refilled
      __builtin_amdgcn_sched_barrier(0);
    }
    // z phase: faces along z; group h holds the kGroup pencils from y0 + kGroup h,
    // whose cells belong to the 16 kGroup threads from 16 kGroup h
#pragma unroll 1
    for (int h = 0; h < 16 / kGroup; ++h) {
      sweep_tile_faces<M, ORDER, 2, 1, true, 16, kGroup, QUIRK>(a, s_f, z0, y0 + h * kGroup, base_e, nz, ny, tid);
      const int rg = r - h * kGroup;
      if (live && rg >= 0 && rg < kGroup) {
#pragma unroll
        for (int v = 0; v < NV; ++v)
          s_acc[v][tid] = fma(s_f[SZ::gidx(v, c + 1, rg)] - s_f[SZ::gidx(v, c, rg)], g.inv_d[2], s_acc[v][tid]);
      }
      __syncthreads();
      __builtin_amdgcn_sched_barrier(0);
    }
    if (live) {
      double fn[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) fn[v] = s_acc[v][tid];
      sweep_apply<M, MODE>(a, q, fn, cs);
    }
  }
  if (MODE >= kSweepStage1) reduce_stats(a.stats, cs.cells, cs.iters, cs.fails, cs.maxit);
}

// The y and x terms of a 2-D stage in one launch (sweep plan 3 on 2-D
// grids): a 16 x 16 (x, y) tile reconstructs its y faces, then its x faces,
// each phase exactly as the separate y and x sweeps do it. Each cell's y net
// flux waits in LDS for the x term, which is added as the separate x sweep
// adds it (one fused multiply-add), and the x phase then applies the sweep's
// mode; so the launch gives the separate sweeps' bits, and the y net flux
// never makes an HBM round trip. The y phase runs along the contiguous axis.
// As in k_sweep_yzf, SRRMHD takes each phase in two groups of 8 pencils; a
// group's 160-176 tile-and-halo cells are one pass of the threads, so its
// split fluxes go through LDS one half at a time as SRMHD's do (46.6-48.4 KB
// instead of 64.5-68.2, 125-156 VGPRs instead of 130-182: 3 blocks per CU
// instead of 2, and the OT 1024^2 steps 3-9 % faster).
template <class M, int ORDER, int MODE>
__global__ __launch_bounds__(kThreads) void k_sweep_xyf(SweepArgs a) {
  constexpr int kGroup = M::NC > 9 ? 8 : 16;  // pencils per group
  using SY = SweepShape<M, ORDER, 1, 0, true, 16, kGroup, true>;
  using SX = SweepShape<M, ORDER, 0, 1, false, 16, kGroup, true>;
  constexpr int NV = M::NC;
  constexpr int kLds = SY::kLds > SX::kLds ? SY::kLds : SX::kLds;
  __shared__ double s_f[kLds];
  __shared__ double s_acc[NV][kThreads];  // each cell's y net flux, then with the x term
  const Geo &g = a.geo;
  const int nx = g.n[0], ny = g.n[1];
  const int tiles_y = (ny + 15) / 16;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int ty = lb % tiles_y, tx = lb / tiles_y;
  const int x0 = tx * 16, y0 = ty * 16;
  const unsigned tid = threadIdx.x;
  // the thread's cell is (x0 + r, y0 + c): neighbouring threads along y, the
  // contiguous axis
  const int r = tid / 16, c = tid % 16;
  const bool live = x0 + r < nx && y0 + c < ny;
  const long long q = (long long)(x0 + r) * g.stride[0] + (long long)(y0 + c) * g.stride[1];

  // y phase: faces along y; group h holds the kGroup x pencils from x0 + kGroup h,
  // whose cells belong to the 16 kGroup threads from 16 kGroup h
#pragma unroll 1
  for (int h = 0; h < 16 / kGroup; ++h) {
    sweep_tile_faces<M, ORDER, 1, 0, true, 16, kGroup, false, true>(a, s_f, y0, x0 + h * kGroup, 0, ny, nx, tid);
    const int rg = r - h * kGroup;
    if (live && rg >= 0 && rg < kGroup) {
#pragma unroll
      for (int v = 0; v < NV; ++v) s_acc[v][tid] = (s_f[SY::gidx(v, c + 1, rg)] - s_f[SY::gidx(v, c, rg)]) * g.inv_d[1];
    }
    __syncthreads();  // every face is read before LDS is refilled
    __builtin_amdgcn_sched_barrier(0);
  }
  CellC2P cs;
  if constexpr (kGroup == 16) {
    // x phase over the whole tile, pencils along y: the thread's cell is tile
    // cell r of pencil c, and it finishes straight from the faces in LDS
    sweep_tile_faces<M, ORDER, 0, 1, false, 16, 16, false>(a, s_f, x0, y0, 0, nx, ny, tid);
    if (live) {
      double fn[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v)
        fn[v] = fma(s_f[SX::gidx(v, r + 1, c)] - s_f[SX::gidx(v, r, c)], g.inv_d[0], s_acc[v][tid]);
      sweep_apply<M, MODE>(a, q, fn, cs);
    }
  } else {
    // x phase: faces along x; group h holds the kGroup y pencils from y0 + kGroup h
#pragma unroll 1
    for (int h = 0; h < 16 / kGroup; ++h) {
      sweep_tile_faces<M, ORDER, 0, 1, false, 16, kGroup, false, true>(a, s_f, x0, y0 + h * kGroup, 0, nx, ny, tid);
      const int cg = c - h * kGroup;
      if (live && cg >= 0 && cg < kGroup) {
#pragma unroll
        for (int v = 0; v < NV; ++v)
          s_acc[v][tid] = fma(s_f[SX::gidx(v, r + 1, cg)] - s_f[SX::gidx(v, r, cg)], g.inv_d[0], s_acc[v][tid]);
      }
      __syncthreads();
      __builtin_amdgcn_sched_barrier(0);
    }
    if (live) {
      double fn[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) fn[v] = s_acc[v][tid];
      sweep_apply<M, MODE>(a, q, fn, cs);
    }
  }
  if (MODE >= kSweepStage1) reduce_stats(a.stats, cs.cells, cs.iters, cs.fails, cs.maxit);
}

// Fused z and y sweeps of a 3-D grid (FVS::F's z and y terms,
// fluxVectorSplitting.cc:92-112). A TYT x TZT tile of one x plane
// reconstructs its z faces, then its y faces, through LDS, and accumulates
// the net flux in the order of the separate sweeps (z, then y as one fused
// multiply-add), so the result is the same bit for bit; the z net flux is
// re-read from L2 rather than making an HBM round trip.
template <class M, int ORDER>
struct YzShape {
  static constexpr int TYT = 16, TZT = 16;  // rows along y, cells along z
  static constexpr int R = Weno<ORDER>::R, NV = M::NC;
  static constexpr int WZ = TZT + 2 * R, HY = TYT + 2 * R;
  static constexpr int kLdsZ = 2 * NV * TYT * WZ, kLdsY = 2 * NV * HY * TZT;
  static constexpr int kLds = kLdsZ > kLdsY ? kLdsZ : kLdsY;
  static_assert(TYT * TZT == kThreads, "one cell per thread");
};

template <class M, int DIR, bool QUIRK>
__device__ __forceinline__ void split_cell(const SweepArgs &a, long long q, double *fp, double *fm) {
  const long long vs = a.geo.vstride;
  double u[M::NC], pl[M::NPL], f[M::NC];
#pragma unroll
  for (int v = 0; v < M::NC; ++v) u[v] = a.U[v * vs + q];
#pragma unroll
  for (int v = 0; v < M::NPL; ++v) pl[v] = a.V[(M::P0 + v) * vs + q];
  M::template flux<DIR, QUIRK>(u, pl, a, q, f);
#pragma unroll
  for (int v = 0; v < M::NC; ++v) {
    fp[v] = 0.5 * (f[v] + u[v]);
    fm[v] = 0.5 * (f[v] - u[v]);
  }
}

template <class M, int ORDER, bool QUIRK>
__global__ __launch_bounds__(kThreads) void k_sweep_yz(SweepArgs a) {
  using S = YzShape<M, ORDER>;
  constexpr int TYT = S::TYT, TZT = S::TZT, R = S::R, NV = S::NV, WZ = S::WZ, HY = S::HY;
  __shared__ double s_f[S::kLds];
  const Geo &g = a.geo;
  const int ny = g.n[1], nz = g.n[2];
  const int tiles_z = (nz + TZT - 1) / TZT, tiles_y = (ny + TYT - 1) / TYT;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int tz = lb % tiles_z;
  const int rest = lb / tiles_z;
  const int ty = rest % tiles_y;
  const int i = rest / tiles_y;
  const int y0 = ty * TYT, z0 = tz * TZT;
  const long long base = (long long)i * g.stride[0], sy = g.stride[1];
  const long long vs = g.vstride;
  const int tid = threadIdx.x;
  const int r = tid / TZT, c = tid % TZT;  // the thread's cell: y0 + r, z0 + c

  // ---- z faces: rows y0 .. y0+TYT-1, cells z0-R .. z0+TZT+R-1
  for (int idx = tid; idx < TYT * WZ; idx += kThreads) {
    const int rr = idx / WZ, cz = idx - rr * WZ;
    const int gy = y0 + rr;
    double fp[NV], fm[NV];
    if (gy < ny) {
      const int gz = remap(min(z0 + cz - R, nz - 1 + R), nz, g.periodic[2]);
      split_cell<M, 2, QUIRK>(a, base + gy * sy + gz, fp, fm);
    } else {
#pragma unroll
      for (int v = 0; v < NV; ++v) fp[v] = fm[v] = 0.0;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      s_f[(v * TYT + rr) * WZ + cz] = fp[v];
      s_f[((NV + v) * TYT + rr) * WZ + cz] = fm[v];
    }
  }
  __syncthreads();
  constexpr int kFacesZ = TYT * (TZT + 1), kItemsZ = NV * kFacesZ;
  constexpr int kPassesZ = (kItemsZ + kThreads - 1) / kThreads;
  double face_z[kPassesZ];
#pragma unroll
  for (int k = 0; k < kPassesZ; ++k) {
    const int item = tid + k * kThreads;
    if (item < kItemsZ) {
      const int v = item / kFacesZ, idx = item - v * kFacesZ;
      const int rr = idx / (TZT + 1), lf = idx - rr * (TZT + 1);
      const double *fpv = s_f + (v * TYT + rr) * WZ + lf;
      const double *fmv = s_f + ((NV + v) * TYT + rr) * WZ + lf;
      face_z[k] = Weno<ORDER>::face([&](int m) { return fpv[m]; }, [&](int m) { return fmv[m]; });
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the passes from piling up their LDS reads
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPassesZ; ++k) {
    const int item = tid + k * kThreads;
    if (item < kItemsZ) s_f[item] = face_z[k];
  }
  __syncthreads();
  // the z net flux goes to fnet now and is read back by the same thread
  // below: the line is still in L2 then, so HBM sees one write of it, and no
  // registers are held across the y phase
  const bool live = y0 + r < ny && z0 + c < nz;
  const long long q = base + (y0 + r) * sy + (z0 + c);
  if (live) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const double *gf = s_f + (v * TYT + r) * (TZT + 1) + c;
      a.fnet[v * vs + q] = (gf[1] - gf[0]) * g.inv_d[2];
    }
  }
  __syncthreads();

  // ---- y faces: rows y0-R .. y0+TYT+R-1, cells z0 .. z0+TZT-1
  for (int idx = tid; idx < HY * TZT; idx += kThreads) {
    const int ry = idx / TZT, cc = idx - ry * TZT;
    const int gz = z0 + cc;
    double fp[NV], fm[NV];
    if (gz < nz) {
      const int gy = remap(min(y0 + ry - R, ny - 1 + R), ny, g.periodic[1]);
      split_cell<M, 1, QUIRK>(a, base + gy * sy + gz, fp, fm);
    } else {
#pragma unroll
      for (int v = 0; v < NV; ++v) fp[v] = fm[v] = 0.0;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      s_f[(v * HY + ry) * TZT + cc] = fp[v];
      s_f[((NV + v) * HY + ry) * TZT + cc] = fm[v];
    }
  }
  __syncthreads();
  constexpr int kFacesY = (TYT + 1) * TZT, kItemsY = NV * kFacesY;
  constexpr int kPassesY = (kItemsY + kThreads - 1) / kThreads;
  double face_y[kPassesY];
#pragma unroll
  for (int k = 0; k < kPassesY; ++k) {
    const int item = tid + k * kThreads;
    if (item < kItemsY) {
      const int v = item / kFacesY, idx = item - v * kFacesY;
      const int lf = idx / TZT, cc = idx - lf * TZT;
      const double *fpv = s_f + (v * HY + lf) * TZT + cc;
      const double *fmv = s_f + ((NV + v) * HY + lf) * TZT + cc;
      face_y[k] = Weno<ORDER>::face([&](int m) { return fpv[m * TZT]; }, [&](int m) { return fmv[m * TZT]; });
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the passes from piling up their LDS reads
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPassesY; ++k) {
    const int item = tid + k * kThreads;
    if (item < kItemsY) s_f[item] = face_y[k];
  }
  __syncthreads();

  if (!live) return;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double *gf = s_f + (v * (TYT + 1) + r) * TZT + c;
    a.fnet[v * vs + q] = fma(gf[TZT] - gf[0], g.inv_d[1], a.fnet[v * vs + q]);
  }
}

// SRMHD::primsToAll over interior cells (srmhd.cc:573-673).
__global__ __launch_bounds__(kThreads) void k_prims_to_all_srmhd(Geo g, const double *V, double *U, double *A,
                                                                  double gam) {
  const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (q >= g.vstride) return;
  const long long vs = g.vstride;
  double P[8], u[9], aux[13];
#pragma unroll
  for (int v = 0; v < 8; ++v) P[v] = V[v * vs + q];
  srmhd_prims_to_all(P, gam, u, aux);
#pragma unroll
  for (int v = 0; v < 9; ++v) U[v * vs + q] = u[v];
#pragma unroll
  for (int v = 0; v < 13; ++v) A[v * vs + q] = aux[v];
}

// SRMHD aux of resident cons from the cons2prim root (v^2, Z) a stage kept:
// the recovery of srmhd.cc:509-557 on the same inputs the stage had, so the
// aux equals what the stage would have written.
__global__ __launch_bounds__(kThreads) void k_aux_srmhd(Geo g, const double *U, const double *root, double *A,
                                                         double gam) {
  const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (q >= g.vstride) return;
  const long long vs = g.vstride;
  double u[9], P[8], aux[13];
#pragma unroll
  for (int v = 0; v < 9; ++v) u[v] = U[v * vs + q];
  srmhd_recover(u, gam, root[q], root[vs + q], P, aux);
#pragma unroll
  for (int v = 0; v < 13; ++v) A[v * vs + q] = aux[v];
}

// Aux of every cell from its cons and kept cons2prim root Z (the values
// srrmhd_c2p_cell writes: srrmhd_recover).
__global__ __launch_bounds__(kThreads) void k_aux_srrmhd(Geo g, const double *U, const double *root, double *A,
                                                          double gam, double sigma) {
  const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (q >= g.vstride) return;
  const long long vs = g.vstride;
  double u[kSrrmhdCons], P[kSrrmhdPrims], aux[kSrrmhdAux];
#pragma unroll
  for (int v = 0; v < kSrrmhdCons; ++v) u[v] = U[v * vs + q];
  srrmhd_recover(u, gam, sigma, root[q], P, aux);
#pragma unroll
  for (int v = 0; v < kSrrmhdAux; ++v) A[v * vs + q] = aux[v];
}

// SRMHD::getPrimitiveVars on resident cons with resident prims as the guess
// (srmhd.cc:390-562).
// cold: start from v^2 = 0, Z = tau + D instead of the previous primitives.
__global__ __launch_bounds__(kThreads) void k_c2p_srmhd(Geo g, const double *U, double *V, double *A, double gam,
                                                         C2PCounters *st, int *fail_count, long long *fail_list,
                                                         unsigned char *fail_flag, int cold,
                                                         unsigned long long *hist) {
  const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
  int cells = 0, iters = 0, fails = 0, maxit = 0, steps = -1;
  if (q < g.vstride) {
    const long long vs = g.vstride;
    double u[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) u[v] = U[v * vs + q];
    double v2, Z;
    if (cold) {
      v2 = 0.0;
      Z = u[4] + u[0];
    } else {
      srmhd_plain_guess(V[q], V[vs + q], V[2 * vs + q], V[3 * vs + q], V[4 * vs + q], gam, v2, Z);
    }
    const double BS = u[5] * u[1] + u[6] * u[2] + u[7] * u[3];
    const double Bsq = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
    const double Ssq = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
    const int its = srmhd_newton(u[0], u[4], Bsq, Ssq, BS, gam, v2, Z);
    double P[8], aux[13];
    srmhd_recover(u, gam, v2, Z, P, aux);
#pragma unroll
    for (int v = 0; v < 8; ++v) V[v * vs + q] = P[v];
#pragma unroll
    for (int v = 0; v < 13; ++v) A[v * vs + q] = aux[v];
    cells = 1;
    steps = its;
    if (its < 0) {
      fails = 1;
      const int slot = atomicAdd(fail_count, 1);
      fail_list[slot] = q;
      fail_flag[q] = 1;
    } else {
      iters = its;
      maxit = its;
    }
  }
  reduce_stats(st, cells, iters, fails, maxit);
  hist_add_block(hist, cells > 0, steps);
}

// Retry of failed cells from the mean root of converged interior face
// neighbours (srmhd.cc:465-507). With `last_cell` >= 0 the system solved is
// that cell's, as the reference's retry reuses the last cell's arguments.
__global__ __launch_bounds__(kThreads) void k_c2p_retry_srmhd(Geo g, const double *U, double *V, double *A,
                                                               double *root, double gam, const int *fail_count,
                                                               const long long *fail_list, unsigned char *fail_flag,
                                                               long long last_cell, C2PCounters *st) {
  const int nfail = *fail_count;
  const long long vs = g.vstride;
  for (int f = blockIdx.x * kThreads + threadIdx.x; f < nfail; f += gridDim.x * kThreads) {
    const long long q = fail_list[f];
    const int i = (int)(q / g.stride[0]);
    const int j = (int)((q / g.stride[1]) % g.n[1]);
    const int k = (int)(q % g.n[2]);
    const int pos[3] = {i, j, k};
    double sv2 = 0.0, sZ = 0.0;
    int cnt = 0;
    for (int ax = 0; ax < 3; ++ax) {
      for (int s = -1; s <= 1; s += 2) {
        const int np = pos[ax] + s;
        if (np < 0 || np >= g.n[ax] || g.n[ax] == 1) continue;
        const long long m = q + s * g.stride[ax];
        if (fail_flag[m]) continue;
        double rho = V[m], vx = V[vs + m], vy = V[2 * vs + m], vz = V[3 * vs + m], p = V[4 * vs + m];
        double v2, Z;
        srmhd_plain_guess(rho, vx, vy, vz, p, gam, v2, Z);
        sv2 += v2;
        sZ += Z;
        ++cnt;
      }
    }
    double v2 = cnt ? sv2 / cnt : 0.0, Z = cnt ? sZ / cnt : 0.0;
    const long long src = last_cell >= 0 ? last_cell : q;
    double u[9];
    for (int v = 0; v < 9; ++v) u[v] = U[v * vs + src];
    const double BS = u[5] * u[1] + u[6] * u[2] + u[7] * u[3];
    const double Bsq = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
    const double Ssq = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
    const int its = srmhd_newton(u[0], u[4], Bsq, Ssq, BS, gam, v2, Z);
    // recovery formulas use this cell's own cons, as srmhd.cc:509-557 does
    for (int v = 0; v < 9; ++v) u[v] = U[v * vs + q];
    double P[8], aux[13];
    srmhd_recover(u, gam, v2, Z, P, aux);
    for (int v = 0; v < 8; ++v) V[v * vs + q] = P[v];
    if (A)
      for (int v = 0; v < 13; ++v) A[v * vs + q] = aux[v];
    else if (root) {
      root[q] = v2;
      root[vs + q] = Z;
    }
    if (st && its < 0) atomicAdd(&st->retry_failures, 1ull);
  }
}

__global__ void k_clear_fail_flags(const int *fail_count, const long long *fail_list, unsigned char *fail_flag) {
  const int nfail = *fail_count;
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nfail; f += gridDim.x * blockDim.x)
    fail_flag[fail_list[f]] = 0;
}

// Per-axis maximum of srmhd_max_speed over interior cells: grid-stride
// per-thread maxima, a butterfly over the 64 lanes with __shfl_xor, one LDS
// slot per wave, then one atomicMax per block on the bit pattern (for
// non-negative doubles the unsigned order of the bits is the numeric order).
__global__ __launch_bounds__(kThreads) void k_max_wavespeed_srmhd(Geo g, const double *V, const double *A,
                                                                   unsigned long long *out) {
  const long long vs = g.vstride;
  double m[3] = {0.0, 0.0, 0.0};
  for (long long q = (long long)blockIdx.x * kThreads + threadIdx.x; q < vs; q += (long long)gridDim.x * kThreads) {
    const double rho = V[q], h = A[q], c = A[3 * vs + q], b2 = A[8 * vs + q], v2 = A[9 * vs + q];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (d < g.dims) m[d] = fmax(m[d], srmhd_max_speed(rho, V[(1 + d) * vs + q], v2, h, c, b2));
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
    for (int off = 32; off > 0; off >>= 1) m[d] = fmax(m[d], __shfl_xor(m[d], off));
  __shared__ double s_m[kThreads / 64][3];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int d = 0; d < 3; ++d) s_m[w][d] = m[d];
  __syncthreads();
  if (threadIdx.x < 3) {
    double r = 0.0;
    for (int k = 0; k < kThreads / 64; ++k) r = fmax(r, s_m[k][threadIdx.x]);
    atomicMax(out + threadIdx.x, (unsigned long long)__double_as_longlong(r));
  }
}

__global__ void k_zero_int(int *p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// Padded reference layout -> interior device layout for nvar variables.
__global__ __launch_bounds__(kThreads) void k_compact(Geo g, PadGeo pg, int nvar, const double *src, double *dst) {
  const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (q >= g.vstride) return;
  const int i = (int)(q / g.stride[0]);
  const int j = (int)((q / g.stride[1]) % g.n[1]);
  const int k = (int)(q % g.n[2]);
  const long long p = ((long long)(i + pg.off[0]) * pg.N[1] + (j + pg.off[1])) * pg.N[2] + (k + pg.off[2]);
  for (int v = 0; v < nvar; ++v) dst[v * g.vstride + q] = src[v * pg.ntot + p];
}

// Interior device layout -> padded reference layout with every ghost cell
// holding the value Bcs::apply gives it (per-axis remap, x then y then z).
// ghosts_zero: leave ghosts at zero instead (FVS::F writes interior only).
__global__ __launch_bounds__(kThreads) void k_expand(Geo g, PadGeo pg, int nvar, const double *src, double *dst,
                                                      int ghosts_zero) {
  const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (p >= pg.ntot) return;
  const int I = (int)(p / ((long long)pg.N[1] * pg.N[2]));
  const int J = (int)((p / pg.N[2]) % pg.N[1]);
  const int K = (int)(p % pg.N[2]);
  const int c[3] = {I - pg.off[0], J - pg.off[1], K - pg.off[2]};
  bool ghost = false;
  long long q = 0;
  for (int ax = 0; ax < 3; ++ax) {
    int x = c[ax];
    if (x < 0 || x >= g.n[ax]) ghost = true;
    x = remap(x, g.n[ax], g.periodic[ax]);
    q += (long long)x * g.stride[ax];
  }
  for (int v = 0; v < nvar; ++v) dst[v * pg.ntot + p] = (ghost && ghosts_zero) ? 0.0 : src[v * g.vstride + q];
}

// Model::fluxVector of the resident state for every padded cell (the padded
// cell's value is its boundary image's, as after Bcs::apply).
template <int DIR, bool QUIRK>
__global__ __launch_bounds__(kThreads) void k_flux_vector_srmhd(Geo g, PadGeo pg, const double *U, const double *V,
                                                                 double *dst) {
  const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (p >= pg.ntot) return;
  const int I = (int)(p / ((long long)pg.N[1] * pg.N[2]));
  const int J = (int)((p / pg.N[2]) % pg.N[1]);
  const int K = (int)(p % pg.N[2]);
  const int c[3] = {I - pg.off[0], J - pg.off[1], K - pg.off[2]};
  long long q = 0;
  for (int ax = 0; ax < 3; ++ax) q += (long long)remap(c[ax], g.n[ax], g.periodic[ax]) * g.stride[ax];
  const long long vs = g.vstride;
  double u[9], f[9];
  for (int v = 0; v < 9; ++v) u[v] = U[v * vs + q];
  srmhd_flux<DIR, QUIRK>(u, V[vs + q], V[2 * vs + q], V[3 * vs + q], V[4 * vs + q], f);
  for (int v = 0; v < 9; ++v) dst[v * pg.ntot + p] = f[v];
}

// Model::sourceTerm of the resident state for every padded cell: only phi
// has a source, -phi / cp^2 (srmhd.cc:233-250).
__global__ __launch_bounds__(kThreads) void k_source_srmhd(Geo g, PadGeo pg, const double *U, double *dst,
                                                            double cp) {
  const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (p >= pg.ntot) return;
  const int I = (int)(p / ((long long)pg.N[1] * pg.N[2]));
  const int J = (int)((p / pg.N[2]) % pg.N[1]);
  const int K = (int)(p % pg.N[2]);
  const int c[3] = {I - pg.off[0], J - pg.off[1], K - pg.off[2]};
  long long q = 0;
  for (int ax = 0; ax < 3; ++ax) q += (long long)remap(c[ax], g.n[ax], g.periodic[ax]) * g.stride[ax];
  for (int v = 0; v < 8; ++v) dst[v * pg.ntot + p] = 0.0;
  dst[8 * pg.ntot + p] = -U[8 * g.vstride + q] / (cp * cp);
}

// RKSplit's source half-step and finalise (rkSplit.cc:36-53): cons += dt
// source(cons) with SRMHD's only source, -phi / cp^2 on phi (srmhd.cc:233-250),
// plus the model extension's source on variables 0..7 when ext is given, then
// cons2prim from the current primitives (srmhd.cc:390-562). Without an
// extension only phi changes, so only phi is written back.
__global__ __launch_bounds__(kThreads) void k_split_srmhd(Geo g, double *U, double *V, double *A, double *root,
                                                           double dt, double gam, double cp, C2PCounters *st,
                                                           int *fail_count, long long *fail_list,
                                                           unsigned char *fail_flag, const double *ext) {
  const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
  int cells = 0, iters = 0, fails = 0, maxit = 0;
  if (q < g.vstride) {
    const long long vs = g.vstride;
    double u[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) u[v] = U[v * vs + q];
    const double src = -u[8] / (cp * cp);
    u[8] = u[8] + dt * src;
    U[8 * vs + q] = u[8];
    if (ext) {
      // the model extension's source on variables 0..7 (rkSplit.cc:15-25,
      // where the model's own source there is zero)
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        u[v] = u[v] + dt * ext[v * vs + q];
        U[v * vs + q] = u[v];
      }
    }
    double v2, Z;
    srmhd_plain_guess(V[q], V[vs + q], V[2 * vs + q], V[3 * vs + q], V[4 * vs + q], gam, v2, Z);
    const double BS = u[5] * u[1] + u[6] * u[2] + u[7] * u[3];
    const double Bsq = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
    const double Ssq = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
    const int its = srmhd_newton(u[0], u[4], Bsq, Ssq, BS, gam, v2, Z);
    double P[8], aux[13];
    srmhd_recover(u, gam, v2, Z, P, aux);
#pragma unroll
    for (int v = 0; v < 8; ++v) V[v * vs + q] = P[v];
    if (A) {
#pragma unroll
      for (int v = 0; v < 13; ++v) A[v * vs + q] = aux[v];
    } else {
      root[q] = v2;
      root[vs + q] = Z;
    }
    cells = 1;
    if (its < 0) {
      fails = 1;
      const int slot = atomicAdd(fail_count, 1);
      fail_list[slot] = q;
      fail_flag[q] = 1;
    } else {
      iters = its;
      maxit = its;
    }
  }
  reduce_stats(st, cells, iters, fails, maxit);
}

// ------------------------------------------------------------- SRRMHD

// SSP2 stage one for every interior cell (SSP2.cc:72-105): solve
// x - cons - dt gam psi(x) = 0, then finalise (cons2prim from the cell's
// previous rho h W^2) and keep source1's non-zero components.
__global__ __launch_bounds__(kThreads) void k_imex1_srrmhd(SweepArgs a) {
  const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
  CellC2P cs;
  if (q < a.geo.vstride) {
    const long long vs = a.geo.vstride;
    double c[kSrrmhdCons], x[kSrrmhdCons], A[kSrrmhdAux];
#pragma unroll
    for (int v = 0; v < kSrrmhdCons; ++v) c[v] = a.Un[v * vs + q];
    const double Zg = srrmhd_guess_Z(a.Vg[0 * vs + q], a.Vg[1 * vs + q], a.Vg[2 * vs + q], a.Vg[3 * vs + q],
                                     a.Vg[4 * vs + q], a.gamma);
    const int its = imex_stage(c, a.dt, a.gamma, a.sigma, a.cp, Zg, x);
    c2p_store_srrmhd(a, q, x, A, cs);
    if (its < 0) cs.fails = 1;
    double s5[5];
    srrmhd_source5(x, A, a.cp, s5);
#pragma unroll
    for (int m = 0; m < 5; ++m) __builtin_nontemporal_store(s5[m], &a.Sout[m * vs + q]);
  }
  reduce_stats(a.stats, cs.cells, cs.iters, cs.fails, cs.maxit);
}

// SRRMHD::primsToAll over interior cells (srrmhd.cc:358-510). The charge
// density is the 4th-order central divergence of E; its stencil reads the
// neighbours' boundary images.
__global__ __launch_bounds__(kThreads) void k_prims_to_all_srrmhd(Geo g, const double *V, double *U, double *A,
                                                                   double gam, double sigma) {
  const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (q >= g.vstride) return;
  const long long vs = g.vstride;
  const int pos[3] = {(int)(q / g.stride[0]), (int)((q / g.stride[1]) % g.n[1]), (int)(q % g.n[2])};
  double P[kSrrmhdPrims], u[kSrrmhdCons], aux[kSrrmhdAux];
#pragma unroll
  for (int v = 0; v < kSrrmhdPrims; ++v) P[v] = V[v * vs + q];
  srrmhd_prims_to_all_local(P, gam, u, aux);
  double div = 0.0;
  for (int ax = 0; ax < g.dims; ++ax) {
    long long qq[4];
    const int offs[4] = {2, 1, -1, -2};
    for (int m = 0; m < 4; ++m) {
      long long r = q - (long long)pos[ax] * g.stride[ax];
      r += (long long)remap(pos[ax] + offs[m], g.n[ax], g.periodic[ax]) * g.stride[ax];
      qq[m] = r;
    }
    const double *E = V + (8 + ax) * vs;
    const double term = (-E[qq[0]] + 8 * E[qq[1]] - 8 * E[qq[2]] + E[qq[3]]) * (g.inv_d[ax] / 12.0);
    div = ax == 0 ? term : div + term;
  }
  u[13] = div;
  double J[3];
  srrmhd_current(u, P, aux, sigma, J);
  aux[4] = J[0];
  aux[5] = J[1];
  aux[6] = J[2];
#pragma unroll
  for (int v = 0; v < kSrrmhdCons; ++v) U[v * vs + q] = u[v];
#pragma unroll
  for (int v = 0; v < kSrrmhdAux; ++v) A[v * vs + q] = aux[v];
}

// SRRMHD::getPrimitiveVars (srrmhd.cc:314-356): secant from the resident
// prims' rho h W^2.
__global__ __launch_bounds__(kThreads) void k_c2p_srrmhd(Geo g, const double *U, double *V, double *A, double gam,
                                                          double sigma, C2PCounters *st) {
  const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
  CellC2P cs;
  if (q < g.vstride) {
    const long long vs = g.vstride;
    double u[kSrrmhdCons], P[kSrrmhdPrims], aux[kSrrmhdAux];
#pragma unroll
    for (int v = 0; v < kSrrmhdCons; ++v) u[v] = U[v * vs + q];
    const double Zg = srrmhd_guess_Z(V[q], V[vs + q], V[2 * vs + q], V[3 * vs + q], V[4 * vs + q], gam);
    cs.add(srrmhd_c2p_cell(u, gam, sigma, Zg, P, aux));
#pragma unroll
    for (int v = 0; v < kSrrmhdPrims; ++v) V[v * vs + q] = P[v];
#pragma unroll
    for (int v = 0; v < kSrrmhdAux; ++v) A[v * vs + q] = aux[v];
  }
  reduce_stats(st, cs.cells, cs.iters, cs.fails, cs.maxit);
}

// --------------------------------------------------------------- RKPlus

template <class M>
__global__ __launch_bounds__(kThreads) void k_rkplus_stage(RkStageArgs a) {
  constexpr int NC = M::NC;
  constexpr bool kRes = NC == kSrrmhdCons;
  const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
  CellC2P cs;
  if (q < a.geo.vstride) {
    const long long vs = a.geo.vstride;
    double r[NC];
#pragma unroll
    for (int v = 0; v < NC; ++v) r[v] = 0.0;
    if (a.fnet) {
      // Model::sourceTerm of the state whose rhs this is (srmhd.cc:233-250,
      // srrmhd.cc:194-207), minus its net flux
      if constexpr (kRes) {
        double u[kSrrmhdCons], A[kSrrmhdAux], s5[5];
#pragma unroll
        for (int v = 0; v < kSrrmhdCons; ++v) u[v] = a.Ub[v * vs + q];
#pragma unroll
        for (int v = 4; v < 7; ++v) A[v] = a.Ab[v * vs + q];
        srrmhd_source5(u, A, a.cp, s5);
#pragma unroll
        for (int m = 0; m < 5; ++m) r[8 + m] = s5[m];
      } else {
        r[8] = -a.Ub[8 * vs + q] / (a.cp * a.cp);
        if (a.ext) {
          // the model extension's source (RKPlus.cc:26-38) on variables 0..7
#pragma unroll
          for (int v = 0; v < 8; ++v) r[v] = r[v] + a.ext[v * vs + q];
        }
      }
#pragma unroll
      for (int v = 0; v < NC; ++v) r[v] = r[v] - a.fnet[v * vs + q];
    }
    double un[NC];
    if (a.half_sum) {
#pragma unroll
      for (int v = 0; v < NC; ++v) un[v] = 0.5 * (a.X[0][v * vs + q] + a.X[1][v * vs + q] + a.dt * r[v]);
    } else {
      const double cr = a.wr * a.dt, cR = a.wR * a.dt;
#pragma unroll
      for (int v = 0; v < NC; ++v) {
        double acc = a.w[0] * a.X[0][v * vs + q];
        for (int m = 1; m < a.nx; ++m) acc += a.w[m] * a.X[m][v * vs + q];
        if (a.R) acc += cR * a.R[v * vs + q];
        acc += cr * r[v];
        un[v] = acc;
      }
    }
    if (a.Rout) {
#pragma unroll
      for (int v = 0; v < NC; ++v) a.Rout[v * vs + q] = r[v];
    }
    if constexpr (kRes) {
      const double Zg = srrmhd_guess_Z(a.Vg[0 * vs + q], a.Vg[1 * vs + q], a.Vg[2 * vs + q], a.Vg[3 * vs + q],
                                       a.Vg[4 * vs + q], a.gamma);
      double P[kSrrmhdPrims], A[kSrrmhdAux];
      cs.add(srrmhd_c2p_cell(un, a.gamma, a.sigma, Zg, P, A));
#pragma unroll
      for (int v = 0; v < kSrrmhdCons; ++v) a.Uout[v * vs + q] = un[v];
#pragma unroll
      for (int v = 0; v < kSrrmhdPrims; ++v) a.Vout[v * vs + q] = P[v];
      if (a.Aout) {
#pragma unroll
        for (int v = 0; v < kSrrmhdAux; ++v) a.Aout[v * vs + q] = A[v];
      }
    } else {
      double v2, Z;
      srmhd_plain_guess(a.Vg[0 * vs + q], a.Vg[1 * vs + q], a.Vg[2 * vs + q], a.Vg[3 * vs + q], a.Vg[4 * vs + q],
                        a.gamma, v2, Z);
      const double BS = un[5] * un[1] + un[6] * un[2] + un[7] * un[3];
      const double Bsq = un[5] * un[5] + un[6] * un[6] + un[7] * un[7];
      const double Ssq = un[1] * un[1] + un[2] * un[2] + un[3] * un[3];
      const int its = srmhd_newton(un[0], un[4], Bsq, Ssq, BS, a.gamma, v2, Z);
      double P[kSrmhdPrims], A[kSrmhdAux];
      srmhd_recover(un, a.gamma, v2, Z, P, A);
#pragma unroll
      for (int v = 0; v < kSrmhdCons; ++v) a.Uout[v * vs + q] = un[v];
#pragma unroll
      for (int v = 0; v < kSrmhdPrims; ++v) a.Vout[v * vs + q] = P[v];
      if (a.Aout) {
#pragma unroll
        for (int v = 0; v < kSrmhdAux; ++v) a.Aout[v * vs + q] = A[v];
      }
      cs.add(its);
      if (its < 0) {
        const int slot = atomicAdd(a.fail_count, 1);
        a.fail_list[slot] = q;
        a.fail_flag[q] = 1;
      }
    }
  }
  reduce_stats(a.stats, cs.cells, cs.iters, cs.fails, cs.maxit);
}

__global__ __launch_bounds__(kThreads) void k_rk410_prepare2(long long n, double *U1, double *U2) {
  const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (q >= n) return;
  const double u2 = 1.0 / 25.0 * U2[q] + 9.0 / 25.0 * U1[q];
  U2[q] = u2;
  U1[q] = 15.0 * u2 - 5.0 * U1[q];
}

__device__ __forceinline__ long long padded_image(const Geo &g, const PadGeo &pg, long long p) {
  const int I = (int)(p / ((long long)pg.N[1] * pg.N[2]));
  const int J = (int)((p / pg.N[2]) % pg.N[1]);
  const int K = (int)(p % pg.N[2]);
  const int c[3] = {I - pg.off[0], J - pg.off[1], K - pg.off[2]};
  long long q = 0;
  for (int ax = 0; ax < 3; ++ax) q += (long long)remap(c[ax], g.n[ax], g.periodic[ax]) * g.stride[ax];
  return q;
}

// SRRMHD::sourceTerm (srrmhd.cc:209-254) of the resident state, every padded cell.
__global__ __launch_bounds__(kThreads) void k_source_srrmhd(Geo g, PadGeo pg, const double *U, const double *V,
                                                             double *dst, double gam, double sigma, double cp) {
  const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (p >= pg.ntot) return;
  const long long q = padded_image(g, pg, p), vs = g.vstride;
  double u[kSrrmhdCons], P[kSrrmhdPrims], aux[kSrrmhdAux];
  for (int v = 0; v < kSrrmhdCons; ++v) u[v] = U[v * vs + q];
  for (int v = 0; v < kSrrmhdPrims; ++v) P[v] = V[v * vs + q];
  double uu[kSrrmhdCons];
  srrmhd_prims_to_all_local(P, gam, uu, aux);
  double J[3];
  srrmhd_current(u, P, aux, sigma, J);
  aux[4] = J[0];
  aux[5] = J[1];
  aux[6] = J[2];
  double s5[5];
  srrmhd_source5(u, aux, cp, s5);
  for (int v = 0; v < 8; ++v) dst[v * pg.ntot + p] = 0.0;
  for (int m = 0; m < 5; ++m) dst[(8 + m) * pg.ntot + p] = s5[m];
  dst[13 * pg.ntot + p] = 0.0;
}

// RKSplit's source half-step and finalise for the resistive model: the source
// of SRRMHD::sourceTerm (srrmhd.cc:209-254) from the current state, cons +=
// dt source, then the secant cons2prim from the current rho h W^2.
__global__ __launch_bounds__(kThreads) void k_split_srrmhd(Geo g, double *U, double *V, double *A, double dt,
                                                            double gam, double sigma, double cp, C2PCounters *st) {
  const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
  int cells = 0, iters = 0, fails = 0, maxit = 0;
  if (q < g.vstride) {
    const long long vs = g.vstride;
    double u[kSrrmhdCons], P[kSrrmhdPrims], aux[kSrrmhdAux];
    for (int v = 0; v < kSrrmhdCons; ++v) u[v] = U[v * vs + q];
    for (int v = 0; v < kSrrmhdPrims; ++v) P[v] = V[v * vs + q];
    double uu[kSrrmhdCons];
    srrmhd_prims_to_all_local(P, gam, uu, aux);
    double J[3];
    srrmhd_current(u, P, aux, sigma, J);
    aux[4] = J[0];
    aux[5] = J[1];
    aux[6] = J[2];
    double s5[5];
    srrmhd_source5(u, aux, cp, s5);
    for (int m = 0; m < 5; ++m) u[8 + m] = u[8 + m] + dt * s5[m];
    const double Zg = srrmhd_guess_Z(P[0], P[1], P[2], P[3], P[4], gam);
    double Pn[kSrrmhdPrims], An[kSrrmhdAux];
    const int its = srrmhd_c2p_cell(u, gam, sigma, Zg, Pn, An);
    for (int m = 0; m < 5; ++m) U[(8 + m) * vs + q] = u[8 + m];
    for (int v = 0; v < kSrrmhdPrims; ++v) V[v * vs + q] = Pn[v];
    for (int v = 0; v < kSrrmhdAux; ++v) A[v * vs + q] = An[v];
    cells = 1;
    if (its < 0) {
      fails = 1;
    } else {
      iters = its;
      maxit = its;
    }
  }
  reduce_stats(st, cells, iters, fails, maxit);
}

template <int DIR, bool QUIRK>
__global__ __launch_bounds__(kThreads) void k_flux_vector_srrmhd(Geo g, PadGeo pg, const double *U, const double *V,
                                                                  double *dst, double gam, double sigma) {
  const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (p >= pg.ntot) return;
  const long long q = padded_image(g, pg, p), vs = g.vstride;
  double u[kSrrmhdCons], f[kSrrmhdCons];
  for (int v = 0; v < kSrrmhdCons; ++v) u[v] = U[v * vs + q];
  srrmhd_flux<DIR, QUIRK>(u, V[q], V[vs + q], V[2 * vs + q], V[3 * vs + q], V[4 * vs + q], gam, sigma, f);
  for (int v = 0; v < kSrrmhdCons; ++v) dst[v * pg.ntot + p] = f[v];
}

template <class M, int ORDER, int DAX, int CAX, bool DFAST, int MODE>
void launch_sweep_t(const SweepArgs &a, int quirk, hipStream_t s) {
  using S = SweepShape<M, ORDER, DAX, CAX, DFAST>;
  const int tiles = ((a.geo.n[DAX] + S::TL - 1) / S::TL) * ((a.geo.n[CAX] + S::TC - 1) / S::TC) * a.geo.n[S::EAX];
  // stored fluxes already carry the z quirk (k_flux_stored)
  if (M::kStored)
    hipLaunchKernelGGL((k_sweep<M, ORDER, DAX, CAX, DFAST, MODE, false>), dim3(tiles), dim3(kThreads), 0, s, a);
  else if (quirk && DAX == 2)
    hipLaunchKernelGGL((k_sweep<M, ORDER, DAX, CAX, DFAST, MODE, true>), dim3(tiles), dim3(kThreads), 0, s, a);
  else
    hipLaunchKernelGGL((k_sweep<M, ORDER, DAX, CAX, DFAST, MODE, false>), dim3(tiles), dim3(kThreads), 0, s, a);
}

template <class M, int ORDER, int MODE>
void launch_sweep_o(const SweepArgs &a, int dir, int quirk, hipStream_t s) {
  if (a.geo.dims == 3) {
    if (dir == 0) launch_sweep_t<M, ORDER, 0, 2, false, MODE>(a, quirk, s);
    else if (dir == 1) launch_sweep_t<M, ORDER, 1, 2, false, MODE>(a, quirk, s);
    else launch_sweep_t<M, ORDER, 2, 1, true, MODE>(a, quirk, s);
  } else {
    if (dir == 0) launch_sweep_t<M, ORDER, 0, 1, false, MODE>(a, quirk, s);
    else launch_sweep_t<M, ORDER, 1, 0, true, MODE>(a, quirk, s);
  }
}

template <class M, int ORDER>
void launch_sweep_m(const SweepArgs &a, int dir, int mode, int quirk, hipStream_t s) {
  if constexpr (M::kStored) {
    if (mode == kSweepWrite) launch_sweep_o<M, ORDER, kSweepWrite>(a, dir, 0, s);
    else if (mode == kSweepAdd) launch_sweep_o<M, ORDER, kSweepAdd>(a, dir, 0, s);
    return;
  }
  switch (mode) {
    case kSweepWrite: launch_sweep_o<M, ORDER, kSweepWrite>(a, dir, quirk, s); break;
    case kSweepAdd: launch_sweep_o<M, ORDER, kSweepAdd>(a, dir, quirk, s); break;
    case kSweepStage1: launch_sweep_o<M, ORDER, kSweepStage1>(a, dir, quirk, s); break;
    case kSweepStage2: launch_sweep_o<M, ORDER, kSweepStage2>(a, dir, quirk, s); break;
    default:
      if constexpr (M::NC == kSrrmhdCons) {
        if (mode == kSweepImex2) launch_sweep_o<M, ORDER, kSweepImex2>(a, dir, quirk, s);
        else if (mode == kSweepImexFinal) launch_sweep_o<M, ORDER, kSweepImexFinal>(a, dir, quirk, s);
      }
      break;
  }
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

template <class M, int ORDER>
void launch_sweep_yz_m(const SweepArgs &a, int quirk, hipStream_t s) {
  using S = YzShape<M, ORDER>;
  const int tiles = ((a.geo.n[1] + S::TYT - 1) / S::TYT) * ((a.geo.n[2] + S::TZT - 1) / S::TZT) * a.geo.n[0];
  if (quirk)
    hipLaunchKernelGGL((k_sweep_yz<M, ORDER, true>), dim3(tiles), dim3(kThreads), 0, s, a);
  else
    hipLaunchKernelGGL((k_sweep_yz<M, ORDER, false>), dim3(tiles), dim3(kThreads), 0, s, a);
}

}  // namespace

void launch_sweep_yz(const SweepArgs &a, int model, int order, int quirk, hipStream_t s) {
  if (model == 0) {
    if (order == 3) launch_sweep_yz_m<SrmhdM, 3>(a, quirk, s);
    else launch_sweep_yz_m<SrmhdM, 5>(a, quirk, s);
  } else {
    if (order == 3) launch_sweep_yz_m<SrrmhdM, 3>(a, quirk, s);
    else launch_sweep_yz_m<SrrmhdM, 5>(a, quirk, s);
  }
}

template <class M, int ORDER, int MODE>
void launch_sweep_yzf_o(const SweepArgs &a, int quirk, hipStream_t s) {
  const int tiles = ((a.geo.n[1] + 15) / 16) * ((a.geo.n[2] + 15) / 16) * a.geo.n[0];
  if (quirk)
    hipLaunchKernelGGL((k_sweep_yzf<M, ORDER, MODE, true>), dim3(tiles), dim3(kThreads), 0, s, a);
  else
    hipLaunchKernelGGL((k_sweep_yzf<M, ORDER, MODE, false>), dim3(tiles), dim3(kThreads), 0, s, a);
}

template <class M, int ORDER>
void launch_sweep_yzf_m(const SweepArgs &a, int mode, int quirk, hipStream_t s) {
  switch (mode) {
    case kSweepAdd: launch_sweep_yzf_o<M, ORDER, kSweepAdd>(a, quirk, s); break;
    case kSweepStage1: launch_sweep_yzf_o<M, ORDER, kSweepStage1>(a, quirk, s); break;
    case kSweepStage2: launch_sweep_yzf_o<M, ORDER, kSweepStage2>(a, quirk, s); break;
    default:
      if constexpr (M::NC == kSrrmhdCons) {
        if (mode == kSweepImex2) launch_sweep_yzf_o<M, ORDER, kSweepImex2>(a, quirk, s);
        else if (mode == kSweepImexFinal) launch_sweep_yzf_o<M, ORDER, kSweepImexFinal>(a, quirk, s);
      }
      break;
  }
}

void launch_sweep_yzf(const SweepArgs &a, int model, int order, int mode, int quirk, hipStream_t s) {
  if (model == 0) {
    if (order == 3) launch_sweep_yzf_m<SrmhdM, 3>(a, mode, quirk, s);
    else launch_sweep_yzf_m<SrmhdM, 5>(a, mode, quirk, s);
  } else {
    if (order == 3) launch_sweep_yzf_m<SrrmhdM, 3>(a, mode, quirk, s);
    else launch_sweep_yzf_m<SrrmhdM, 5>(a, mode, quirk, s);
  }
}

template <class M, int ORDER, int MODE>
void launch_sweep_xyf_o(const SweepArgs &a, hipStream_t s) {
  const int tiles = ((a.geo.n[0] + 15) / 16) * ((a.geo.n[1] + 15) / 16);
  hipLaunchKernelGGL((k_sweep_xyf<M, ORDER, MODE>), dim3(tiles), dim3(kThreads), 0, s, a);
}

template <class M, int ORDER>
void launch_sweep_xyf_m(const SweepArgs &a, int mode, hipStream_t s) {
  switch (mode) {
    case kSweepAdd: launch_sweep_xyf_o<M, ORDER, kSweepAdd>(a, s); break;
    case kSweepStage1: launch_sweep_xyf_o<M, ORDER, kSweepStage1>(a, s); break;
    case kSweepStage2: launch_sweep_xyf_o<M, ORDER, kSweepStage2>(a, s); break;
    default:
      if constexpr (M::NC == kSrrmhdCons) {
        if (mode == kSweepImex2) launch_sweep_xyf_o<M, ORDER, kSweepImex2>(a, s);
        else if (mode == kSweepImexFinal) launch_sweep_xyf_o<M, ORDER, kSweepImexFinal>(a, s);
      }
      break;
  }
}

void launch_sweep_xyf(const SweepArgs &a, int model, int order, int mode, hipStream_t s) {
  if (model == 0) {
    if (order == 3) launch_sweep_xyf_m<SrmhdM, 3>(a, mode, s);
    else launch_sweep_xyf_m<SrmhdM, 5>(a, mode, s);
  } else {
    if (order == 3) launch_sweep_xyf_m<SrrmhdM, 3>(a, mode, s);
    else launch_sweep_xyf_m<SrrmhdM, 5>(a, mode, s);
  }
}

void launch_sweep(const SweepArgs &a, int model, int order, int dir, int mode, int quirk, hipStream_t s) {
  if (model == 2) {
    if (order == 3)
      launch_sweep_m<StoredM<SrmhdM>, 3>(a, dir, mode, quirk, s);
    else
      launch_sweep_m<StoredM<SrmhdM>, 5>(a, dir, mode, quirk, s);
  } else if (model == 3) {
    if (order == 3)
      launch_sweep_m<StoredM<SrrmhdM>, 3>(a, dir, mode, quirk, s);
    else
      launch_sweep_m<StoredM<SrrmhdM>, 5>(a, dir, mode, quirk, s);
  } else if (model == 0) {
    if (order == 3)
      launch_sweep_m<SrmhdM, 3>(a, dir, mode, quirk, s);
    else
      launch_sweep_m<SrmhdM, 5>(a, dir, mode, quirk, s);
  } else {
    if (order == 3)
      launch_sweep_m<SrrmhdM, 3>(a, dir, mode, quirk, s);
    else
      launch_sweep_m<SrrmhdM, 5>(a, dir, mode, quirk, s);
  }
}

void launch_prims_to_all_srmhd(const Geo &g, const double *V, double *U, double *A, double gam, hipStream_t s) {
  hipLaunchKernelGGL(k_prims_to_all_srmhd, dim3(blocks_for(g.vstride)), dim3(kThreads), 0, s, g, V, U, A, gam);
}

void launch_aux_srmhd(const Geo &g, const double *U, const double *root, double *A, double gam, hipStream_t s) {
  hipLaunchKernelGGL(k_aux_srmhd, dim3(blocks_for(g.vstride)), dim3(kThreads), 0, s, g, U, root, A, gam);
}

void launch_aux_srrmhd(const Geo &g, const double *U, const double *root, double *A, double gam, double sigma,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_aux_srrmhd, dim3(blocks_for(g.vstride)), dim3(kThreads), 0, s, g, U, root, A, gam, sigma);
}

void launch_c2p_srmhd(const Geo &g, const double *U, double *V, double *A, double gam, C2PCounters *st,
                      int *fail_count, long long *fail_list, unsigned char *fail_flag, int cold,
                      unsigned long long *hist, hipStream_t s) {
  hipLaunchKernelGGL(k_c2p_srmhd, dim3(blocks_for(g.vstride)), dim3(kThreads), 0, s, g, U, V, A, gam, st, fail_count,
                     fail_list, fail_flag, cold, hist);
}

void launch_c2p_retry_srmhd(const Geo &g, const double *U, double *V, double *A, double *root, double gam,
                            const int *fail_count, const long long *fail_list, unsigned char *fail_flag,
                            long long last_cell, C2PCounters *st, hipStream_t s) {
  hipLaunchKernelGGL(k_c2p_retry_srmhd, dim3(64), dim3(kThreads), 0, s, g, U, V, A, root, gam, fail_count,
                     fail_list, fail_flag, last_cell, st);
  hipLaunchKernelGGL(k_clear_fail_flags, dim3(64), dim3(kThreads), 0, s, fail_count, fail_list, fail_flag);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, const_cast<int *>(fail_count));
}

void launch_max_wavespeed_srmhd(const Geo &g, const double *V, const double *A, unsigned long long *out,
                                hipStream_t s) {
  const long long blocks = std::min<long long>(blocks_for(g.vstride), 4096);
  hipLaunchKernelGGL(k_max_wavespeed_srmhd, dim3((unsigned)blocks), dim3(kThreads), 0, s, g, V, A, out);
}

void launch_split_srmhd(const Geo &g, double *U, double *V, double *A, double *root, double dt, double gam,
                        double cp, C2PCounters *st, int *fail_count, long long *fail_list, unsigned char *fail_flag,
                        const double *ext, hipStream_t s) {
  hipLaunchKernelGGL(k_split_srmhd, dim3(blocks_for(g.vstride)), dim3(kThreads), 0, s, g, U, V, A, root, dt, gam,
                     cp, st, fail_count, fail_list, fail_flag, ext);
}

void launch_split_srrmhd(const Geo &g, double *U, double *V, double *A, double dt, double gam, double sigma,
                         double cp, C2PCounters *st, hipStream_t s) {
  hipLaunchKernelGGL(k_split_srrmhd, dim3(blocks_for(g.vstride)), dim3(kThreads), 0, s, g, U, V, A, dt, gam, sigma,
                     cp, st);
}

void launch_rkplus_stage(const RkStageArgs &a, int model, hipStream_t s) {
  const dim3 grid(blocks_for(a.geo.vstride)), block(kThreads);
  if (model == 0)
    hipLaunchKernelGGL(k_rkplus_stage<SrmhdM>, grid, block, 0, s, a);
  else
    hipLaunchKernelGGL(k_rkplus_stage<SrrmhdM>, grid, block, 0, s, a);
}

void launch_flux_stored(const Geo &g, int model, const double *U, const double *V, const double *A, double *F,
                        int quirk, hipStream_t s) {
  hipLaunchKernelGGL(k_flux_stored, dim3(blocks_for(g.vstride)), dim3(kThreads), 0, s, g, model, U, V, A, F, quirk);
}

void launch_rk410_prepare2(const Geo &g, int ncons, double *U1, double *U2, hipStream_t s) {
  const long long n = (long long)ncons * g.vstride;
  hipLaunchKernelGGL(k_rk410_prepare2, dim3(blocks_for(n)), dim3(kThreads), 0, s, n, U1, U2);
}

void launch_zero_int(int *p, hipStream_t s) { hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, p); }

void launch_compact(const Geo &g, const PadGeo &pg, int nvar, const double *src, double *dst, hipStream_t s) {
  hipLaunchKernelGGL(k_compact, dim3(blocks_for(g.vstride)), dim3(kThreads), 0, s, g, pg, nvar, src, dst);
}

void launch_expand(const Geo &g, const PadGeo &pg, int nvar, const double *src, double *dst, int ghosts_zero,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_expand, dim3(blocks_for(pg.ntot)), dim3(kThreads), 0, s, g, pg, nvar, src, dst, ghosts_zero);
}

void launch_imex1_srrmhd(const SweepArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_imex1_srrmhd, dim3(blocks_for(a.geo.vstride)), dim3(kThreads), 0, s, a);
}

void launch_prims_to_all_srrmhd(const Geo &g, const double *V, double *U, double *A, double gam, double sigma,
                                hipStream_t s) {
  hipLaunchKernelGGL(k_prims_to_all_srrmhd, dim3(blocks_for(g.vstride)), dim3(kThreads), 0, s, g, V, U, A, gam,
                     sigma);
}

void launch_c2p_srrmhd(const Geo &g, const double *U, double *V, double *A, double gam, double sigma,
                       C2PCounters *st, hipStream_t s) {
  hipLaunchKernelGGL(k_c2p_srrmhd, dim3(blocks_for(g.vstride)), dim3(kThreads), 0, s, g, U, V, A, gam, sigma, st);
}

void launch_source_srrmhd(const Geo &g, const PadGeo &pg, const double *U, const double *V, double *dst,
                          double gam, double sigma, double cp, hipStream_t s) {
  hipLaunchKernelGGL(k_source_srrmhd, dim3(blocks_for(pg.ntot)), dim3(kThreads), 0, s, g, pg, U, V, dst, gam, sigma,
                     cp);
}

void launch_flux_vector_srrmhd(const Geo &g, const PadGeo &pg, const double *U, const double *V, double *dst,
                               int dir, int quirk, double gam, double sigma, hipStream_t s) {
  const dim3 grid(blocks_for(pg.ntot)), block(kThreads);
  if (dir == 0)
    hipLaunchKernelGGL((k_flux_vector_srrmhd<0, false>), grid, block, 0, s, g, pg, U, V, dst, gam, sigma);
  else if (dir == 1)
    hipLaunchKernelGGL((k_flux_vector_srrmhd<1, false>), grid, block, 0, s, g, pg, U, V, dst, gam, sigma);
  else if (quirk)
    hipLaunchKernelGGL((k_flux_vector_srrmhd<2, true>), grid, block, 0, s, g, pg, U, V, dst, gam, sigma);
  else
    hipLaunchKernelGGL((k_flux_vector_srrmhd<2, false>), grid, block, 0, s, g, pg, U, V, dst, gam, sigma);
}

void launch_source_srmhd(const Geo &g, const PadGeo &pg, const double *U, double *dst, double cp, hipStream_t s) {
  hipLaunchKernelGGL(k_source_srmhd, dim3(blocks_for(pg.ntot)), dim3(kThreads), 0, s, g, pg, U, dst, cp);
}

void launch_flux_vector_srmhd(const Geo &g, const PadGeo &pg, const double *U, const double *V, double *dst, int dir,
                              int quirk, hipStream_t s) {
  const dim3 grid(blocks_for(pg.ntot)), block(kThreads);
  if (dir == 0)
    hipLaunchKernelGGL((k_flux_vector_srmhd<0, false>), grid, block, 0, s, g, pg, U, V, dst);
  else if (dir == 1)
    hipLaunchKernelGGL((k_flux_vector_srmhd<1, false>), grid, block, 0, s, g, pg, U, V, dst);
  else if (quirk)
    hipLaunchKernelGGL((k_flux_vector_srmhd<2, true>), grid, block, 0, s, g, pg, U, V, dst);
  else
    hipLaunchKernelGGL((k_flux_vector_srmhd<2, false>), grid, block, 0, s, g, pg, U, V, dst);
}

}  // namespace mgpu
