// Reference-API operators outside the fused time step, on gfx950 (FP64):
//
// * FVS::fluxReconstruction (fluxVectorSplitting.cc:4-56) on given padded
//   arrays: the Lax-Friedrichs split of (f, cons) and the WENO face flux at
//   i - 1/2 along one axis, over the reference's loop range [shift, N - shift)
//   of that axis and every padded index of the others (weno.cc:23-118).
// * Model::getPrimitiveVarsSingleCell (srmhd.cc:297-388, srrmhd.cc:256-312)
//   and Model::sourceTermSingleCell (srmhd.cc:212-223, srrmhd.cc:194-207) on
//   lists of independent cells, each cell's variables contiguous as the
//   reference passes one cell ([cell][var]).
//
// The cell-list cons2prim writes one partial counter record per block and a
// one-block kernel sums them, so the counts need no atomics and no zeroed
// buffer.
#include <hip/hip_runtime.h>

#include <mutex>

#include "method_gpu.h"
#include "mgpu_kernels.hpp"

namespace mgpu {

namespace {

constexpr int kCellThreads = kCellBlock;

// Counter records of the allocation-free cell-list cons2prim: one per block
// of kNoAllocMaxCells cells, then the total. Every GPU holds its own copy for
// the life of the process; the host reaches it only through the address that
// noalloc_scratch_setup looked up on that GPU.
constexpr long long kNoAllocRecords = kNoAllocMaxCells / kCellBlock + 1;
static_assert(kNoAllocMaxCells % kCellBlock == 0, "whole blocks");
static __device__ __align__(256) CellCounters scratch[kNoAllocRecords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
CellCounters *g_scratch_ptr[kMaxDevices] = {};

// Sum over a block of one cell counter set; thread 0 gets the totals.
__device__ __forceinline__ void block_counts(int &cells, int &iters, int &fails, int &maxit) {
  __shared__ int s_red[kCellThreads / 64][4];
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
    cells = iters = fails = maxit = 0;
    for (int k = 0; k < kCellThreads / 64; ++k) {
      cells += s_red[k][0];
      iters += s_red[k][1];
      fails += s_red[k][2];
      maxit = max(maxit, s_red[k][3]);
    }
  }
}

__device__ __forceinline__ void store_partial(CellCounters *part, int cells, int iters, int fails, int maxit) {
  block_counts(cells, iters, fails, maxit);
  if (threadIdx.x == 0) {
    CellCounters c;
    c.cells = (unsigned long long)cells;
    c.iterations = (unsigned long long)iters;
    c.failures = (unsigned long long)fails;
    c.max_iterations = (unsigned long long)maxit;
    part[blockIdx.x] = c;
  }
}

// SRMHD::getPrimitiveVarsSingleCell per cell: Bx, By, Bz copied, BS, B^2, S^2
// formed, the start (v^2, rho h / (1 - v^2)) taken from the cell's prims and
// aux h (srmhd.cc:318-343), Newton to the 1e-12 scaled residual, then the
// recovery of srmhd.cc:351-386. A cell that does not converge still gets the
// last iterate's recovery and is counted as a failure.
__global__ __launch_bounds__(kCellThreads) void k_c2p_cells_srmhd(long long n, const double *cons, double *prims,
                                                                   double *aux, double gam, CellCounters *part) {
  const long long c = (long long)blockIdx.x * kCellThreads + threadIdx.x;
  int cells = 0, iters = 0, fails = 0, maxit = 0;
  if (c < n) {
    double u[kSrmhdCons];
#pragma unroll
    for (int v = 0; v < kSrmhdCons; ++v) u[v] = cons[c * kSrmhdCons + v];
    const double *P0 = prims + c * kSrmhdPrims;
    double v2 = P0[1] * P0[1] + P0[2] * P0[2] + P0[3] * P0[3];
    double Z = P0[0] * aux[c * kSrmhdAux] / (1.0 - v2);
    const double BS = u[5] * u[1] + u[6] * u[2] + u[7] * u[3];
    const double Bsq = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
    const double Ssq = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
    const int its = srmhd_newton(u[0], u[4], Bsq, Ssq, BS, gam, v2, Z);
    double P[kSrmhdPrims], A[kSrmhdAux];
    srmhd_recover(u, gam, v2, Z, P, A);
#pragma unroll
    for (int v = 0; v < kSrmhdPrims; ++v) prims[c * kSrmhdPrims + v] = P[v];
#pragma unroll
    for (int v = 0; v < kSrmhdAux; ++v) aux[c * kSrmhdAux + v] = A[v];
    cells = 1;
    if (its < 0) {
      fails = 1;
    } else {
      iters = its;
      maxit = its;
    }
  }
  store_partial(part, cells, iters, fails, maxit);
}

// SRRMHD::getPrimitiveVarsSingleCell per cell: the secant on Z from the
// cell's aux rhohWsq (srrmhd.cc:256-312), sigma constant.
__global__ __launch_bounds__(kCellThreads) void k_c2p_cells_srrmhd(long long n, const double *cons, double *prims,
                                                                    double *aux, double gam, double sigma,
                                                                    CellCounters *part) {
  const long long c = (long long)blockIdx.x * kCellThreads + threadIdx.x;
  int cells = 0, iters = 0, fails = 0, maxit = 0;
  if (c < n) {
    double u[kSrrmhdCons];
#pragma unroll
    for (int v = 0; v < kSrrmhdCons; ++v) u[v] = cons[c * kSrrmhdCons + v];
    const double Zg = aux[c * kSrrmhdAux + 10];
    double P[kSrrmhdPrims], A[kSrrmhdAux];
    const int its = srrmhd_c2p_cell(u, gam, sigma, Zg, P, A);
#pragma unroll
    for (int v = 0; v < kSrrmhdPrims; ++v) prims[c * kSrrmhdPrims + v] = P[v];
#pragma unroll
    for (int v = 0; v < kSrrmhdAux; ++v) aux[c * kSrrmhdAux + v] = A[v];
    cells = 1;
    if (its < 0) {
      fails = 1;
    } else {
      iters = its;
      maxit = its;
    }
  }
  store_partial(part, cells, iters, fails, maxit);
}

// Sum of nb per-block records into *total (one block; fixed order per lane,
// so the totals do not depend on scheduling).
__global__ __launch_bounds__(kCellThreads) void k_sum_partials(const CellCounters *part, int nb,
                                                                CellCounters *total) {
  unsigned long long c = 0, it = 0, f = 0, m = 0;
  for (int b = threadIdx.x; b < nb; b += kCellThreads) {
    const CellCounters p = part[b];
    c += p.cells;
    it += p.iterations;
    f += p.failures;
    m = p.max_iterations > m ? p.max_iterations : m;
  }
  __shared__ unsigned long long s[kCellThreads][4];
  s[threadIdx.x][0] = c;
  s[threadIdx.x][1] = it;
  s[threadIdx.x][2] = f;
  s[threadIdx.x][3] = m;
  __syncthreads();
  for (int h = kCellThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s[threadIdx.x][0] += s[threadIdx.x + h][0];
      s[threadIdx.x][1] += s[threadIdx.x + h][1];
      s[threadIdx.x][2] += s[threadIdx.x + h][2];
      s[threadIdx.x][3] = s[threadIdx.x + h][3] > s[threadIdx.x][3] ? s[threadIdx.x + h][3] : s[threadIdx.x][3];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    CellCounters t;
    t.cells = s[0][0];
    t.iterations = s[0][1];
    t.failures = s[0][2];
    t.max_iterations = s[0][3];
    *total = t;
  }
}

// Model::sourceTermSingleCell per cell. SRMHD: only phi, -phi / cp^2
// (srmhd.cc:212-223). SRRMHD: -J from the cell's aux, qch - psi / cp^2 and
// -phi / cp^2 (srrmhd.cc:194-207).
__global__ __launch_bounds__(kCellThreads) void k_source_cells(int model, long long n, const double *cons,
                                                                const double *aux, double cp, double *src) {
  const long long c = (long long)blockIdx.x * kCellThreads + threadIdx.x;
  if (c >= n) return;
  if (model == MGPU_SRMHD) {
    double *s = src + c * kSrmhdCons;
    for (int v = 0; v < 8; ++v) s[v] = 0.0;
    s[8] = -cons[c * kSrmhdCons + 8] / (cp * cp);
  } else {
    const double *u = cons + c * kSrrmhdCons;
    const double *A = aux + c * kSrrmhdAux;
    double *s = src + c * kSrrmhdCons;
    for (int v = 0; v < 8; ++v) s[v] = 0.0;
    s[8] = -A[4];
    s[9] = -A[5];
    s[10] = -A[6];
    s[11] = u[13] - u[11] / (cp * cp);
    s[12] = -u[12] / (cp * cp);
    s[13] = 0.0;
  }
}

// FVS::fluxReconstruction of variables [0, nvars) along axis DIR on padded
// arrays (variable stride pg.ntot): f+- = (f +- cons) / 2 at every stencil
// cell, the upwind WENO of f+ and the downwind WENO of f- at face i - 1/2,
// each with its own division, then their sum (fluxVectorSplitting.cc:24-50).
// Outside [R, N - R) along DIR the face flux is 0, as the reference's
// zero-initialised temporaries leave it.
template <int ORDER, int DIR>
__global__ __launch_bounds__(kCellThreads) void k_flux_recon(PadGeo pg, int nvars, const double *cons,
                                                              const double *f, double *frecon) {
  constexpr int R = Weno<ORDER>::R;
  const long long p = (long long)blockIdx.x * kCellThreads + threadIdx.x;
  if (p >= pg.ntot) return;
  const long long plane = (long long)pg.N[1] * pg.N[2];
  const int idx = DIR == 0 ? (int)(p / plane) : (DIR == 1 ? (int)((p / pg.N[2]) % pg.N[1]) : (int)(p % pg.N[2]));
  const int N = pg.N[DIR];
  const long long st = DIR == 0 ? plane : (DIR == 1 ? (long long)pg.N[2] : 1);
  const bool inside = idx >= R && idx < N - R;
  for (int v = 0; v < nvars; ++v) {
    double out = 0.0;
    if (inside) {
      const long long b = (long long)v * pg.ntot + p - (long long)R * st;
      auto fp = [&](int m) { return 0.5 * (f[b + m * st] + cons[b + m * st]); };
      auto fm = [&](int m) { return 0.5 * (f[b + m * st] - cons[b + m * st]); };
      double nu, du, nd, dd;
      Weno<ORDER>::up(fp, nu, du);
      Weno<ORDER>::down(fm, nd, dd);
      out = nu / du + nd / dd;
    }
    frecon[(long long)v * pg.ntot + p] = out;
  }
}

template <int ORDER>
void launch_flux_recon_o(const PadGeo &pg, int dir, int nvars, const double *cons, const double *f, double *frecon,
                         hipStream_t s) {
  const dim3 grid((unsigned)((pg.ntot + kCellThreads - 1) / kCellThreads));
  if (dir == 0)
    hipLaunchKernelGGL((k_flux_recon<ORDER, 0>), grid, dim3(kCellThreads), 0, s, pg, nvars, cons, f, frecon);
  else if (dir == 1)
    hipLaunchKernelGGL((k_flux_recon<ORDER, 1>), grid, dim3(kCellThreads), 0, s, pg, nvars, cons, f, frecon);
  else
    hipLaunchKernelGGL((k_flux_recon<ORDER, 2>), grid, dim3(kCellThreads), 0, s, pg, nvars, cons, f, frecon);
}

}  // namespace

long long cells_blocks(long long n) { return (n + kCellThreads - 1) / kCellThreads; }

int noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return 0;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != device) return -1;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
  g_scratch_ptr[device] = static_cast<CellCounters *>(p);
  return 0;
}

CellCounters *noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  return g_scratch_ptr[device];
}

void launch_flux_recon(const PadGeo &pg, int order, int dir, int nvars, const double *cons, const double *f,
                       double *frecon, hipStream_t s) {
  if (order == 3)
    launch_flux_recon_o<3>(pg, dir, nvars, cons, f, frecon, s);
  else
    launch_flux_recon_o<5>(pg, dir, nvars, cons, f, frecon, s);
}

void launch_c2p_cells(int model, long long n, const double *cons, double *prims, double *aux, double gam,
                      double sigma, CellCounters *partial, CellCounters *total, hipStream_t s) {
  const long long nb = cells_blocks(n);
  if (nb > 0) {
    if (model == MGPU_SRMHD)
      hipLaunchKernelGGL(k_c2p_cells_srmhd, dim3((unsigned)nb), dim3(kCellThreads), 0, s, n, cons, prims, aux, gam,
                         partial);
    else
      hipLaunchKernelGGL(k_c2p_cells_srrmhd, dim3((unsigned)nb), dim3(kCellThreads), 0, s, n, cons, prims, aux, gam,
                         sigma, partial);
  }
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kCellThreads), 0, s, (const CellCounters *)partial, (int)nb,
                     total);
}

void launch_source_cells(int model, long long n, const double *cons, const double *aux, double cp, double *src,
                         hipStream_t s) {
  const long long nb = cells_blocks(n);
  if (nb > 0)
    hipLaunchKernelGGL(k_source_cells, dim3((unsigned)nb), dim3(kCellThreads), 0, s, model, n, cons, aux, cp, src);
}

}  // namespace mgpu

