// The REGIME resistive extension of ideal SRMHD (REGIME.cc; Wright & Hawke
// 2019) on the GPU: its source term, added by RKSplit::setSource
// (rkSplit.cc:15-25) and RKPlus::rhs (RKPlus.cc:26-38).
//
// The reference forms the source on the ghost-padded grid with loop ranges
// that leave zeros near the grid edge (q outside [2, N-2), the E-flux
// reconstruction outside [shift, N-shift), K outside [1, N-1)). Those zeros
// reach the first and last interior cells through the centred differences,
// so the kernels here work on the same padded index space: a padded cell
// reads its primitives from the interior cell Bcs::apply copies there, and
// every range test is the reference's. The quantities, over padded cells:
//   E = -v x B                                    (REGIME.cc:150-154)
//   q = 4th-order centred div E where it fits     (REGIME.cc:186-212)
//   per axis, the WENO face flux of (E, B-flux)   (REGIME.cc:293-370)
//   K, then per axis D = M K, M = -df/dw dw/ds_b  (REGIME.cc:226-290,551-774)
//   source = centred difference of D, interior    (REGIME.cc:92-138)
// The default plan is two launches: k_regime_fq (every axis' face fluxes
// and q, E formed from the primitives where it is used), then k_regime_ds
// (K and D on chip, differenced into the source). The field-keeping plan
// stores each quantity from its own launch for mgpu_regime_fields.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mgpu_device.hpp"
#include "mgpu_kernels.hpp"

// No implicit a*b+c contraction in this file's arithmetic: whether the
// compiler fuses a product into an add depends on the surrounding code, and
// the one-launch and two-launch forms of D must round identically. (The
// shared WENO above keeps its explicit fma form.)
#pragma clang fp contract(off)

namespace mgpu {
namespace {

constexpr int kRegimeThreads = 256;

// The grid spacings as the reference's Data holds them, and 1 / (2 d) for
// the centred differences of K and of D (the reference divides by 2 d; the
// product differs from that quotient by at most one rounding).
struct Spacing {
  double d[3];
  double ih2[3];
};

// n / d for 0 <= n < 2^31 by a multiply-high and a shift: with l the
// smallest integer with d <= 2^l and m = ceil(2^(31+l) / d) (< 2^32 for
// d >= 2), n m = 2^(31+l) (n/d + delta) with 0 <= delta < 2^-l <= 1/d, so
// (n m) >> (31 + l) is floor(n / d) exactly (Granlund & Montgomery 1994,
// sec. 4). d == 1 passes n through.
struct FastDiv {
  unsigned d, m, s;  // s = l - 1
};

FastDiv make_fastdiv(unsigned d) {
  FastDiv f = {d, 0u, 0u};
  if (d < 2) return f;
  unsigned l = 0;
  while ((1ull << l) < d) ++l;
  f.m = (unsigned)(((1ull << (31 + l)) + d - 1) / d);
  f.s = l - 1;
  return f;
}

__device__ __forceinline__ unsigned fdiv(unsigned n, const FastDiv &f) {
  return f.d < 2 ? n : __umulhi(n, f.m) >> f.s;
}

// Divisors of the padded index: one plane (N1 N2) and one row (N2).
struct PadDiv {
  FastDiv plane, row;
};

// (i, j, k) of padded cell p (the padded grid holds fewer than 2^31 cells).
__device__ __forceinline__ void padded_pos(const PadDiv &dv, long long p, int pos[3]) {
  const unsigned u = (unsigned)p;
  const unsigned i = fdiv(u, dv.plane);
  const unsigned r = u - i * dv.plane.d;
  const unsigned j = fdiv(r, dv.row);
  pos[0] = (int)i;
  pos[1] = (int)j;
  pos[2] = (int)(r - j * dv.row.d);
}

// interior cell whose values Bcs::apply copies into padded position pos
__device__ __forceinline__ long long image_of(const Geo &g, const PadGeo &pg, const int pos[3]) {
  long long q = 0;
  for (int ax = 0; ax < 3; ++ax) q += (long long)remap(pos[ax] - pg.off[ax], g.n[ax], g.periodic[ax]) * g.stride[ax];
  return q;
}

__device__ __forceinline__ long long pstride(const PadGeo &pg, int ax) {
  return ax == 0 ? (long long)pg.N[1] * pg.N[2] : (ax == 1 ? pg.N[2] : 1);
}

__device__ __forceinline__ void load_prims(const double *V, long long vs, long long q, double P[8]) {
#pragma unroll
  for (int v = 0; v < 8; ++v) P[v] = V[v * vs + q];
}

// E = -v x B from one cell's primitives (REGIME.cc:150-154). Every kernel
// forms E through this function, so a value formed on the fly equals the
// stored one bit for bit.
__device__ __forceinline__ void regime_e_of(const double P[8], double E[3]) {
  E[0] = -(P[2] * P[7] - P[3] * P[6]);
  E[1] = -(P[3] * P[5] - P[1] * P[7]);
  E[2] = -(P[1] * P[6] - P[2] * P[5]);
}

__global__ __launch_bounds__(kRegimeThreads) void k_regime_e(Geo g, PadGeo pg, PadDiv dv, const double *V, double *E) {
  const long long p = (long long)xcd_block(blockIdx.x, gridDim.x) * kRegimeThreads + threadIdx.x;
  if (p >= pg.ntot) return;
  int pos[3];
  padded_pos(dv, p, pos);
  double P[8], e[3];
  load_prims(V, g.vstride, image_of(g, pg, pos), P);
  regime_e_of(P, e);
  E[p] = e[0];
  E[pg.ntot + p] = e[1];
  E[2 * pg.ntot + p] = e[2];
}

__global__ __launch_bounds__(kRegimeThreads) void k_regime_q(Geo g, PadGeo pg, PadDiv dv, Spacing sp, const double *E,
                                                             double *Q, int run) {
  const long long p = (long long)chunk_block(blockIdx.x, gridDim.x, run) * kRegimeThreads + threadIdx.x;
  if (p >= pg.ntot) return;
  int pos[3];
  padded_pos(dv, p, pos);
  bool in = true;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
    if (ax < g.dims) in = in && pos[ax] >= 2 && pos[ax] < pg.N[ax] - 2;
  double acc = 0.0;
  if (in) {
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      if (ax >= g.dims) break;
      const double *Ea = E + ax * pg.ntot;
      const long long s = pstride(pg, ax);
      const double t = (-Ea[p + 2 * s] + 8 * Ea[p + s] - 8 * Ea[p - s] + Ea[p - 2 * s]) / (12 * sp.d[ax]);
      acc = ax == 0 ? t : acc + t;
    }
  }
  Q[p] = acc;
}

// B-flux of the stiff E equations along axis ax (REGIME.cc:299-303,324-328,349-353)
__device__ __forceinline__ void stiff_flux(int ax, const double P[8], double fb[3]) {
  if (ax == 0) {
    fb[0] = 0.0;
    fb[1] = P[7];
    fb[2] = -P[6];
  } else if (ax == 1) {
    fb[0] = -P[7];
    fb[1] = 0.0;
    fb[2] = P[5];
  } else {
    fb[0] = P[6];
    fb[1] = -P[5];
    fb[2] = 0.0;
  }
}

// FVS::fluxReconstruction of (cons = E, f = fb) along axis AX with 3
// variables (fluxVectorSplitting.cc:4-56): the face value at the left face of
// each padded cell with [shift, N - shift) along AX, zero elsewhere. The axis
// is a template parameter so every per-axis array index is a constant.
template <int ORDER, int AX>
__global__ __launch_bounds__(kRegimeThreads) void k_regime_flux(Geo g, PadGeo pg, PadDiv dv, const double *V, const double *E,
                                                                double *F, int run) {
  constexpr int R = Weno<ORDER>::R;
  const long long p = (long long)chunk_block(blockIdx.x, gridDim.x, run) * kRegimeThreads + threadIdx.x;
  if (p >= pg.ntot) return;
  int pos[3];
  padded_pos(dv, p, pos);
  if (pos[AX] < R || pos[AX] >= pg.N[AX] - R) {
    for (int v = 0; v < 3; ++v) F[v * pg.ntot + p] = 0.0;
    return;
  }
  const long long s = pstride(pg, AX);
  double fp[3][2 * R], fm[3][2 * R];
#pragma unroll
  for (int m = 0; m < 2 * R; ++m) {
    const int cp[3] = {pos[0] + (AX == 0 ? m - R : 0), pos[1] + (AX == 1 ? m - R : 0),
                       pos[2] + (AX == 2 ? m - R : 0)};
    double P[8], fb[3];
    load_prims(V, g.vstride, image_of(g, pg, cp), P);
    stiff_flux(AX, P, fb);
    const long long pc = p + (m - R) * s;
    for (int v = 0; v < 3; ++v) {
      const double e = E[v * pg.ntot + pc];
      fp[v][m] = 0.5 * (fb[v] + e);
      fm[v][m] = 0.5 * (fb[v] - e);
    }
  }
  for (int v = 0; v < 3; ++v)
    F[v * pg.ntot + p] = Weno<ORDER>::face([&](int m) { return fp[v][m]; }, [&](int m) { return fm[v][m]; });
}

// Tile of padded cells per block of k_regime_fq (the same shapes as the
// source's tiles).
template <int DIMS>
struct FqTile {
  static constexpr int T[3] = {DIMS == 3 ? 4 : (DIMS == 2 ? 16 : 256), DIMS == 3 ? 8 : (DIMS == 2 ? 16 : 1),
                               DIMS == 3 ? 8 : 1};
  static_assert(T[0] * T[1] * T[2] == kRegimeThreads, "one padded cell per thread");
};

// Stencil offsets along an axis of k_regime_fq: the face's -R .. R - 1 and
// q's fourth-order -2 .. 2.
template <int ORDER>
struct FqHalo {
  static constexpr int R = Weno<ORDER>::R;
  static constexpr int LO = R > 2 ? R : 2, HI = R - 1 > 2 ? R - 1 : 2;
};

// Cells of a tile extended by the halo along axis AX.
template <int ORDER, int DIMS, int AX>
struct FqExt {
  static constexpr int D0 = FqTile<DIMS>::T[0] + (AX == 0 ? FqHalo<ORDER>::LO + FqHalo<ORDER>::HI : 0);
  static constexpr int D1 = FqTile<DIMS>::T[1] + (AX == 1 ? FqHalo<ORDER>::LO + FqHalo<ORDER>::HI : 0);
  static constexpr int D2 = FqTile<DIMS>::T[2] + (AX == 2 ? FqHalo<ORDER>::LO + FqHalo<ORDER>::HI : 0);
  static constexpr int kCells = D0 * D1 * D2;
};

template <int ORDER, int DIMS>
constexpr int fq_lds_cells() {
  int m = FqExt<ORDER, DIMS, 0>::kCells;
  if (DIMS > 1 && FqExt<ORDER, DIMS, 1>::kCells > m) m = FqExt<ORDER, DIMS, 1>::kCells;
  if (DIMS > 2 && FqExt<ORDER, DIMS, 2>::kCells > m) m = FqExt<ORDER, DIMS, 2>::kCells;
  return m;
}

// One axis of k_regime_fq: E and B of the tile extended along AX go to LDS
// (E formed from the primitives by regime_e_of), then each thread forms the
// face flux of its cell along AX into F's rows for AX (zero outside
// [R, N - R) along AX, as k_regime_flux) and, when its cell is inside
// [2, N - 2) along every axis (qin), adds the AX term of q to qacc as
// k_regime_q sums it.
template <int ORDER, int DIMS, int AX, int NL>
__device__ __forceinline__ void regime_fq_phase(const Geo &g, const PadGeo &pg, const Spacing &sp, const double *V,
                                                const int o[3], const int c[3], bool live, bool qin, double &qacc,
                                                double *F, double (&s_eb)[6][NL]) {
  using X = FqExt<ORDER, DIMS, AX>;
  constexpr int R = FqHalo<ORDER>::R, LO = FqHalo<ORDER>::LO;
  for (int idx = threadIdx.x; idx < X::kCells; idx += kRegimeThreads) {
    const int l2 = idx % X::D2, l1 = (idx / X::D2) % X::D1, l0 = idx / (X::D2 * X::D1);
    int cp[3] = {o[0] + l0, o[1] + l1, o[2] + l2};
    cp[AX] -= LO;
    bool in = true;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) in = in && cp[ax] >= 0 && cp[ax] < pg.N[ax];
    double e[3] = {0.0, 0.0, 0.0}, B[3] = {0.0, 0.0, 0.0};
    if (in) {
      double P[8];
      load_prims(V, g.vstride, image_of(g, pg, cp), P);
      regime_e_of(P, e);
      B[0] = P[5];
      B[1] = P[6];
      B[2] = P[7];
    }
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      s_eb[v][idx] = e[v];
      s_eb[3 + v][idx] = B[v];
    }
  }
  __syncthreads();
  if (live) {
    const int p0[3] = {o[0] + c[0], o[1] + c[1], o[2] + c[2]};
    const long long p = ((long long)p0[0] * pg.N[1] + p0[1]) * pg.N[2] + p0[2];
    const bool fin = p0[AX] >= R && p0[AX] < pg.N[AX] - R;
    // LDS slot of the cell at offset m - LO along AX from this thread's
    constexpr int kStep = AX == 0 ? X::D1 * X::D2 : (AX == 1 ? X::D2 : 1);
    const int base = (c[0] * X::D1 + c[1]) * X::D2 + c[2];  // offset -LO
    const long long nt = pg.ntot;
    if (fin) {
      double fp[3][2 * R], fm[3][2 * R];
#pragma unroll
      for (int m = 0; m < 2 * R; ++m) {
        const int sl = base + (m + LO - R) * kStep;
        double P[8] = {0.0, 0.0, 0.0, 0.0, 0.0, s_eb[3][sl], s_eb[4][sl], s_eb[5][sl]}, fb[3];
        stiff_flux(AX, P, fb);
#pragma unroll
        for (int v = 0; v < 3; ++v) {
          const double e = s_eb[v][sl];
          fp[v][m] = 0.5 * (fb[v] + e);
          fm[v][m] = 0.5 * (fb[v] - e);
        }
      }
#pragma unroll
      for (int v = 0; v < 3; ++v)
        F[((long long)AX * 3 + v) * nt + p] =
            Weno<ORDER>::face([&](int m) { return fp[v][m]; }, [&](int m) { return fm[v][m]; });
    } else {
#pragma unroll
      for (int v = 0; v < 3; ++v) F[((long long)AX * 3 + v) * nt + p] = 0.0;
    }
    if (qin) {
      const double *ea = &s_eb[AX][base + LO * kStep];  // E_AX at this cell
      const double t = (-ea[2 * kStep] + 8 * ea[kStep] - 8 * ea[-kStep] + ea[-2 * kStep]) / (12 * sp.d[AX]);
      qacc = AX == 0 ? t : qacc + t;
    }
  }
  __syncthreads();  // the next axis refills LDS
}

// The E-face fluxes of every active axis (REGIME.cc:293-370) and q
// (REGIME.cc:186-212) of each padded cell in one launch: the values
// k_regime_e, k_regime_q and the per-axis k_regime_flux launches give, bit
// for bit (E through regime_e_of, the same sums in the same order), without
// E going through HBM. A block takes a tile of padded cells and, one axis at
// a time, stages E and B of the tile extended by the stencil halo along that
// axis in LDS, so each cell's primitives are read about twice per axis
// instead of once per stencil point. The source's one-launch plan runs this,
// then k_regime_ds.
template <int ORDER, int DIMS>
__global__ __launch_bounds__(kRegimeThreads) void k_regime_fq(Geo g, PadGeo pg, Spacing sp, const double *V,
                                                              double *F, double *Q, int run) {
  using T = FqTile<DIMS>;
  constexpr int NL = fq_lds_cells<ORDER, DIMS>();
  __shared__ double s_eb[6][NL];
  const int nt0 = (pg.N[0] + T::T[0] - 1) / T::T[0], nt1 = (pg.N[1] + T::T[1] - 1) / T::T[1],
            nt2 = (pg.N[2] + T::T[2] - 1) / T::T[2];
  const int lb = chunk_block(blockIdx.x, gridDim.x, run);
  const int t2 = lb % nt2, t1 = (lb / nt2) % nt1, t0 = lb / (nt2 * nt1);
  if (t0 >= nt0) return;  // whole block: no barrier is skipped by part of it
  const int o[3] = {t0 * T::T[0], t1 * T::T[1], t2 * T::T[2]};
  const int tid = threadIdx.x;
  const int c[3] = {tid / (T::T[2] * T::T[1]), (tid / T::T[2]) % T::T[1], tid % T::T[2]};
  bool live = true, qin = true;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) live = live && o[ax] + c[ax] < pg.N[ax];
#pragma unroll
  for (int ax = 0; ax < DIMS; ++ax) qin = qin && o[ax] + c[ax] >= 2 && o[ax] + c[ax] < pg.N[ax] - 2;
  double qacc = 0.0;
  regime_fq_phase<ORDER, DIMS, 0, NL>(g, pg, sp, V, o, c, live, qin, qacc, F, s_eb);
  if constexpr (DIMS > 1) regime_fq_phase<ORDER, DIMS, 1, NL>(g, pg, sp, V, o, c, live, qin, qacc, F, s_eb);
  if constexpr (DIMS > 2) regime_fq_phase<ORDER, DIMS, 2, NL>(g, pg, sp, V, o, c, live, qin, qacc, F, s_eb);
  if (live) Q[((long long)(o[0] + c[0]) * pg.N[1] + o[1] + c[1]) * pg.N[2] + o[2] + c[2]] = qin ? qacc : 0.0;
}

// dw/ds_b (12 x 3, rows 0, 4 and 11 zero) times alpha (REGIME.cc:226-290)
__device__ __forceinline__ void regime_dwdsb(double q, double sig, const double P[8], const double E[3], double alpha,
                                             double dw[12][3]) {
  const double qsigsq = q * q + sig * sig, sigcu = sig * sig * sig, qch = q;
  for (int l = 0; l < 12; ++l)
    for (int m = 0; m < 3; ++m) dw[l][m] = 0.0;
  dw[1][0] = -qch * (qch * qch + sig * sig * (1 + P[5] * P[5]));
  dw[1][1] = -sig * (P[5] * P[6] * qch * sig - P[7] * qsigsq);
  dw[1][2] = -sig * (P[5] * P[7] * qch * sig + P[6] * qsigsq);
  dw[2][0] = -sig * (P[5] * P[6] * qch * sig + P[7] * qsigsq);
  dw[2][1] = -qch * (qch * qch + sig * sig * (1 + P[6] * P[6]));
  dw[2][2] = -sig * (P[6] * P[7] * qch * sig - P[5] * qsigsq);
  dw[3][0] = -sig * (P[5] * P[7] * qch * sig - P[6] * qsigsq);
  dw[3][1] = -sig * (P[6] * P[7] * qch * sig + P[5] * qsigsq);
  dw[3][2] = -qch * (qch * qch + sig * sig * (1 + P[7] * P[7]));
  dw[5][0] = -P[5] * sigcu * E[0];
  dw[5][1] = -P[6] * sigcu * E[0] - sig * qsigsq * P[3];
  dw[5][2] = -P[7] * sigcu * E[0] + sig * qsigsq * P[2];
  dw[6][0] = -P[5] * sigcu * E[1] + sig * qsigsq * P[3];
  dw[6][1] = -P[6] * sigcu * E[1];
  dw[6][2] = -P[7] * sigcu * E[1] - sig * qsigsq * P[1];
  dw[7][0] = -P[5] * sigcu * E[2] - sig * qsigsq * P[2];
  dw[7][1] = -P[6] * sigcu * E[2] + sig * qsigsq * P[1];
  dw[7][2] = -P[7] * sigcu * E[2];
  dw[8][0] = -P[5] * P[5] * sigcu - sig * qsigsq;
  dw[8][1] = -P[5] * P[6] * sigcu;
  dw[8][2] = -P[5] * P[7] * sigcu;
  dw[9][0] = -P[6] * P[5] * sigcu;
  dw[9][1] = -P[6] * P[6] * sigcu - sig * qsigsq;
  dw[9][2] = -P[6] * P[7] * sigcu;
  dw[10][0] = -P[7] * P[5] * sigcu;
  dw[10][1] = -P[7] * P[6] * sigcu;
  dw[10][2] = -P[7] * P[7] * sigcu - sig * qsigsq;
  for (int l = 0; l < 12; ++l)
    for (int m = 0; m < 3; ++m) dw[l][m] *= alpha;
}

// Rows 1..4, columns 5..10 of df/dw along axis ax (REGIME.cc:404-548), and
// its (row, ax+1) entries of rows 0 and 4.
__device__ __forceinline__ void regime_dfdw(int ax, double gam, const double P[8], const double E[3], double J[5][6],
                                            double &j0, double &j4) {
  for (int l = 0; l < 5; ++l)
    for (int n = 0; n < 6; ++n) J[l][n] = 0.0;
  // column index n - 5 for w entries 5..10 (B, E)
  if (ax == 0) {
    j0 = P[0];
    J[1][0] = -P[5]; J[1][1] = P[6]; J[1][2] = P[7]; J[1][3] = -E[0]; J[1][4] = E[1]; J[1][5] = E[2];
    J[2][0] = -P[6]; J[2][1] = -P[5]; J[2][3] = -E[1]; J[2][4] = -E[0];
    J[3][0] = -P[7]; J[3][2] = -P[5]; J[3][3] = -E[2]; J[3][5] = -E[0];
    J[4][1] = -E[2]; J[4][2] = E[1]; J[4][4] = P[7]; J[4][5] = -P[6];
  } else if (ax == 1) {
    j0 = P[0];
    J[1][0] = -P[6]; J[1][1] = -P[5]; J[1][3] = -E[1]; J[1][4] = -E[0];
    J[2][0] = P[5]; J[2][1] = -P[6]; J[2][2] = P[7]; J[2][3] = E[0]; J[2][4] = -E[1]; J[2][5] = E[2];
    J[3][1] = -P[7]; J[3][2] = -P[6]; J[3][4] = -E[2]; J[3][5] = -E[1];
    J[4][0] = E[2]; J[4][2] = -E[0]; J[4][3] = -P[7]; J[4][5] = P[5];
  } else {
    j0 = P[0];
    J[1][0] = -P[7]; J[1][2] = -P[5]; J[1][3] = -E[2]; J[1][5] = -E[0];
    J[2][1] = -P[7]; J[2][2] = -P[6]; J[2][4] = -E[2]; J[2][5] = -E[1];
    J[3][0] = P[5]; J[3][1] = P[6]; J[3][2] = -P[7]; J[3][3] = E[0]; J[3][4] = E[1]; J[3][5] = -E[2];
    J[4][0] = -E[1]; J[4][1] = E[0]; J[4][3] = P[6]; J[4][4] = -P[5];
  }
  j4 = gam * P[4] / (gam - 1);
}

// K at padded cell p (REGIME.cc:293-370): centred differences of the E-face
// fluxes, zero outside [1, N - 1) along each axis.
__device__ __forceinline__ void regime_k(const Geo &g, const PadGeo &pg, const Spacing &sp, const double *F,
                                         long long p, const int pos[3], double K[3]) {
  const long long nt = pg.ntot;
  K[0] = K[1] = K[2] = 0.0;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    if (ax >= g.dims || pos[ax] < 1 || pos[ax] >= pg.N[ax] - 1) continue;
    const long long s = pstride(pg, ax);
    const double *Fa = F + (long long)ax * 3 * nt;
    for (int v = 0; v < 3; ++v) K[v] += (Fa[v * nt + p + s] - Fa[v * nt + p - s]) * sp.ih2[ax];
  }
}

// The per-cell quantities D is formed from: primitives, E, q, alpha
// (REGIME.cc:141-224) and dw/ds_b (REGIME.cc:226-290).
struct RegimeCell {
  double P[8], Ec[3], q, alpha;
  double dw[12][3];
};

__device__ __forceinline__ void regime_cell(const Geo &g, const PadGeo &pg, const double *V, const double *Q,
                                            long long p, const int pos[3], double sig, RegimeCell &c) {
  load_prims(V, g.vstride, image_of(g, pg, pos), c.P);
  regime_e_of(c.P, c.Ec);
  c.q = Q[p];
  const double qs = c.q * c.q + sig * sig;
  c.alpha = 1 / (qs * (qs + sig * sig * (c.P[5] * c.P[5] + c.P[6] * c.P[6] + c.P[7] * c.P[7])));
  regime_dwdsb(c.q, sig, c.P, c.Ec, c.alpha, c.dw);
}

// D = M K along axis AX with M as REGIME::set_Dx/Dy/Dz form it
// (REGIME.cc:551-774): M = -df/dw dw/ds_b.
template <int AX>
__device__ __forceinline__ void regime_d_axis(const RegimeCell &c, const double K[3], double gam, double D[8]) {
  double J[5][6], j0, j4;
  regime_dfdw(AX, gam, c.P, c.Ec, J, j0, j4);
  constexpr int a = 1 + AX;
  double M[8][3];
  for (int l = 0; l < 8; ++l)
    for (int m = 0; m < 3; ++m) M[l][m] = 0.0;
  for (int m = 0; m < 3; ++m) {
    M[0][m] -= j0 * c.dw[a][m];
    M[4][m] -= j4 * c.dw[a][m];
  }
  // the induction rows with one unit entry each
  if (AX == 0) {
    for (int m = 0; m < 3; ++m) M[6][m] -= (-c.dw[10][m]);
    for (int m = 0; m < 3; ++m) M[7][m] -= c.dw[9][m];
  } else if (AX == 1) {
    for (int m = 0; m < 3; ++m) M[5][m] -= c.dw[10][m];
    for (int m = 0; m < 3; ++m) M[7][m] -= (-c.dw[8][m]);
  } else {
    for (int m = 0; m < 3; ++m) M[5][m] -= (-c.dw[9][m]);
    for (int m = 0; m < 3; ++m) M[6][m] -= c.dw[8][m];
  }
  for (int l = 1; l < 5; ++l)
    for (int m = 0; m < 3; ++m)
      for (int n = 0; n < 6; ++n) M[l][m] = fma(-J[l][n], c.dw[5 + n][m], M[l][m]);
  for (int l = 0; l < 8; ++l) D[l] = fma(M[l][2], K[2], fma(M[l][1], K[1], M[l][0] * K[0]));
}

// K (REGIME.cc:293-370), then D = M K per active axis, into the work arrays
// (the fields the C ABI reads back, and the two-launch source).
__global__ __launch_bounds__(kRegimeThreads) void k_regime_d(Geo g, PadGeo pg, PadDiv dv, Spacing sp, const double *V,
                                                             const double *Q, const double *F,
                                                             double *D, double *Kout, double *Aout, double gam,
                                                             double sig, int run) {
  const long long p = (long long)chunk_block(blockIdx.x, gridDim.x, run) * kRegimeThreads + threadIdx.x;
  if (p >= pg.ntot) return;
  const long long nt = pg.ntot;
  int pos[3];
  padded_pos(dv, p, pos);
  double K[3];
  regime_k(g, pg, sp, F, p, pos, K);
  RegimeCell c;
  regime_cell(g, pg, V, Q, p, pos, sig, c);
  for (int v = 0; v < 3; ++v) Kout[v * nt + p] = K[v];
  Aout[p] = c.alpha;
  double Dl[8];
  regime_d_axis<0>(c, K, gam, Dl);
  for (int l = 0; l < 8; ++l) D[l * nt + p] = Dl[l];
  if (g.dims > 1) {
    regime_d_axis<1>(c, K, gam, Dl);
    for (int l = 0; l < 8; ++l) D[(8 + l) * nt + p] = Dl[l];
  }
  if (g.dims > 2) {
    regime_d_axis<2>(c, K, gam, Dl);
    for (int l = 0; l < 8; ++l) D[(16 + l) * nt + p] = Dl[l];
  }
}

// D at padded position ppos along axis AX, zero outside the padded grid.
template <int AX>
__device__ __forceinline__ void regime_d_at(const Geo &g, const PadGeo &pg, const Spacing &sp, const double *V,
                                            const double *Q, const double *F, const int ppos[3],
                                            double gam, double sig, double D[8]) {
  bool in = true;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) in = in && ppos[ax] >= 0 && ppos[ax] < pg.N[ax];
  if (!in) {
    for (int l = 0; l < 8; ++l) D[l] = 0.0;
    return;
  }
  const long long p = ((long long)ppos[0] * pg.N[1] + ppos[1]) * pg.N[2] + ppos[2];
  double K[3];
  regime_k(g, pg, sp, F, p, ppos, K);
  RegimeCell c;
  regime_cell(g, pg, V, Q, p, ppos, sig, c);
  regime_d_axis<AX>(c, K, gam, D);
}

// Interior cells per block of the fused D + source kernel, by grid dimension.
template <int DIMS>
struct DsTile {
  static constexpr int T0 = DIMS == 3 ? 4 : (DIMS == 2 ? 16 : 256);
  static constexpr int T1 = DIMS == 3 ? 8 : (DIMS == 2 ? 16 : 1);
  static constexpr int T2 = DIMS == 3 ? 8 : 1;
  // D along axis a is needed on the tile widened by one cell each way along a
  static constexpr int kItems0 = (T0 + 2) * T1 * T2;
  static constexpr int kItems1 = DIMS > 1 ? T0 * (T1 + 2) * T2 : 0;
  static constexpr int kItems2 = DIMS > 2 ? T0 * T1 * (T2 + 2) : 0;
  static constexpr int kItems = kItems0 + kItems1 + kItems2;
  static_assert(T0 * T1 * T2 == kRegimeThreads, "one interior cell per thread");
};

// D = M K and the source's centred differences of it in one launch: each
// block forms D along every axis on its tile widened by one cell along that
// axis, in LDS, then differences it (REGIME.cc:92-138), so D, K and alpha
// never go to HBM. LDS holds, per axis a, D_a on the tile widened along a.
// Pass 1: every thread forms K and the cell quantities of its own tile cell
// once and D along each axis there. Pass 2: the cells one step outside the
// tile's faces, each only along the axis crossing that face. D is the same
// per-cell computation as k_regime_d's.
template <int DIMS>
__global__ __launch_bounds__(kRegimeThreads) void k_regime_ds(Geo g, PadGeo pg, Spacing sp, const double *V,
                                                              const double *Q, const double *F,
                                                              double *src, double gam, double sig, int run) {
  using T = DsTile<DIMS>;
  constexpr int T0 = T::T0, T1 = T::T1, T2 = T::T2;
  constexpr int B1 = T::kItems0, B2 = T::kItems0 + T::kItems1;  // first LDS slot of axes 1 and 2
  __shared__ double sD[8][T::kItems];
  const int nt0 = (g.n[0] + T0 - 1) / T0, nt1 = (g.n[1] + T1 - 1) / T1, nt2 = (g.n[2] + T2 - 1) / T2;
  const int lb = chunk_block(blockIdx.x, gridDim.x, run);
  const int t2 = lb % nt2, t1 = (lb / nt2) % nt1, t0 = lb / (nt2 * nt1);
  if (t0 >= nt0) return;  // whole block: no barrier is skipped by part of it
  const int o[3] = {t0 * T0 + pg.off[0], t1 * T1 + pg.off[1], t2 * T2 + pg.off[2]};  // padded tile origin
  const int tid = threadIdx.x;
  const int c2 = tid % T2, c1 = (tid / T2) % T1, c0 = tid / (T2 * T1);
  // LDS slot of widened-tile cell (i0, i1, i2) of axis 0, 1, 2
  auto slot0 = [&](int i0, int i1, int i2) { return (i0 * T1 + i1) * T2 + i2; };
  auto slot1 = [&](int i0, int i1, int i2) { return B1 + (i0 * (T1 + 2) + i1) * T2 + i2; };
  auto slot2 = [&](int i0, int i1, int i2) { return B2 + (i0 * T1 + i1) * (T2 + 2) + i2; };
  {
    // pass 1: the thread's own tile cell, D along every axis
    const int pp[3] = {o[0] + c0, o[1] + c1, o[2] + c2};
    bool in = true;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) in = in && pp[ax] < pg.N[ax];
    double D0[8] = {}, D1[8] = {}, D2[8] = {};
    if (in) {
      const long long p = ((long long)pp[0] * pg.N[1] + pp[1]) * pg.N[2] + pp[2];
      double K[3];
      regime_k(g, pg, sp, F, p, pp, K);
      RegimeCell c;
      regime_cell(g, pg, V, Q, p, pp, sig, c);
      regime_d_axis<0>(c, K, gam, D0);
      if (DIMS > 1) regime_d_axis<(DIMS > 1 ? 1 : 0)>(c, K, gam, D1);
      if (DIMS > 2) regime_d_axis<(DIMS > 2 ? 2 : 0)>(c, K, gam, D2);
    }
    for (int l = 0; l < 8; ++l) sD[l][slot0(c0 + 1, c1, c2)] = D0[l];
    if (DIMS > 1)
      for (int l = 0; l < 8; ++l) sD[l][slot1(c0, c1 + 1, c2)] = D1[l];
    if (DIMS > 2)
      for (int l = 0; l < 8; ++l) sD[l][slot2(c0, c1, c2 + 1)] = D2[l];
  }
  {
    // pass 2: the face neighbours of the tile, one axis each
    constexpr int H0 = 2 * T1 * T2, H1 = DIMS > 1 ? 2 * T0 * T2 : 0, H2 = DIMS > 2 ? 2 * T0 * T1 : 0;
    for (int h = tid; h < H0 + H1 + H2; h += kRegimeThreads) {
      double D[8];
      int slot;
      if (h < H0) {
        const int side = h / (T1 * T2), r = h % (T1 * T2), i1 = r / T2, i2 = r % T2;
        const int pp[3] = {o[0] + (side ? T0 : -1), o[1] + i1, o[2] + i2};
        regime_d_at<0>(g, pg, sp, V, Q, F, pp, gam, sig, D);
        slot = slot0(side ? T0 + 1 : 0, i1, i2);
      } else if (DIMS > 1 && h < H0 + H1) {
        const int hh = h - H0, side = hh / (T0 * T2), r = hh % (T0 * T2), i0 = r / T2, i2 = r % T2;
        const int pp[3] = {o[0] + i0, o[1] + (side ? T1 : -1), o[2] + i2};
        regime_d_at<(DIMS > 1 ? 1 : 0)>(g, pg, sp, V, Q, F, pp, gam, sig, D);
        slot = slot1(i0, side ? T1 + 1 : 0, i2);
      } else {
        const int hh = h - H0 - H1, side = hh / (T0 * T1), r = hh % (T0 * T1), i0 = r / T1, i1 = r % T1;
        const int pp[3] = {o[0] + i0, o[1] + i1, o[2] + (side ? T2 : -1)};
        regime_d_at<(DIMS > 2 ? 2 : 0)>(g, pg, sp, V, Q, F, pp, gam, sig, D);
        slot = slot2(i0, i1, side ? T2 + 1 : 0);
      }
      for (int l = 0; l < 8; ++l) sD[l][slot] = D[l];
    }
  }
  __syncthreads();
  const int c[3] = {o[0] - pg.off[0] + c0, o[1] - pg.off[1] + c1, o[2] - pg.off[2] + c2};
  if (c[0] >= g.n[0] || c[1] >= g.n[1] || c[2] >= g.n[2]) return;
  // LDS slots of the cell's two neighbours along each axis
  const int lo[3] = {slot0(c0, c1, c2), slot1(c0, c1, c2), slot2(c0, c1, c2)};
  const int hi[3] = {slot0(c0 + 2, c1, c2), slot1(c0, c1 + 2, c2), slot2(c0, c1, c2 + 2)};
  const long long q = c[0] * g.stride[0] + c[1] * g.stride[1] + (long long)c[2];
  for (int v = 0; v < 8; ++v) {
    double acc = 0.0;
#pragma unroll
    for (int ax = 0; ax < DIMS; ++ax) {
      const double t = (sD[v][hi[ax]] - sD[v][lo[ax]]) * sp.ih2[ax];
      acc = ax == 0 ? t : acc + t;
    }
    src[v * g.vstride + q] = acc;
  }
}

// sourceExtension over interior cells (REGIME.cc:92-138): variables 0..7.
__global__ __launch_bounds__(kRegimeThreads) void k_regime_source(Geo g, PadGeo pg, Spacing sp, const double *D,
                                                                  double *src, int run) {
  const long long q = (long long)chunk_block(blockIdx.x, gridDim.x, run) * kRegimeThreads + threadIdx.x;
  if (q >= g.vstride) return;
  const long long nt = pg.ntot;
  const int c[3] = {(int)(q / g.stride[0]), (int)((q / g.stride[1]) % g.n[1]), (int)(q % g.n[2])};
  const long long p =
      ((long long)(c[0] + pg.off[0]) * pg.N[1] + (c[1] + pg.off[1])) * pg.N[2] + (c[2] + pg.off[2]);
  for (int v = 0; v < 8; ++v) {
    double acc = 0.0;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      if (ax >= g.dims) break;
      const long long s = pstride(pg, ax);
      const double *Dv = D + ((long long)ax * 8 + v) * nt;
      const double t = (Dv[p + s] - Dv[p - s]) * sp.ih2[ax];
      acc = ax == 0 ? t : acc + t;
    }
    src[v * g.vstride + q] = acc;
  }
}

// dw/ds_b of every padded cell into dw[12*3][cell] (REGIME.cc:226-290), from
// the fields k_regime_d left (q, alpha).
__global__ __launch_bounds__(kRegimeThreads) void k_regime_dwdsb(Geo g, PadGeo pg, PadDiv dv, const double *V, const double *E,
                                                                 const double *Q, const double *A, double sig,
                                                                 double *out) {
  const long long p = (long long)xcd_block(blockIdx.x, gridDim.x) * kRegimeThreads + threadIdx.x;
  if (p >= pg.ntot) return;
  const long long nt = pg.ntot;
  int pos[3];
  padded_pos(dv, p, pos);
  double P[8];
  load_prims(V, g.vstride, image_of(g, pg, pos), P);
  const double Ec[3] = {E[p], E[nt + p], E[2 * nt + p]};
  double dw[12][3];
  regime_dwdsb(Q[p], sig, P, Ec, A[p], dw);
  for (int l = 0; l < 12; ++l)
    for (int m = 0; m < 3; ++m) out[(l * 3 + m) * nt + p] = dw[l][m];
}

unsigned blocks(long long n) { return (unsigned)((n + kRegimeThreads - 1) / kRegimeThreads); }

PadDiv pad_div(const PadGeo &pg) { return {make_fastdiv((unsigned)(pg.N[1] * pg.N[2])), make_fastdiv((unsigned)pg.N[2])}; }

}  // namespace

// E (3), q, face fluxes (3 axes x 3), D (3 axes x 8), K (3), alpha
size_t regime_work_doubles(const PadGeo &pg) { return (size_t)pg.ntot * (3 + 1 + 3 * 3 + 3 * 8 + 3 + 1); }

RegimeFields regime_fields_of(const PadGeo &pg, double *work) {
  const long long nt = pg.ntot;
  RegimeFields f;
  f.E = work;
  f.q = f.E + 3 * nt;
  f.F = f.q + nt;
  f.D = f.F + 9 * nt;
  f.K = f.D + 24 * nt;
  f.alpha = f.K + 3 * nt;
  return f;
}

void launch_regime_dwdsb(const Geo &g, const PadGeo &pg, const double *V, const double *work, double sigma,
                         double *out, hipStream_t s) {
  const RegimeFields f = regime_fields_of(pg, const_cast<double *>(work));
  const PadDiv dv = pad_div(pg);
  hipLaunchKernelGGL(k_regime_dwdsb, dim3(blocks(pg.ntot)), dim3(kRegimeThreads), 0, s, g, pg, dv, V, f.E, f.q, f.alpha,
                     sigma, out);
}

void launch_regime_source(const Geo &g, const PadGeo &pg, const double dx[3], const double *V, double *work,
                          double *src, double gam, double sigma, int order, bool fused, hipStream_t s) {
  const long long nt = pg.ntot;
  const Spacing sp = {{dx[0], dx[1], dx[2]}, {1.0 / (2 * dx[0]), 1.0 / (2 * dx[1]), 1.0 / (2 * dx[2])}};
  const RegimeFields fl = regime_fields_of(pg, work);
  double *E = fl.E, *Q = fl.q, *F = fl.F, *D = fl.D;
  const PadDiv dv = pad_div(pg);
  // an eighth of the blocks of one padded (stencil kernels) or interior
  // plane: each XCD then keeps the same rows of every plane, so both its
  // x and its y neighbours stay in its L2
  const int run_pad = (int)std::max<long long>(1, (long long)pg.N[1] * pg.N[2] / kRegimeThreads / 8);
  const int run_int = (int)std::max<long long>(1, (long long)g.n[1] * g.n[2] / kRegimeThreads / 8);
  const dim3 grid(blocks(nt)), block(kRegimeThreads);
  if (fused) {
    // face fluxes and q in one launch, E formed where it is used
    {
      const int n1 = (pg.N[1] + 7) / 8, n2 = (pg.N[2] + 7) / 8;  // 3-D tiles 4 x 8 x 8
      if (g.dims == 3) {
        const unsigned nb = (unsigned)(((pg.N[0] + 3) / 4) * n1 * n2);
        const int run = std::max(1, n1 * n2 / 8);
        if (order == 3) hipLaunchKernelGGL((k_regime_fq<3, 3>), dim3(nb), block, 0, s, g, pg, sp, V, F, Q, run);
        else hipLaunchKernelGGL((k_regime_fq<5, 3>), dim3(nb), block, 0, s, g, pg, sp, V, F, Q, run);
      } else if (g.dims == 2) {
        const int m1 = (pg.N[1] + 15) / 16;
        const unsigned nb = (unsigned)(((pg.N[0] + 15) / 16) * m1);
        const int run = std::max(1, m1 / 8);
        if (order == 3) hipLaunchKernelGGL((k_regime_fq<3, 2>), dim3(nb), block, 0, s, g, pg, sp, V, F, Q, run);
        else hipLaunchKernelGGL((k_regime_fq<5, 2>), dim3(nb), block, 0, s, g, pg, sp, V, F, Q, run);
      } else {
        const unsigned nb = (unsigned)((pg.N[0] + 255) / 256);
        if (order == 3) hipLaunchKernelGGL((k_regime_fq<3, 1>), dim3(nb), block, 0, s, g, pg, sp, V, F, Q, 1);
        else hipLaunchKernelGGL((k_regime_fq<5, 1>), dim3(nb), block, 0, s, g, pg, sp, V, F, Q, 1);
      }
    }
    if (g.dims == 3) {
      using T = DsTile<3>;
      const unsigned nb = (unsigned)(((g.n[0] + T::T0 - 1) / T::T0) * ((g.n[1] + T::T1 - 1) / T::T1) *
                                     ((g.n[2] + T::T2 - 1) / T::T2));
      const int run = std::max(1, ((g.n[1] + T::T1 - 1) / T::T1) * ((g.n[2] + T::T2 - 1) / T::T2) / 8);
      hipLaunchKernelGGL(k_regime_ds<3>, dim3(nb), block, 0, s, g, pg, sp, V, Q, F, src, gam, sigma, run);
    } else if (g.dims == 2) {
      using T = DsTile<2>;
      const unsigned nb = (unsigned)(((g.n[0] + T::T0 - 1) / T::T0) * ((g.n[1] + T::T1 - 1) / T::T1));
      const int run = std::max(1, ((g.n[1] + T::T1 - 1) / T::T1) / 8);
      hipLaunchKernelGGL(k_regime_ds<2>, dim3(nb), block, 0, s, g, pg, sp, V, Q, F, src, gam, sigma, run);
    } else {
      using T = DsTile<1>;
      const unsigned nb = (unsigned)((g.n[0] + T::T0 - 1) / T::T0);
      hipLaunchKernelGGL(k_regime_ds<1>, dim3(nb), block, 0, s, g, pg, sp, V, Q, F, src, gam, sigma, 1);
    }
    return;
  }
  // the field-keeping plan: E, q and each axis' face fluxes as separate
  // launches that store them, then D, K and alpha stored, then the source
  hipLaunchKernelGGL(k_regime_e, grid, block, 0, s, g, pg, dv, V, E);
  hipLaunchKernelGGL(k_regime_q, grid, block, 0, s, g, pg, dv, sp, E, Q, run_pad);
  if (order == 3) {
    hipLaunchKernelGGL((k_regime_flux<3, 0>), grid, block, 0, s, g, pg, dv, V, E, F, run_pad);
    if (g.dims > 1) hipLaunchKernelGGL((k_regime_flux<3, 1>), grid, block, 0, s, g, pg, dv, V, E, F + 3 * nt, run_pad);
    if (g.dims > 2) hipLaunchKernelGGL((k_regime_flux<3, 2>), grid, block, 0, s, g, pg, dv, V, E, F + 6 * nt, run_pad);
  } else {
    hipLaunchKernelGGL((k_regime_flux<5, 0>), grid, block, 0, s, g, pg, dv, V, E, F, run_pad);
    if (g.dims > 1) hipLaunchKernelGGL((k_regime_flux<5, 1>), grid, block, 0, s, g, pg, dv, V, E, F + 3 * nt, run_pad);
    if (g.dims > 2) hipLaunchKernelGGL((k_regime_flux<5, 2>), grid, block, 0, s, g, pg, dv, V, E, F + 6 * nt, run_pad);
  }
  hipLaunchKernelGGL(k_regime_d, dim3(blocks(nt)), dim3(kRegimeThreads), 0, s, g, pg, dv, sp, V, Q, F, D, fl.K,
                     fl.alpha, gam, sigma, run_pad);
  hipLaunchKernelGGL(k_regime_source, dim3(blocks(g.vstride)), dim3(kRegimeThreads), 0, s, g, pg, sp, D, src,
                     run_int);
}

}  // namespace mgpu
