// Device-side numerics of the MI355X time-step engine: grid geometry with
// boundary index remapping, WENO3/WENO5 reconstruction and the ideal SRMHD
// model (flux vector, prims -> all, Newton cons2prim). FP64 throughout.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mgpu {

// Interior-only device layout [var][i][j][k], k fastest. 2-D grids keep a
// unit z extent. Ghost cells are never stored: a read outside the interior is
// redirected to the interior cell the reference's Bcs::apply would have
// copied there (periodic wrap or outflow clamp, boundaryConds.cc:5-145,590-730).
struct Geo {
  int n[3];             // interior extents; n[2] == 1 in 2-D
  long long stride[3];  // element strides of i, j, k
  long long vstride;    // elements per variable plane = n0*n1*n2
  int periodic[3];      // 1 periodic, 0 outflow
  double inv_d[3];      // 1/dx, 1/dy, 1/dz
  int dims;
};

// Logical block of hardware block bid out of nb. Blocks are dispatched round
// robin over the 8 XCDs; this gives each XCD a contiguous run of logical
// blocks, so work on neighbouring cells shares that XCD's L2.
__device__ __forceinline__ int xcd_block(int bid, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int x = bid & 7, local = bid >> 3;
  return x * q + (x < r ? x : r) + local;
}

// Logical block of hardware block bid out of nb for grids whose neighbours
// lie both nearby (rows) and far apart (planes): the 8 XCDs sweep the grid
// together in chunks of 8 * run logical blocks, each XCD taking the same
// contiguous run of `run` blocks of every chunk. With a chunk of one plane an
// XCD keeps the same rows of every plane, so the neighbours along rows and
// across planes are both in its own L2. Blocks past the last whole chunk
// keep their own index.
__device__ __forceinline__ int chunk_block(int bid, int nb, int run) {
  const int chunk = 8 * run, full = nb / chunk * chunk;
  if (bid >= full) return bid;
  const int x = bid & 7, local = bid >> 3;
  return local / run * chunk + x * run + local % run;
}

__device__ __forceinline__ int remap(int i, int n, int periodic) {
  if (i < 0) return periodic ? i + n : 0;
  if (i >= n) return periodic ? i - n : n - 1;
  return i;
}

// ------------------------------------------------------------------ WENO
// Jiang-Shu WENO with nonlinear weights alpha_k = d_k / (eps + beta_k^2),
// eps = 1e-16, as in the reference (wenoUpwinds.cc:5-84). The normalised
// weights are formed with one division: sum_k d_k q_k / s_k over sum_k d_k / s_k
// is multiplied through by s_0 s_1 (s_2) and divided by d_0, which is the same
// quantity up to rounding. The indicators and candidate values are written in
// first differences of the stencil so that they share their terms.

constexpr double kWenoEps = 1e-16;

// Upwind reconstruction at the right face of v1 from (v0, v1, v2), as a
// numerator/denominator pair (value = n / d); the downwind value at the left
// face of v1 is the same function of (v2, v1, v0). Linear weights (1/3, 2/3)
// over 1/3 give t0 = s1, t1 = 2 s0; candidates q0 = v1 + (v1 - v0)/2 and
// q1 = (v1 + v2)/2, so t1 q1 = s0 (v1 + v2).
__device__ __forceinline__ void weno3_nd(double v0, double v1, double v2, double &n, double &d) {
  const double a0 = v1 - v0, a1 = v2 - v1;
  const double b0 = a0 * a0, b1 = a1 * a1;
  const double s0 = fma(b0, b0, kWenoEps), s1 = fma(b1, b1, kWenoEps);
  n = fma(s1, fma(0.5, a0, v1), s0 * (v1 + v2));
  d = fma(2.0, s0, s1);
}

// Upwind reconstruction at the right face of v2 from (v0..v4) as n / d. With
// the differences d01 = v0 - v1, d12, d23, d34, the second differences are
// a_k = d(k)(k+1) - d(k+1)(k+2) and the one-sided ones
//   e0 = v0 - 4 v1 + 3 v2 = d01 - 3 d12,  e1 = v1 - v3 = d12 + d23,
//   e2 = 3 v2 - 4 v3 + v4 = 3 d23 - d34,
// so 4 beta_k = 13/3 a_k^2 + e_k^2 (eps scaled by 16 to match, which leaves
// the weights unchanged). Linear weights (0.1, 0.6, 0.3) over 0.1 are
// (1, 6, 3); the candidates times those weights are
//   q0     = (2 v0 - 7 v1 + 11 v2)/6 = a0/3 + v2 - d12/2,
//   6 q1   = -v1 + 5 v2 + 2 v3      = 6 v2 - d12 - 2 d23,
//   3 q2   = (2 v2 + 5 v3 - v4)/2   = 3 v3 + d23 + d34/2.
__device__ __forceinline__ void weno5_nd(double v0, double v1, double v2, double v3, double v4, double &n,
                                         double &d) {
  constexpr double c13 = 13.0 / 3.0;
  constexpr double eps16 = 16.0 * kWenoEps;
  const double d01 = v0 - v1, d12 = v1 - v2, d23 = v2 - v3, d34 = v3 - v4;
  const double a0 = d01 - d12, a1 = d12 - d23, a2 = d23 - d34;
  const double e0 = fma(-3.0, d12, d01), e1 = d12 + d23, e2 = fma(3.0, d23, -d34);
  const double b0 = fma(c13 * a0, a0, e0 * e0);
  const double b1 = fma(c13 * a1, a1, e1 * e1);
  const double b2 = fma(c13 * a2, a2, e2 * e2);
  const double s0 = fma(b0, b0, eps16), s1 = fma(b1, b1, eps16), s2 = fma(b2, b2, eps16);
  const double q0 = fma(1.0 / 3.0, a0, fma(-0.5, d12, v2));
  const double q1 = fma(-2.0, d23, fma(6.0, v2, -d12));
  const double q2 = fma(3.0, v3, fma(0.5, d34, d23));
  const double p0 = s1 * s2, p1 = s0 * s2, p2 = s0 * s1;
  n = fma(p0, q0, fma(p1, q1, p2 * q2));
  d = fma(6.0, p1, fma(3.0, p2, p0));
}

// a / b for a positive normal b whose reciprocal is normal too (the WENO
// denominators, products of eps-bounded positive sums): the hardware
// reciprocal estimate refined by two Newton-Raphson steps, then one residual
// correction of the quotient. It skips the operand scaling and special-case
// fix-up of the general IEEE division and can differ from it in the last bit.
__device__ __forceinline__ double div_pos(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  r = fma(fma(-b, r, 1.0), r, r);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}

template <int ORDER>
struct Weno;

// Face value = upwind reconstruction of f+ plus downwind reconstruction of
// f-, combined over one division: (nu dd + nd du) / (du dd). The two halves
// are also available apart (up, down, combine) for kernels that stage f+ and
// f- through LDS one after the other; the arithmetic is the same.
__device__ __forceinline__ double weno_combine(double nu, double du, double nd, double dd) {
  return div_pos(fma(nu, dd, nd * du), du * dd);
}

template <>
struct Weno<3> {
  static constexpr int R = 2;  // stencil radius = (order + 1) / 2
  // Face between cells R-1 and R of p[0 .. 2R-1] (upwind on fp, downwind on fm).
  template <class LP>
  __device__ __forceinline__ static void up(LP fp, double &n, double &d) {
    weno3_nd(fp(0), fp(1), fp(2), n, d);
  }
  template <class LM>
  __device__ __forceinline__ static void down(LM fm, double &n, double &d) {
    weno3_nd(fm(3), fm(2), fm(1), n, d);
  }
  template <class LP, class LM>
  __device__ __forceinline__ static double face(LP fp, LM fm) {
    double nu, du, nd, dd;
    up(fp, nu, du);
    down(fm, nd, dd);
    return weno_combine(nu, du, nd, dd);
  }
};

template <>
struct Weno<5> {
  static constexpr int R = 3;
  template <class LP>
  __device__ __forceinline__ static void up(LP fp, double &n, double &d) {
    weno5_nd(fp(0), fp(1), fp(2), fp(3), fp(4), n, d);
  }
  template <class LM>
  __device__ __forceinline__ static void down(LM fm, double &n, double &d) {
    weno5_nd(fm(5), fm(4), fm(3), fm(2), fm(1), n, d);
  }
  template <class LP, class LM>
  __device__ __forceinline__ static double face(LP fp, LM fm) {
    double nu, du, nd, dd;
    up(fp, nu, du);
    down(fm, nd, dd);
    return weno_combine(nu, du, nd, dd);
  }
};

// ----------------------------------------------------------------- SRMHD
// Variable order (srmhd.cc:36-53): cons D, Sx, Sy, Sz, tau, Bx, By, Bz, phi;
// prims rho, vx, vy, vz, p, Bx, By, Bz; aux h, W, e, c, b0, bx, by, bz, bsq,
// vsq, BS, Bsq, Ssq.
constexpr int kSrmhdCons = 9, kSrmhdPrims = 8, kSrmhdAux = 13;

// Flux vector in direction DIR from the cell's cons and (vx, vy, vz, p)
// (SRMHD::fluxVector, srmhd.cc:72-206). W and b^mu are formed from the
// primitives as SRMHD::primsToAll does (srmhd.cc:596-622). QUIRK_PHI_Z: the
// z flux of phi is By as at srmhd.cc:200.
template <int DIR, bool QUIRK_PHI_Z>
__device__ __forceinline__ void srmhd_flux(const double u[9], double vx, double vy, double vz, double p,
                                           double f[9]) {
  const double Bx = u[5], By = u[6], Bz = u[7];
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double iW = sqrt(1.0 - v2);  // 1/W
  const double W = 1.0 / iW;
  const double b0 = W * (vx * Bx + vy * By + vz * Bz);
  const double bx = fma(Bx, iW, b0 * vx), by = fma(By, iW, b0 * vy), bz = fma(Bz, iW, b0 * vz);
  const double bsq = (Bx * Bx + By * By + Bz * Bz + b0 * b0) * (iW * iW);
  const double vd = DIR == 0 ? vx : (DIR == 1 ? vy : vz);
  const double Bd = DIR == 0 ? Bx : (DIR == 1 ? By : Bz);
  const double ptot = p + 0.5 * bsq;
  const double BdiW = Bd * iW;
  f[0] = u[0] * vd;
  f[1] = fma(u[1], vd, -bx * BdiW);
  f[2] = fma(u[2], vd, -by * BdiW);
  f[3] = fma(u[3], vd, -bz * BdiW);
  f[1 + DIR] += ptot;
  f[4] = fma(u[4] + ptot, vd, -b0 * BdiW);
  f[5] = DIR == 0 ? u[8] : fma(Bx, vd, -Bd * vx);
  f[6] = DIR == 1 ? u[8] : fma(By, vd, -Bd * vy);
  f[7] = DIR == 2 ? u[8] : fma(Bz, vd, -Bd * vz);
  f[8] = (DIR == 2 && QUIRK_PHI_Z) ? By : Bd;
}

// SRMHD::primsToAll for one cell (srmhd.cc:573-673).
__device__ __forceinline__ void srmhd_prims_to_all(const double P[8], double gam, double U[9], double A[13]) {
  U[5] = P[5];
  U[6] = P[6];
  U[7] = P[7];
  A[11] = P[5] * P[5] + P[6] * P[6] + P[7] * P[7];
  U[8] = 0.0;
  A[9] = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
  A[1] = 1.0 / sqrt(1.0 - A[9]);
  A[4] = A[1] * (P[1] * P[5] + P[2] * P[6] + P[3] * P[7]);
  A[5] = P[5] / A[1] + A[4] * P[1];
  A[6] = P[6] / A[1] + A[4] * P[2];
  A[7] = P[7] / A[1] + A[4] * P[3];
  A[8] = (A[11] + A[4] * A[4]) / (A[1] * A[1]);
  A[0] = 1.0 + P[4] / P[0] * (gam / (gam - 1.0));
  A[2] = P[4] / (P[0] * (gam - 1.0));
  A[3] = sqrt(A[2] * gam * (gam - 1.0) / A[0]);
  U[0] = P[0] * A[1];
  const double e = (P[0] * A[0] + A[8]) * A[1] * A[1];
  U[1] = e * P[1] - A[4] * A[5];
  U[2] = e * P[2] - A[4] * A[6];
  U[3] = e * P[3] - A[4] * A[7];
  A[12] = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
  A[10] = P[5] * U[1] + P[6] * U[2] + P[7] * U[3];
  U[4] = e - (P[4] + A[8] * 0.5) - A[4] * A[4] - U[0];
}

constexpr double kNewtonTol = 1e-12;
constexpr int kNewtonMaxIt = 60;

// Newton on the 2x2 cons2prim system of srmhd.cc:269-295 in x = (v^2, Z =
// rho h W^2), with its analytic Jacobian. Converged when the scaled residual
// max(|f0|/(Z+B^2)^2, |f1|/(Z+B^2)) <= 1e-12; steps leaving v^2 in [0,1) or
// Z > 0 are pulled back half way. Returns Newton steps taken or -1.
__device__ __forceinline__ int srmhd_newton(double D, double tau, double Bsq, double Ssq, double BS, double gam,
                                            double &v2, double &Z) {
  if (!(v2 >= 0.0 && v2 < 1.0)) v2 = 0.0;
  if (!(Z > 0.0)) Z = D + tau + Bsq;
  const double gf = (gam - 1.0) / gam;
  const double BS2 = BS * BS;
  for (int it = 0; it <= kNewtonMaxIt; ++it) {
    const double s = 1.0 - v2, sq = sqrt(s);
    const double p = gf * (Z * s - D * sq);
    const double ZB = Z + Bsq;
    const double iZ = 1.0 / Z;
    const double iZ2 = iZ * iZ;
    const double f0 = ZB * ZB * v2 - (2.0 * Z + Bsq) * BS2 * iZ2 - Ssq;
    const double f1 = ZB - p - 0.5 * Bsq * s - 0.5 * BS2 * iZ2 - D - tau;
    const double iZB = 1.0 / ZB;
    const double r = fmax(fabs(f0) * iZB * iZB, fabs(f1) * iZB);
    if (r <= kNewtonTol) return it;
    if (!(r == r) || it == kNewtonMaxIt) break;
    const double BS2iZ3 = BS2 * iZ2 * iZ;
    const double J00 = ZB * ZB;
    const double J01 = 2.0 * ZB * (v2 + BS2iZ3);
    const double J10 = gf * (Z - 0.5 * D / sq) + 0.5 * Bsq;
    const double J11 = 1.0 - gf * s + BS2iZ3;
    const double idet = 1.0 / (J00 * J11 - J01 * J10);
    const double d0 = (J11 * f0 - J01 * f1) * idet;
    const double d1 = (J00 * f1 - J10 * f0) * idet;
    double nv2 = v2 - d0, nZ = Z - d1;
    if (nv2 < 0.0) nv2 = 0.5 * v2;
    if (nv2 >= 1.0) nv2 = 0.5 * (v2 + 1.0);
    if (nZ <= 0.0) nZ = 0.5 * Z;
    v2 = nv2;
    Z = nZ;
  }
  return -1;
}

// Recovery of prims and aux from the root (srmhd.cc:509-557).
__device__ __forceinline__ void srmhd_recover(const double U[9], double gam, double v2, double Z, double P[8],
                                              double A[13]) {
  const double BS = U[5] * U[1] + U[6] * U[2] + U[7] * U[3];
  const double Bsq = U[5] * U[5] + U[6] * U[6] + U[7] * U[7];
  const double Ssq = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
  const double W = 1.0 / sqrt(1.0 - v2);
  const double rho = U[0] / W;
  const double h = Z / (rho * W * W);
  const double p = (h - 1.0) * rho * (gam - 1.0) / gam;
  const double e = p / (rho * (gam - 1.0));
  const double den = 1.0 / (Z * (Bsq + Z));
  const double vx = (U[5] * BS + U[1] * Z) * den;
  const double vy = (U[6] * BS + U[2] * Z) * den;
  const double vz = (U[7] * BS + U[3] * Z) * den;
  const double b0 = W * (U[5] * vx + U[6] * vy + U[7] * vz);
  const double iW = 1.0 / W;
  P[0] = rho;
  P[1] = vx;
  P[2] = vy;
  P[3] = vz;
  P[4] = p;
  P[5] = U[5];
  P[6] = U[6];
  P[7] = U[7];
  A[0] = h;
  A[1] = W;
  A[2] = e;
  A[3] = sqrt(e * gam * (gam - 1.0) / h);
  A[4] = b0;
  A[5] = U[5] * iW + b0 * vx;
  A[6] = U[6] * iW + b0 * vy;
  A[7] = U[7] * iW + b0 * vz;
  A[8] = (Bsq + b0 * b0) * (iW * iW);
  A[9] = vx * vx + vy * vy + vz * vz;
  A[10] = BS;
  A[11] = Bsq;
  A[12] = Ssq;
}

// Plain starting guess from the previous primitives (srmhd.cc:438-442), with
// h re-formed from p and rho as primsToAll does (srmhd.cc:625).
__device__ __forceinline__ void srmhd_plain_guess(double rho, double vx, double vy, double vz, double p, double gam,
                                                  double &v2, double &Z) {
  v2 = vx * vx + vy * vy + vz * vz;
  const double h = 1.0 + p / rho * (gam / (gam - 1.0));
  Z = rho * h / (1.0 - v2);
}

// Newton-step histogram of one block's cells: bin min(steps, 63), failed
// cells (steps < 0) in bin 63. Every thread of the block calls it; hist null
// (the histogram is off) makes it a no-op.
constexpr int kHistBins = 64;
__device__ __forceinline__ void hist_add_block(unsigned long long *hist, bool counted, int steps) {
  if (!hist) return;
  __shared__ unsigned s_hist[kHistBins];
  for (int b = threadIdx.x; b < kHistBins; b += blockDim.x) s_hist[b] = 0u;
  __syncthreads();
  if (counted) atomicAdd(&s_hist[steps < 0 || steps >= kHistBins - 1 ? kHistBins - 1 : steps], 1u);
  __syncthreads();
  for (int b = threadIdx.x; b < kHistBins; b += blockDim.x)
    if (s_hist[b]) atomicAdd(&hist[b], (unsigned long long)s_hist[b]);
}

// RK2 stage updates on one conserved value (RK2.cc:31-91), written with
// explicit fused multiply-adds so that every kernel rounds them alike:
// stage one U1 = U - dt F(U), stage two U = (U + U1 - dt F(U1)) / 2.
__device__ __forceinline__ double rk2_stage1(double u, double f, double dt) { return fma(-dt, f, u); }
__device__ __forceinline__ double rk2_stage2(double un, double u1, double f, double dt) {
  return 0.5 * fma(-dt, f, un + u1);
}

// Largest |characteristic speed| along one axis of an ideal SRMHD state: the
// fast magnetosonic speed taken as isotropic in the fluid frame,
//   a2 = cs2 + va2 - cs2 va2,  cs2 = c^2,  va2 = b2 / (rho h + b2),
// carried to the lab frame,
//   lambda = [v_d (1 - a2) +- sqrt(a2 (1 - v2) (1 - v2 a2 - v_d^2 (1 - a2)))] / (1 - v2 a2)
// (Gammie, McKinney & Toth 2003, ApJ 589, 444, sec. 3.3, in flat space). c, h,
// b2 and v2 are the SRMHD aux entries 3, 0, 8 and 9.
__device__ __host__ inline double srmhd_max_speed(double rho, double vd, double v2, double h, double c, double b2) {
  const double cs2 = c * c;
  const double va2 = b2 / (rho * h + b2);
  const double a2 = cs2 + va2 - cs2 * va2;
  const double den = 1.0 - v2 * a2;
  const double disc = a2 * (1.0 - v2) * (den - vd * vd * (1.0 - a2));
  const double sq = sqrt(disc > 0.0 ? disc : 0.0);
  const double b = vd * (1.0 - a2);
  return fmax(fabs(b + sq), fabs(b - sq)) / den;
}

}  // namespace mgpu
