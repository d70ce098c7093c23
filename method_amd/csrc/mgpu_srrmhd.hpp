// Device-side resistive SRMHD (SRRMHD) model and the SSP2 IMEX implicit stage.
// Variable order (srrmhd.cc:36-63): cons D, Sx, Sy, Sz, tau, Bx, By, Bz, Ex,
// Ey, Ez, psi, phi, qch; prims rho, vx, vy, vz, p, Bx, By, Bz, Ex, Ey, Ez; aux
// h, W, e, c, Jx, Jy, Jz, Bsq, Esq, vsq, rhohWsq, vE, Sbarx, Sbary, Sbarz,
// Sbarsq, tauBar.
#pragma once

#include "mgpu_device.hpp"

namespace mgpu {

constexpr int kSrrmhdCons = 14, kSrrmhdPrims = 11, kSrrmhdAux = 17;

// Flux vector in direction DIR (SRRMHD::fluxVector, srrmhd.cc:65-192) from the
// cell's cons (D, S, E, B, psi, phi), velocity and pressure, Z = rho h W^2,
// hs = (B^2 + E^2) / 2 and the current J along DIR. QUIRK: in z, the Sy flux
// takes the Sz formula and the Sz flux is the y-direction Sz flux left in the
// reference's shared flux buffer (srrmhd.cc:163-171 after the y call).
template <int DIR, bool QUIRK>
__device__ __forceinline__ void srrmhd_flux_core(const double u[14], double vx, double vy, double vz, double p,
                                                 double Z, double hs, double J, double f[14]) {
  const double Bx = u[5], By = u[6], Bz = u[7], Ex = u[8], Ey = u[9], Ez = u[10];
  if (DIR == 0) {
    f[0] = u[0] * vx;
    f[1] = Z * vx * vx + p - Ex * Ex - Bx * Bx + hs;
    f[2] = Z * vx * vy - Ex * Ey - Bx * By;
    f[3] = Z * vx * vz - Ex * Ez - Bx * Bz;
    f[4] = u[1] - u[0] * vx;
    f[5] = u[12];
    f[6] = -Ez;
    f[7] = Ey;
    f[8] = u[11];
    f[9] = Bz;
    f[10] = -By;
    f[11] = Ex;
    f[12] = Bx;
    f[13] = J;
  } else if (DIR == 1) {
    f[0] = u[0] * vy;
    f[1] = Z * vx * vy - Ey * Ex - By * Bx;
    f[2] = Z * vy * vy + p - Ey * Ey - By * By + hs;
    f[3] = Z * vy * vz - Ey * Ez - By * Bz;
    f[4] = u[2] - u[0] * vy;
    f[5] = Ez;
    f[6] = u[12];
    f[7] = -Ex;
    f[8] = -Bz;
    f[9] = u[11];
    f[10] = Bx;
    f[11] = Ey;
    f[12] = By;
    f[13] = J;
  } else {
    f[0] = u[0] * vz;
    f[1] = Z * vx * vz - Ez * Ex - Bz * Bx;
    const double szz = Z * vz * vz + p - Ez * Ez - Bz * Bz + hs;
    const double syz = Z * vy * vz - Ey * Ez - By * Bz;
    f[2] = QUIRK ? szz : syz;
    f[3] = QUIRK ? syz : szz;
    f[4] = u[3] - u[0] * vz;
    f[5] = -Ey;
    f[6] = Ex;
    f[7] = u[12];
    f[8] = By;
    f[9] = -Bx;
    f[10] = u[11];
    f[11] = Ez;
    f[12] = Bz;
    f[13] = J;
  }
}


// The flux with Z, B^2, E^2 and J formed from the primitives as
// SRRMHD::primsToAll / getPrimitiveVarsSingleCell do (srrmhd.cc:303-311,380-408).
template <int DIR, bool QUIRK>
__device__ __forceinline__ void srrmhd_flux(const double u[14], double rho, double vx, double vy, double vz, double p,
                                            double gam, double sigma, double f[14]) {
  const double Bx = u[5], By = u[6], Bz = u[7], Ex = u[8], Ey = u[9], Ez = u[10];
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double W = 1.0 / sqrt(1.0 - v2);
  const double e = p / (rho * (gam - 1.0));
  const double h = 1.0 + e * gam;
  const double Z = rho * h * W * W;
  const double hs = 0.5 * (Bx * Bx + By * By + Bz * Bz + Ex * Ex + Ey * Ey + Ez * Ez);
  const double vE = vx * Ex + vy * Ey + vz * Ez;
  const double q = u[13];
  const double ws = W * sigma;
  const double J = DIR == 0   ? q * vx + ws * (Ex + (vy * Bz - vz * By) - vE * vx)
                   : DIR == 1 ? q * vy + ws * (Ey + (vz * Bx - vx * Bz) - vE * vy)
                              : q * vz + ws * (Ez + (vx * By - vy * Bx) - vE * vz);
  srrmhd_flux_core<DIR, QUIRK>(u, vx, vy, vz, p, Z, hs, J, f);
}

// The flux exactly as SRRMHD::fluxVector forms it from stored arrays: Z, B^2,
// E^2 and J taken from aux (rhohWsq, Bsq, Esq, J; srrmhd.cc:78-99), E and B from
// the cons, velocity and pressure from the prims. Equal to srrmhd_flux when the
// three arrays belong to one state.
template <int DIR, bool QUIRK>
__device__ __forceinline__ void srrmhd_flux_stored(const double u[14], const double P[11], const double A[17],
                                                   double f[14]) {
  srrmhd_flux_core<DIR, QUIRK>(u, P[1], P[2], P[3], P[4], A[10], 0.5 * (A[7] + A[8]), A[4 + DIR], f);
}

// Residual of the recovery equation in Z = rho h W^2 (srrmhd.cc:512-537).
__device__ __forceinline__ double srrmhd_c2p_residual(double Z, double SbSq, double D, double tauBar, double gam) {
  const double vsq = SbSq / (Z * Z);
  if (vsq >= 1.0 || Z < 0) return 1.0e6;
  const double W = 1.0 / sqrt(1.0 - vsq);
  const double rho = D / W;
  const double h = Z / (rho * W * W);
  const double p = (gam - 1.0) * (h - rho) / gam;
  if (rho < 0 || p < 0 || W < 1 || h < 0) return 1.0e6;
  return (1.0 - (gam - 1.0) / (W * W * gam)) * Z + ((gam - 1.0) / (W * gam) - 1.0) * D - tauBar;
}

// The reference's secant iteration on Z (srrmhd.cc:539-581): first step 1e-4,
// stop at |residual| < 1e-12, at most 50 iterations; returns iterations or -1
// (Z then holds the best iterate).
__device__ __forceinline__ int srrmhd_secant(double &Z, double SbSq, double D, double tauBar, double gam) {
  double x0 = Z, x1 = x0 + 1.0e-4;
  double f0 = srrmhd_c2p_residual(x0, SbSq, D, tauBar, gam);
  double f1 = srrmhd_c2p_residual(x1, SbSq, D, tauBar, gam);
  double bestX = x0, bestF = f0;
  for (int it = 0; it < 50; ++it) {
    if (fabs(f0) < 1.0e-12) {
      Z = x0;
      return it;
    }
    const double x2 = x1 - f1 * (x1 - x0) / (f1 - f0);
    x1 = x0;
    x0 = x2;
    f1 = f0;
    f0 = srrmhd_c2p_residual(x0, SbSq, D, tauBar, gam);
    if (fabs(f0) < fabs(bestF)) {
      bestX = x0;
      bestF = f0;
    }
  }
  Z = bestX;
  return -1;
}

// The aux entries that depend on the cons only (srrmhd.cc:512-538): B^2, E^2,
// S - E x B and its square, tau - (B^2 + E^2)/2.
__device__ __forceinline__ void srrmhd_aux_of_cons(const double u[14], double A[17]) {
  A[7] = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
  A[8] = u[8] * u[8] + u[9] * u[9] + u[10] * u[10];
  A[12] = u[1] - (u[9] * u[7] - u[10] * u[6]);
  A[13] = u[2] - (u[10] * u[5] - u[8] * u[7]);
  A[14] = u[3] - (u[8] * u[6] - u[9] * u[5]);
  A[15] = A[12] * A[12] + A[13] * A[13] + A[14] * A[14];
  A[16] = u[4] - 0.5 * (A[7] + A[8]);
}

// Prims and aux from the cons and the root Z = rho h W^2 (srrmhd.cc:539-581,
// after the secant): a pure function of (u, Z), so an aux formed later from a
// kept root equals the one the cons2prim would have written, bit for bit.
__device__ __forceinline__ void srrmhd_recover(const double u[14], double gam, double sigma, double Z, double P[11],
                                               double A[17]) {
  P[5] = u[5];
  P[6] = u[6];
  P[7] = u[7];
  P[8] = u[8];
  P[9] = u[9];
  P[10] = u[10];
  srrmhd_aux_of_cons(u, A);
  A[10] = Z;
  A[9] = A[15] / (Z * Z);
  A[1] = 1.0 / sqrt(1.0 - A[9]);
  P[0] = u[0] / A[1];
  A[0] = Z / (P[0] * A[1] * A[1]);
  A[2] = (A[0] - 1.0) / gam;
  A[3] = sqrt((A[2] * gam * (gam - 1.0)) / A[0]);
  P[4] = P[0] * A[2] * (gam - 1.0);
  const double iZ = 1.0 / Z;
  P[1] = A[12] * iZ;
  P[2] = A[13] * iZ;
  P[3] = A[14] * iZ;
  A[11] = P[1] * u[8] + P[2] * u[9] + P[3] * u[10];
  const double ws = A[1] * sigma;
  A[4] = u[13] * P[1] + ws * (u[8] + (P[2] * u[7] - P[3] * u[6]) - A[11] * P[1]);
  A[5] = u[13] * P[2] + ws * (u[9] + (P[3] * u[5] - P[1] * u[7]) - A[11] * P[2]);
  A[6] = u[13] * P[3] + ws * (u[10] + (P[1] * u[6] - P[2] * u[5]) - A[11] * P[3]);
}

// SRRMHD::getPrimitiveVarsSingleCell (srrmhd.cc:256-312) with guess Zg.
// Writes prims P[11] and aux A[17]; returns secant iterations or -1.
__device__ __forceinline__ int srrmhd_c2p_cell(const double u[14], double gam, double sigma, double Zg, double P[11],
                                               double A[17]) {
  double Ac[17];
  srrmhd_aux_of_cons(u, Ac);
  double Z = Zg;
  const int its = srrmhd_secant(Z, Ac[15], u[0], Ac[16], gam);
  srrmhd_recover(u, gam, sigma, Z, P, A);
  return its;
}

// Starting guess Z = rho h W^2 from previous primitives (the reference keeps
// it as aux rhohWsq and starts the secant there, srrmhd.cc:337).
__device__ __forceinline__ double srrmhd_guess_Z(double rho, double vx, double vy, double vz, double p, double gam) {
  const double W2 = 1.0 / (1.0 - (vx * vx + vy * vy + vz * vz));
  return (rho + p * gam / (gam - 1.0)) * W2;
}

// Non-zero source components 8..12 of SRRMHD::sourceTermSingleCell
// (srrmhd.cc:194-207): -J, qch - psi/cp^2, -phi/cp^2.
__device__ __forceinline__ void srrmhd_source5(const double u[14], const double A[17], double cp, double s[5]) {
  const double icp2 = 1.0 / (cp * cp);
  s[0] = -A[4];
  s[1] = -A[5];
  s[2] = -A[6];
  s[3] = u[13] - u[11] * icp2;
  s[4] = -u[12] * icp2;
}

// SRRMHD::primsToAll for one cell without the charge density and current,
// which need the neighbours' E (srrmhd.cc:358-413, 490-508).
__device__ __forceinline__ void srrmhd_prims_to_all_local(const double P[11], double gam, double U[14], double A[17]) {
  U[11] = U[12] = 0.0;
  U[5] = P[5];
  U[6] = P[6];
  U[7] = P[7];
  U[8] = P[8];
  U[9] = P[9];
  U[10] = P[10];
  A[7] = P[5] * P[5] + P[6] * P[6] + P[7] * P[7];
  A[8] = P[8] * P[8] + P[9] * P[9] + P[10] * P[10];
  A[9] = P[1] * P[1] + P[2] * P[2] + P[3] * P[3];
  A[11] = P[1] * U[8] + P[2] * U[9] + P[3] * U[10];
  A[1] = 1.0 / sqrt(1.0 - A[9]);
  U[0] = P[0] * A[1];
  A[2] = P[4] / (P[0] * (gam - 1.0));
  A[0] = 1.0 + A[2] * gam;
  A[3] = sqrt((A[2] * gam * (gam - 1.0)) / A[0]);
  A[10] = P[0] * A[0] * A[1] * A[1];
  U[1] = A[10] * P[1] + U[9] * U[7] - U[10] * U[6];
  U[2] = A[10] * P[2] + U[10] * U[5] - U[8] * U[7];
  U[3] = A[10] * P[3] + U[8] * U[6] - U[9] * U[5];
  U[4] = A[10] - P[4] + 0.5 * (A[7] + A[8]) - U[0];
  A[12] = A[10] * P[1];
  A[13] = A[10] * P[2];
  A[14] = A[10] * P[3];
  A[15] = A[12] * A[12] + A[13] * A[13] + A[14] * A[14];
  A[16] = A[10] - P[4] - U[0];
}

// Current J = qch v + W sigma (E + v x B - (v.E) v) (srrmhd.cc:459-488).
__device__ __forceinline__ void srrmhd_current(const double U[14], const double P[11], const double A[17],
                                               double sigma, double J[3]) {
  const double ws = A[1] * sigma;
  J[0] = U[13] * P[1] + ws * (U[8] + P[2] * U[7] - P[3] * U[6] - A[11] * P[1]);
  J[1] = U[13] * P[2] + ws * (U[9] + P[3] * U[5] - P[1] * U[7] - A[11] * P[2]);
  J[2] = U[13] * P[3] + ws * (U[10] + P[1] * U[6] - P[2] * U[5] - A[11] * P[3]);
}

// ------------------------------------------------------- SSP2 implicit stage
// IMEX SSP2(222) constants held by the reference's integrator (IMEX2Args.h:24-26).
constexpr double kSsp2Gam = 0.2928932188134525;
constexpr double kSsp2Om2Gam = 0.4142135623730950;

// Stiff part of the stage equations of SSP2.cc:194-264 in y = (Ex, Ey, Ez, Z):
// E + dtg J(x) = c_E together with the recovery equation in Z. Every other
// component of the stage value equals its explicit part c; psi and phi are
// linear and solved in closed form by imex_stage.
__device__ __forceinline__ bool imex_residual(const double c[14], double dtg, double gam, double sigma,
                                              const double y[4], double F[4]) {
  const double Ex = y[0], Ey = y[1], Ez = y[2], Z = y[3];
  const double Bx = c[5], By = c[6], Bz = c[7];
  const double Sbx = c[1] - (Ey * Bz - Ez * By);
  const double Sby = c[2] - (Ez * Bx - Ex * Bz);
  const double Sbz = c[3] - (Ex * By - Ey * Bx);
  const double tauBar = c[4] - 0.5 * (Bx * Bx + By * By + Bz * Bz + Ex * Ex + Ey * Ey + Ez * Ez);
  if (!(Z > 0.0)) return false;
  const double iZ = 1.0 / Z;
  const double vx = Sbx * iZ, vy = Sby * iZ, vz = Sbz * iZ;
  const double vsq = vx * vx + vy * vy + vz * vz;
  if (!(vsq < 1.0)) return false;
  const double W = 1.0 / sqrt(1.0 - vsq);
  F[3] = (1.0 - (gam - 1.0) / (W * W * gam)) * Z + ((gam - 1.0) / (W * gam) - 1.0) * c[0] - tauBar;
  const double vE = vx * Ex + vy * Ey + vz * Ez;
  const double q = c[13], ws = W * sigma;
  F[0] = Ex - c[8] + dtg * (q * vx + ws * (Ex + (vy * Bz - vz * By) - vE * vx));
  F[1] = Ey - c[9] + dtg * (q * vy + ws * (Ey + (vz * Bx - vx * Bz) - vE * vy));
  F[2] = Ez - c[10] + dtg * (q * vz + ws * (Ez + (vx * By - vy * Bx) - vE * vz));
  return true;
}

// imex_residual and its Jacobian J[i][j] = dF_i/dy_j at y, in closed form:
// with Sb = S - E x B, v = Sb / Z, W = (1 - v^2)^-1/2, tauBar = tau - (B^2 +
// E^2)/2 and G = E + v x B - (v.E) v, the residual is F_E = E - c_E +
// dtg (q v + sigma W G) and F_Z = (1 - g/(gam W^2)) Z + (g/(gam W) - 1) D -
// tauBar (g = gam - 1); dv/dE_k = (dSb/dE_k)/Z, dv/dZ = -v/Z,
// dW = W^3 (v.dv), dtauBar/dE_k = -E_k. F is formed exactly as
// imex_residual forms it.
__device__ __forceinline__ bool imex_residual_jac(const double c[14], double dtg, double gam, double sigma,
                                                  const double y[4], double F[4], double J[4][4]) {
  const double Ex = y[0], Ey = y[1], Ez = y[2], Z = y[3];
  const double Bx = c[5], By = c[6], Bz = c[7];
  const double Sbx = c[1] - (Ey * Bz - Ez * By);
  const double Sby = c[2] - (Ez * Bx - Ex * Bz);
  const double Sbz = c[3] - (Ex * By - Ey * Bx);
  const double tauBar = c[4] - 0.5 * (Bx * Bx + By * By + Bz * Bz + Ex * Ex + Ey * Ey + Ez * Ez);
  if (!(Z > 0.0)) return false;
  const double iZ = 1.0 / Z;
  const double vx = Sbx * iZ, vy = Sby * iZ, vz = Sbz * iZ;
  const double vsq = vx * vx + vy * vy + vz * vz;
  if (!(vsq < 1.0)) return false;
  const double W = 1.0 / sqrt(1.0 - vsq);
  F[3] = (1.0 - (gam - 1.0) / (W * W * gam)) * Z + ((gam - 1.0) / (W * gam) - 1.0) * c[0] - tauBar;
  const double vE = vx * Ex + vy * Ey + vz * Ez;
  const double q = c[13], ws = W * sigma;
  const double G[3] = {Ex + (vy * Bz - vz * By) - vE * vx, Ey + (vz * Bx - vx * Bz) - vE * vy,
                       Ez + (vx * By - vy * Bx) - vE * vz};
  F[0] = Ex - c[8] + dtg * (q * vx + ws * (Ex + (vy * Bz - vz * By) - vE * vx));
  F[1] = Ey - c[9] + dtg * (q * vy + ws * (Ey + (vz * Bx - vx * Bz) - vE * vy));
  F[2] = Ez - c[10] + dtg * (q * vz + ws * (Ez + (vx * By - vy * Bx) - vE * vz));
  const double v[3] = {vx, vy, vz}, E[3] = {Ex, Ey, Ez};
  // dSb/dE_k: columns (0, Bz, -By), (-Bz, 0, Bx), (By, -Bx, 0)
  const double dSb[3][3] = {{0.0, -Bz, By}, {Bz, 0.0, -Bx}, {-By, Bx, 0.0}};  // [component][k]
  const double g = gam - 1.0, W2 = W * W, W3 = W2 * W;
  const double cW = 2.0 * g / (W3 * gam) * Z - g / (W2 * gam) * c[0];  // dF_Z/dW
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double dv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) dv[i] = a < 3 ? dSb[i][a] * iZ : -v[i] * iZ;
    const double dW = W3 * (v[0] * dv[0] + v[1] * dv[1] + v[2] * dv[2]);
    J[3][a] = cW * dW + (a == 3 ? 1.0 - g / (W2 * gam) : E[a]);
    const double dvE = dv[0] * E[0] + dv[1] * E[1] + dv[2] * E[2] + (a < 3 ? v[a] : 0.0);
    const double dvxB[3] = {dv[1] * Bz - dv[2] * By, dv[2] * Bx - dv[0] * Bz, dv[0] * By - dv[1] * Bx};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double dG = (a == k ? 1.0 : 0.0) + dvxB[k] - dvE * v[k] - vE * dv[k];
      J[k][a] = (a == k ? 1.0 : 0.0) + dtg * (q * dv[k] + sigma * (dW * G[k] + W * dG));
    }
  }
  return true;
}

// Newton with the closed-form Jacobian and partial pivoting; steps that leave
// the physical region are halved. Returns iterations or -1. The residual and Jacobian
// the step search forms at the accepted point serve the next iteration.
__device__ __forceinline__ int imex_solve(const double c[14], double dtg, double gam, double sigma, double y[4]) {
  double F[4], J[4][4], M[4][5], dprev = HUGE_VAL;
  if (!imex_residual_jac(c, dtg, gam, sigma, y, F, J)) return -1;
  for (int it = 0; it < 60; ++it) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int j = 0; j < 4; ++j) M[i][j] = J[i][j];
      M[i][4] = -F[i];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      // partial pivoting by unrolled conditional row swaps (static indices)
#pragma unroll
      for (int i = k + 1; i < 4; ++i) {
        const bool sw = fabs(M[i][k]) > fabs(M[k][k]);
#pragma unroll
        for (int j = 0; j < 5; ++j) {
          const double t = M[k][j], u = M[i][j];
          M[k][j] = sw ? u : t;
          M[i][j] = sw ? t : u;
        }
      }
      if (M[k][k] == 0.0) return -1;
      const double inv = 1.0 / M[k][k];
#pragma unroll
      for (int i = k + 1; i < 4; ++i) {
        const double m = M[i][k] * inv;
#pragma unroll
        for (int j = k; j < 5; ++j) M[i][j] -= m * M[k][j];
      }
    }
    double d[4];
#pragma unroll
    for (int k = 3; k >= 0; --k) {
      double s = M[k][4];
#pragma unroll
      for (int j = k + 1; j < 4; ++j) s -= M[k][j] * d[j];
      d[k] = s / M[k][k];
    }
    double ymax = 1.0, dmax = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      ymax = fmax(ymax, fabs(y[j]));
      dmax = fmax(dmax, fabs(d[j]));
    }
    // Converged after this step: it is below 1e-14 |y|, or below 1e-8 |y| and a
    // thousandth of the previous step, so that convergence is quadratic and
    // the next correction would be below about 1e-14 |y| too.
    const bool last = dmax <= 1e-14 * ymax || (dmax <= 1e-8 * ymax && dmax <= 1e-3 * dprev);
    double lam = 1.0, yn[4];
    bool ok = false;
    for (int tries = 0; tries < 40; ++tries) {
#pragma unroll
      for (int j = 0; j < 4; ++j) yn[j] = y[j] + lam * d[j];
      if (last ? imex_residual(c, dtg, gam, sigma, yn, F) : imex_residual_jac(c, dtg, gam, sigma, yn, F, J)) {
        ok = true;
        break;
      }
      lam *= 0.5;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) y[j] = yn[j];
    if (last) return it + 1;
    if (!ok) return -1;  // no physical point along the step
    dprev = dmax;
  }
  return -1;
}

// One implicit stage of one cell: c is the explicit part (cons for stage one,
// cons - dt (flux1 - om2gam source1) for stage two), Zg the starting Z.
__device__ __forceinline__ int imex_stage(const double c[14], double dt, double gam, double sigma, double cp, double Zg,
                                          double x[14]) {
  const double dtg = dt * kSsp2Gam;
  double y[4] = {c[8], c[9], c[10], Zg};
  const int its = imex_solve(c, dtg, gam, sigma, y);
#pragma unroll
  for (int v = 0; v < 14; ++v) x[v] = c[v];
  x[8] = y[0];
  x[9] = y[1];
  x[10] = y[2];
  const double k = 1.0 + dtg / (cp * cp);
  x[11] = (c[11] + dtg * x[13]) / k;
  x[12] = c[12] / k;
  return its;
}

}  // namespace mgpu
