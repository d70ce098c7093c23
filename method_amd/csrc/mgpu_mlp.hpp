// Learned starting guess for the SRMHD cons2prim Newton solve: an 8-64-64-2
// GELU multilayer perceptron. Features and output decoding are shared by the
// training kernels and the fused inference inside the cons2prim kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mgpu {

constexpr int kMlpIn = 8, kMlpHid = 64, kMlpOut = 2;
// Flat FP32 parameter vector: W1 [in][hid], b1, W2 [hid][hid], b2, W3 [hid][out], b3.
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpIn * kMlpHid;
constexpr int kMlpW2 = kMlpB1 + kMlpHid;
constexpr int kMlpB2 = kMlpW2 + kMlpHid * kMlpHid;
constexpr int kMlpW3 = kMlpB2 + kMlpHid;
constexpr int kMlpB3 = kMlpW3 + kMlpHid * kMlpOut;
constexpr int kMlpParams = kMlpB3 + kMlpOut;  // 4866
constexpr int kMlpWeights = kMlpIn * kMlpHid + kMlpHid * kMlpHid + kMlpHid * kMlpOut;  // 4736

// Inference form in device memory. The BF16 weights are stored as the 24 MFMA
// 16x16x16 A-operand fragments the inference reads, each in lane order (lane
// l holds 4 consecutive inputs 4 (l / 16) .. + 3 of output row l % 16 of its
// 16 x 16 tile of the transposed weights; W1 padded to 16 inputs, W3 to 16
// outputs), so a wave reads a fragment as one contiguous 512-byte row:
//   fragment mt       W1 tile (outputs 16 mt ..),
//   fragment 4+4mt+kt W2 tile (outputs 16 mt .., inputs 16 kt ..),
//   fragment 20+kt    W3 tile (inputs 16 kt ..).
// Then the FP32 biases and the output de-normalisation.
constexpr int kMlpFrags = 24;
struct MlpDev {
  uint16_t frag[kMlpFrags][64][4];
  float b1[kMlpHid], b2[kMlpHid], b3[kMlpOut];
  float ymean[kMlpOut], ystd[kMlpOut];
};

// Inputs from the conserved variables (D, Sx, Sy, Sz, tau, Bx, By, Bz) with
// E = tau + D: log D / 5, S / E, log(E / D) / 4, B / sqrt(E). Every input is
// O(1) over the training ranges, and all but the first are invariant under the
// scaling rho, p, B^2 -> lambda (rho, p, B^2).
__device__ __host__ inline void mlp_features(double D, double Sx, double Sy, double Sz, double tau, double Bx,
                                             double By, double Bz, float f[8]) {
  const double E = tau + D;
  const double iE = 1.0 / E, isE = 1.0 / sqrt(E);
  f[0] = (float)(log(D) * 0.2);
  f[1] = (float)(Sx * iE);
  f[2] = (float)(Sy * iE);
  f[3] = (float)(Sz * iE);
  f[4] = (float)(log(E / D) * 0.25);
  f[5] = (float)(Bx * isE);
  f[6] = (float)(By * isE);
  f[7] = (float)(Bz * isE);
}

// Targets: log W and log(p / E).
__device__ __host__ inline void mlp_targets(double W, double p, double D, double tau, float t[2]) {
  t[0] = (float)log(W);
  t[1] = (float)log(p / (tau + D));
}

// GELU, x Phi(x), with the normal distribution function Phi taken as a
// clamped odd polynomial: Phi(x) ~ 1/2 + t P(t^2), t = clamp(x, -c, c),
// P of degree 4, its coefficients a minimax fit of |x (Phi~(x) - Phi(x))|
// over the real line under the constraint Phi~(c) = 1 (largest error
// 1.22e-3, below the BF16 rounding of the activation it feeds; x beyond c
// gives x or 0). Eight FP32 operations and no transcendental; the training
// kernels use the same function and its exact derivative.
constexpr float kGeluC = 3.38253573f;
constexpr float kGeluP0 = 3.93638938e-01f, kGeluP1 = -5.91898904e-02f, kGeluP2 = 6.53467361e-03f,
                kGeluP3 = -3.95613815e-04f, kGeluP4 = 9.83280758e-06f;

__device__ __forceinline__ float gelu_poly(float t2) {
  float p = fmaf(t2, kGeluP4, kGeluP3);
  p = fmaf(t2, p, kGeluP2);
  p = fmaf(t2, p, kGeluP1);
  return fmaf(t2, p, kGeluP0);
}

__device__ __forceinline__ float gelu(float x) {
  const float t = __builtin_amdgcn_fmed3f(x, -kGeluC, kGeluC);
  return x * fmaf(t, gelu_poly(t * t), 0.5f);
}

// d/dx of gelu: Phi~(x) + x Phi~'(x), Phi~' = P(t^2) + 2 t^2 P'(t^2) inside the
// clamp and 0 outside.
__device__ __forceinline__ float gelu_grad(float x) {
  const float t = __builtin_amdgcn_fmed3f(x, -kGeluC, kGeluC);
  const float t2 = t * t;
  const float P = gelu_poly(t2);
  float dP = fmaf(t2, 4.0f * kGeluP4, 3.0f * kGeluP3);
  dP = fmaf(t2, dP, 2.0f * kGeluP2);
  dP = fmaf(t2, dP, kGeluP1);
  const float dphi = fabsf(x) < kGeluC ? fmaf(2.0f * t2, dP, P) : 0.0f;
  return fmaf(x, dphi, fmaf(t, P, 0.5f));
}

// The network inputs of mlp_features in FP32 for inference, from the same
// conserved values: hardware log2, reciprocal and reciprocal square root, and
// log(E / D) as log(1 + tau / D) so that it keeps its relative accuracy when
// tau << D.
constexpr float kLn2 = 0.6931471805599453f;
__device__ __forceinline__ void mlp_features_fast(const double u[9], float f[8]) {
  const float D = (float)u[0], tau = (float)u[4];
  const float E = D + tau;
  const float iE = __builtin_amdgcn_rcpf(E), isE = __builtin_amdgcn_rsqf(E);
  f[0] = __builtin_amdgcn_logf(D) * (0.2f * kLn2);
  f[1] = (float)u[1] * iE;
  f[2] = (float)u[2] * iE;
  f[3] = (float)u[3] * iE;
  f[4] = __builtin_amdgcn_logf(1.0f + tau * __builtin_amdgcn_rcpf(D)) * (0.25f * kLn2);
  f[5] = (float)u[5] * isE;
  f[6] = (float)u[6] * isE;
  f[7] = (float)u[7] * isE;
}

// Decoded guess (v^2, Z = rho h W^2) for the Newton solve from the two
// normalised outputs, in FP32 (the guess only seeds the FP64 Newton solve):
// W = exp(log W) clamped to [1, e^6], p = E exp(log(p / E)),
// v^2 = 1 - 1 / W^2 and Z = rho h W^2 = D W h with h = 1 + p W / D gam/(gam-1).
constexpr float kLog2e = 1.4426950408889634f;
__device__ __forceinline__ void mlp_decode_guess_fast(float y0, float y1, const float ym[2], const float ys[2],
                                                      double D, double tau, float gfac, double &v2, double &Z) {
  const float lw = fminf(fmaxf(fmaf(y0, ys[0], ym[0]), 0.0f), 6.0f);
  const float W = __builtin_amdgcn_exp2f(lw * kLog2e);
  const float iW = __builtin_amdgcn_rcpf(W);
  const float Df = (float)D, Ef = (float)(D + tau);
  const float p = Ef * __builtin_amdgcn_exp2f(fmaf(y1, ys[1], ym[1]) * kLog2e);
  const float h = fmaf(p * W * __builtin_amdgcn_rcpf(Df), gfac, 1.0f);
  v2 = (double)fmaf(-iW, iW, 1.0f);
  Z = (double)(Df * W * h);
}

struct Geo;
struct C2PCounters;
// Packs FP32 parameters (the flat layout above) and output scales into MlpDev.
void mlp_pack(const float *params, const float *ymean, const float *ystd, MlpDev *out);
// SRMHD cons2prim of every interior cell with the learned starting guess.
// hist: null, or 64 Newton-step counters (both solves of a fallback summed).
void launch_c2p_learned_srmhd(const Geo &g, const double *U, double *V, double *A, const MlpDev *m, double gam,
                              C2PCounters *st, unsigned long long *hist, hipStream_t s);

}  // namespace mgpu
