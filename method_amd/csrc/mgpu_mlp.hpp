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

// Inference form in device memory: weights transposed to [out][in] in BF16 (W1
// padded to 16 inputs, W3 to 16 outputs) as MFMA A operands, FP32 biases and
// the output de-normalisation.
struct MlpDev {
  uint16_t w1t[kMlpHid * 16];
  uint16_t w2t[kMlpHid * kMlpHid];
  uint16_t w3t[16 * kMlpHid];
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

// tanh-form GELU written through the logistic function:
// 0.5 x (1 + tanh(c (x + 0.044715 x^3))) = x sigmoid(2 c (x + 0.044715 x^3)),
// with the logistic's exp taken as a base-2 exponential (log2 e folded into
// the constants) and its reciprocal as the hardware reciprocal: one exp2, one
// rcp and five FP32 operations, for training and inference alike.
constexpr float kGeluK = 1.5957691216057308f;  // 2 sqrt(2/pi)
constexpr float kGeluA = 0.044715f;
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kGeluK2 = -kGeluK * kLog2e;           // exp(-u) = exp2(x (kGeluK2 + kGeluKA2 x^2))
constexpr float kGeluKA2 = -kGeluK * kGeluA * kLog2e;

__device__ __forceinline__ float gelu_sigmoid(float x) {
  const float t = x * fmaf(kGeluKA2, x * x, kGeluK2);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(t));
}

__device__ __forceinline__ float gelu(float x) { return x * gelu_sigmoid(x); }

__device__ __forceinline__ float gelu_grad(float x) {
  const float s = gelu_sigmoid(x);
  return s + x * s * (1.0f - s) * kGeluK * (1.0f + 3.0f * kGeluA * x * x);
}

// Decoded guess (v^2, Z = rho h W^2) for the Newton solve from the two outputs.
__device__ __forceinline__ void mlp_decode_guess(float y0, float y1, const MlpDev &m, double D, double tau, double gam,
                                                 double &v2, double &Z) {
  double lw = (double)(y0 * m.ystd[0] + m.ymean[0]);
  if (lw < 0.0) lw = 0.0;
  if (lw > 6.0) lw = 6.0;
  const double W = exp(lw);
  const double p = (tau + D) * exp((double)(y1 * m.ystd[1] + m.ymean[1]));
  v2 = 1.0 - 1.0 / (W * W);
  const double rho = D / W;
  const double h = 1.0 + p / rho * (gam / (gam - 1.0));
  Z = rho * h * W * W;
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
