// Learned cons2prim guess: training kernels (FP32 GEMMs, GELU, Adam), the
// BF16 MFMA inference and the cons2prim kernel that uses it, plus their C ABI.
// No reference counterpart: the reference starts its SRMHD root find from the
// previous primitives (srmhd.cc:438-442); this network replaces that guess and
// the cell falls back to it when Newton does not converge.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "method_gpu.h"
#include "mgpu_kernels.hpp"
#include "mgpu_mlp.hpp"

using namespace mgpu;

namespace mgpu {
int engine_fail(const char *fmt, ...);  // sets the thread's error text (mgpu_engine.hip)
}

#define MLP_HIP_OK(expr)                                                                         \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess) return mgpu::engine_fail("%s failed: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// Four FP32 values as one BF16 MFMA operand, rounded to nearest even by the
// hardware conversion, two values per instruction.
typedef float floatx2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ bf16x4 bf4(float a, float b, float c, float d) {
  const bf16x2v lo = __builtin_convertvector((floatx2{a, b}), bf16x2v);
  const bf16x2v hi = __builtin_convertvector((floatx2{c, d}), bf16x2v);
  return __builtin_bit_cast(bf16x4, (uint2{__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi)}));
}

// FP32 -> BF16 bits, round to nearest even, in integer arithmetic (host).
__host__ __device__ inline uint16_t f2bf(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

// ------------------------------------------------------------- training

// C[M x N] (+)= op(A)[M x K] op(B)[K x N], row-major, FP32, 16x16 LDS tiles.
__global__ void k_gemm(int M, int N, int K, const float *A, int lda, int ta, const float *B, int ldb, int tb,
                       float *C, int ldc, int accumulate) {
  __shared__ float sa[16][17], sb[16][17];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
  float acc = 0.0f;
  for (int k0 = 0; k0 < K; k0 += 16) {
    const int ka = k0 + tx, kb = k0 + ty;
    sa[ty][tx] = (row < M && ka < K) ? (ta ? A[(long long)ka * lda + row] : A[(long long)row * lda + ka]) : 0.0f;
    sb[ty][tx] = (col < N && kb < K) ? (tb ? B[(long long)col * ldb + kb] : B[(long long)kb * ldb + col]) : 0.0f;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) acc = fmaf(sa[ty][k], sb[k][tx], acc);
    __syncthreads();
  }
  if (row < M && col < N) {
    float *c = C + (long long)row * ldc + col;
    *c = accumulate ? *c + acc : acc;
  }
}

// H = gelu(A + b) in place on A's copy; A keeps the pre-activation.
__global__ void k_bias_gelu(int rows, int cols, float *Apre, const float *b, float *H) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)rows * cols) return;
  const float a = Apre[i] + b[i % cols];
  Apre[i] = a;
  H[i] = gelu(a);
}

__global__ void k_bias(int rows, int cols, float *Y, const float *b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)rows * cols) return;
  Y[i] += b[i % cols];
}

// dY = 2 (Y - T) / (rows * cols), and the batch loss into loss[0].
__global__ void k_mse_grad(int n, const float *Y, const float *T, float *dY, float *loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < n) {
    const float d = Y[i] - T[i];
    dY[i] = 2.0f * d / (float)n;
    l = d * d / (float)n;
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}

__global__ void k_gelu_back(int n, const float *dH, const float *Apre, float *dA) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dA[i] = dH[i] * gelu_grad(Apre[i]);
}

// db[c] = sum over rows of dA[r][c]; one block per column.
__global__ void k_colsum(int rows, int cols, const float *dA, float *db) {
  const int c = blockIdx.x;
  float s = 0.0f;
  for (int r = threadIdx.x; r < rows; r += blockDim.x) s += dA[(long long)r * cols + c];
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] = red[0];
}

__global__ void k_adam(int n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Inputs and normalised targets of n samples from (D, Sx, Sy, Sz, tau, Bx, By,
// Bz) and (W, p), both row-major double.
__global__ void k_make_data(long long n, const double *cons8, const double *wp, float *X, float *T) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *u = cons8 + 8 * i;
  mlp_features(u[0], u[1], u[2], u[3], u[4], u[5], u[6], u[7], X + 8 * i);
  mlp_targets(wp[2 * i], wp[2 * i + 1], u[0], u[4], T + 2 * i);
}

__global__ void k_normalise_targets(long long n, float *T, float m0, float s0, float m1, float s1) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T[2 * i] = (T[2 * i] - m0) / s0;
  T[2 * i + 1] = (T[2 * i + 1] - m1) / s1;
}

// ------------------------------------------------------------ inference

// The network's constants as a block stages them in LDS: the 24 weight
// fragments in lane order (a wave reads fragment f as one contiguous row,
// free of bank conflicts), the biases and the output scales.
struct alignas(16) MlpShared {
  uint2 frag[kMlpFrags][64];
  float b1[kMlpHid], b2[kMlpHid];
  float b3[kMlpOut], ymean[kMlpOut], ystd[kMlpOut];
};
static_assert(sizeof(MlpDev::frag) == sizeof(MlpShared::frag), "fragment layout");

__device__ __forceinline__ void mlp_stage_shared(const MlpDev &m, MlpShared &s) {
  const int t = threadIdx.x;
  const uint4 *src = reinterpret_cast<const uint4 *>(&m.frag[0][0][0]);
  uint4 *dst = reinterpret_cast<uint4 *>(&s.frag[0][0]);
  for (int k = t; k < kMlpFrags * 64 / 2; k += blockDim.x) dst[k] = src[k];
  if (t < kMlpHid) s.b1[t] = m.b1[t];
  else if (t < 2 * kMlpHid) s.b2[t - kMlpHid] = m.b2[t - kMlpHid];
  else if (t < 2 * kMlpHid + kMlpOut) {
    const int k = t - 2 * kMlpHid;
    s.b3[k] = m.b3[k];
    s.ymean[k] = m.ymean[k];
    s.ystd[k] = m.ystd[k];
  }
}

__device__ __forceinline__ bf16x4 frag(const MlpShared &sh, int f, int lane) {
  return __builtin_bit_cast(bf16x4, sh.frag[f][lane]);
}

// Runs the network for the 64 cells of one wave (one per lane, `feat` their
// inputs) on MFMA 16x16x16 BF16 with FP32 accumulation, and returns this
// lane's two outputs. The weights are the A operand (16 outputs x 16 inputs
// per tile) and the activations the B operand (16 inputs x 16 cells), so a
// layer's accumulator tile (lane l: rows 4(l/16)..+3 of column l%16) is
// already the next layer's B operand; each accumulator starts from the bias
// of its rows. lds: 64*8 + 64*2 floats private to the wave.
__device__ __forceinline__ void mlp_wave(const MlpShared &sh, const float feat[8], float *lds, float out[2]) {
  const int lane = threadIdx.x & 63;
  const int lr = lane & 15, lq = lane >> 4;
  float *sfeat = lds, *sout = lds + 64 * 8;
  *reinterpret_cast<floatx4 *>(sfeat + lane * 8) = floatx4{feat[0], feat[1], feat[2], feat[3]};
  *reinterpret_cast<floatx4 *>(sfeat + lane * 8 + 4) = floatx4{feat[4], feat[5], feat[6], feat[7]};
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    // B operand of layer 1: inputs 4 lq .. 4 lq + 3 of cell 16 t + lr (zero past 8)
    bf16x4 x = {0, 0, 0, 0};
    if (lq < 2) {
      const floatx4 f = *reinterpret_cast<const floatx4 *>(sfeat + (16 * t + lr) * 8 + 4 * lq);
      x = bf4(f[0], f[1], f[2], f[3]);
    }
    bf16x4 h1[4], h2[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      floatx4 acc = *reinterpret_cast<const floatx4 *>(sh.b1 + 16 * mt + 4 * lq);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(frag(sh, mt, lane), x, acc, 0, 0, 0);
      h1[mt] = bf4(gelu(acc[0]), gelu(acc[1]), gelu(acc[2]), gelu(acc[3]));
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      floatx4 acc = *reinterpret_cast<const floatx4 *>(sh.b2 + 16 * mt + 4 * lq);
#pragma unroll
      for (int kt = 0; kt < 4; ++kt)
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(frag(sh, 4 + 4 * mt + kt, lane), h1[kt], acc, 0, 0, 0);
      h2[mt] = bf4(gelu(acc[0]), gelu(acc[1]), gelu(acc[2]), gelu(acc[3]));
    }
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(frag(sh, 20 + kt, lane), h2[kt], acc, 0, 0, 0);
    if (lq == 0) {  // rows 0 and 1 are the outputs of cell 16 t + lr
      sout[(16 * t + lr) * 2] = acc[0] + sh.b3[0];
      sout[(16 * t + lr) * 2 + 1] = acc[1] + sh.b3[1];
    }
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_s_waitcnt(0);
  out[0] = sout[lane * 2];
  out[1] = sout[lane * 2 + 1];
  // the next group's feature writes must not overtake this group's reads
  __builtin_amdgcn_wave_barrier();
}

constexpr int kThr = 256;
// 256-cell groups per block of the fused cons2prim: the constants a block
// stages serve four groups of its cells.
constexpr int kLearnedGroups = 4;

// Network outputs (normalised) for n cells given as (D, Sx, Sy, Sz, tau, Bx,
// By, Bz), with the inputs the fused cons2prim forms (mlp_features_fast).
__global__ __launch_bounds__(kThr) void k_mlp_predict(long long n, const double *cons8, const MlpDev *m, float *y) {
  __shared__ __align__(16) float lds[kThr / 64][64 * 10];
  __shared__ MlpShared sh;
  mlp_stage_shared(*m, sh);
  __syncthreads();
  const long long i = (long long)blockIdx.x * kThr + threadIdx.x;
  float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (i < n) {
    double u[9];
#pragma unroll
    for (int v = 0; v < 8; ++v) u[v] = cons8[8 * i + v];
    u[8] = 0.0;
    mlp_features_fast(u, f);
  }
  float out[2];
  mlp_wave(sh, f, lds[threadIdx.x >> 6], out);
  if (i < n) {
    y[2 * i] = out[0];
    y[2 * i + 1] = out[1];
  }
}

// SRMHD cons2prim with the learned starting guess; cells whose Newton solve
// does not converge restart from the previous primitives (the reference's
// guess, srmhd.cc:438-442) and count as fallbacks. A block takes
// kLearnedGroups consecutive runs of 256 cells and stages the network's
// constants in LDS once for all of them.
__global__ __launch_bounds__(kThr) void k_c2p_learned_srmhd(Geo g, const double *U, double *V, double *A,
                                                             const MlpDev *m, double gam, C2PCounters *st,
                                                             unsigned long long *hist) {
  __shared__ __align__(16) float lds[kThr / 64][64 * 10];
  __shared__ MlpShared sh;
  __shared__ unsigned s_hist[kHistBins];
  mlp_stage_shared(*m, sh);
  if (hist)
    for (int b = threadIdx.x; b < kHistBins; b += kThr) s_hist[b] = 0u;
  __syncthreads();
  const long long vs = g.vstride;
  const float gfac = (float)(gam / (gam - 1.0));
  const float ym[2] = {sh.ymean[0], sh.ymean[1]}, ys[2] = {sh.ystd[0], sh.ystd[1]};
  int cells = 0, iters = 0, fails = 0, maxit = 0, fallback = 0;
  for (int grp = 0; grp < kLearnedGroups; ++grp) {
    const long long q = ((long long)blockIdx.x * kLearnedGroups + grp) * kThr + threadIdx.x;
    const bool live = q < vs;
    double u[9];
    float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (live) {
#pragma unroll
      for (int v = 0; v < 9; ++v) u[v] = U[v * vs + q];
      mlp_features_fast(u, f);
    }
    float y[2];
    mlp_wave(sh, f, lds[threadIdx.x >> 6], y);
    if (live) {
      double v2, Z;
      mlp_decode_guess_fast(y[0], y[1], ym, ys, u[0], u[4], gfac, v2, Z);
      const double BS = u[5] * u[1] + u[6] * u[2] + u[7] * u[3];
      const double Bsq = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
      const double Ssq = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
      int its = srmhd_newton(u[0], u[4], Bsq, Ssq, BS, gam, v2, Z);
      int total = its < 0 ? kNewtonMaxIt : its;
      if (its < 0) {
        ++fallback;
        srmhd_plain_guess(V[q], V[vs + q], V[2 * vs + q], V[3 * vs + q], V[4 * vs + q], gam, v2, Z);
        its = srmhd_newton(u[0], u[4], Bsq, Ssq, BS, gam, v2, Z);
        total += its < 0 ? kNewtonMaxIt : its;
      }
      double P[8], aux[13];
      srmhd_recover(u, gam, v2, Z, P, aux);
#pragma unroll
      for (int v = 0; v < 8; ++v) V[v * vs + q] = P[v];
#pragma unroll
      for (int v = 0; v < 13; ++v) A[v * vs + q] = aux[v];
      ++cells;
      iters += total;
      maxit = max(maxit, total);
      fails += its < 0;
      if (hist) atomicAdd(&s_hist[its < 0 || total >= kHistBins - 1 ? kHistBins - 1 : total], 1u);
    }
  }
  if (hist) {
    __syncthreads();
    for (int b = threadIdx.x; b < kHistBins; b += kThr)
      if (s_hist[b]) atomicAdd(&hist[b], (unsigned long long)s_hist[b]);
  }
  if (st) {
    for (int off = 32; off > 0; off >>= 1) {
      cells += __shfl_xor(cells, off);
      iters += __shfl_xor(iters, off);
      fails += __shfl_xor(fails, off);
      fallback += __shfl_xor(fallback, off);
      maxit = max(maxit, __shfl_xor(maxit, off));
    }
    if ((threadIdx.x & 63) == 0 && cells > 0) {
      C2PCounters *b = st + ((blockIdx.x * 4 + (threadIdx.x >> 6)) & (kStatBuckets - 1));
      atomicAdd(&b->cells, (unsigned long long)cells);
      atomicAdd(&b->iterations, (unsigned long long)iters);
      if (fails) atomicAdd(&b->failures, (unsigned long long)fails);
      if (fallback) atomicAdd(&b->retry_failures, (unsigned long long)fallback);
      atomicMax(&b->max_iterations, (unsigned long long)maxit);
    }
  }
}

inline unsigned nblk(long long n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void mgpu::launch_c2p_learned_srmhd(const Geo &g, const double *U, double *V, double *A, const MlpDev *m, double gam,
                                    C2PCounters *st, unsigned long long *hist, hipStream_t s) {
  hipLaunchKernelGGL(k_c2p_learned_srmhd, dim3(nblk(g.vstride, kThr * kLearnedGroups)), dim3(kThr), 0, s, g, U, V, A,
                     m, gam, st, hist);
}

// Packs FP32 parameters and output scales into the inference form.
void mgpu::mlp_pack(const float *params, const float *ymean, const float *ystd, MlpDev *out) {
  memset(out, 0, sizeof(MlpDev));
  // transposed weights [out][in]: W1 padded to 16 inputs, W3 to 16 outputs
  auto w1t = [&](int o, int i) { return i < kMlpIn ? params[kMlpW1 + i * kMlpHid + o] : 0.0f; };
  auto w2t = [&](int o, int i) { return params[kMlpW2 + i * kMlpHid + o]; };
  auto w3t = [&](int o, int i) { return o < kMlpOut ? params[kMlpW3 + i * kMlpOut + o] : 0.0f; };
  for (int lane = 0; lane < 64; ++lane) {
    const int lr = lane & 15, lq = lane >> 4;
    for (int j = 0; j < 4; ++j) {
      for (int mt = 0; mt < 4; ++mt) {
        out->frag[mt][lane][j] = f2bf(w1t(16 * mt + lr, 4 * lq + j));
        for (int kt = 0; kt < 4; ++kt)
          out->frag[4 + 4 * mt + kt][lane][j] = f2bf(w2t(16 * mt + lr, 16 * kt + 4 * lq + j));
      }
      for (int kt = 0; kt < 4; ++kt) out->frag[20 + kt][lane][j] = f2bf(w3t(lr, 16 * kt + 4 * lq + j));
    }
  }
  for (int k = 0; k < kMlpHid; ++k) {
    out->b1[k] = params[kMlpB1 + k];
    out->b2[k] = params[kMlpB2 + k];
  }
  for (int k = 0; k < kMlpOut; ++k) {
    out->b3[k] = params[kMlpB3 + k];
    out->ymean[k] = ymean[k];
    out->ystd[k] = ystd[k];
  }
}

// ----------------------------------------------------------------- C ABI

struct mgpu_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  float *X = nullptr, *T = nullptr;  // training set
  long long n = 0;
  int cap = 0;  // batch capacity of the work buffers
  float *A1 = nullptr, *H1 = nullptr, *A2 = nullptr, *H2 = nullptr, *Y = nullptr, *dY = nullptr, *dH = nullptr,
        *dA = nullptr, *loss = nullptr;
  float ymean[2] = {0.f, 0.f}, ystd[2] = {1.f, 1.f};
  long long adam_t = 0;
  MlpDev *dev_pack = nullptr;
};

namespace {

int mlp_ensure_batch(mgpu_mlp *m, int batch) {
  if (batch <= m->cap) return 0;
  float **bufs[] = {&m->A1, &m->H1, &m->A2, &m->H2, &m->dH, &m->dA};
  for (auto b : bufs) {
    if (*b) hipFree(*b);
    MLP_HIP_OK(hipMalloc(b, sizeof(float) * batch * kMlpHid));
  }
  if (m->Y) hipFree(m->Y);
  if (m->dY) hipFree(m->dY);
  MLP_HIP_OK(hipMalloc(&m->Y, sizeof(float) * batch * 2));
  MLP_HIP_OK(hipMalloc(&m->dY, sizeof(float) * batch * 2));
  m->cap = batch;
  return 0;
}

void gemm(hipStream_t s, int M, int N, int K, const float *A, int lda, int ta, const float *B, int ldb, int tb,
          float *C, int ldc, int acc) {
  dim3 grid((N + 15) / 16, (M + 15) / 16), block(16, 16);
  hipLaunchKernelGGL(k_gemm, grid, block, 0, s, M, N, K, A, lda, ta, B, ldb, tb, C, ldc, acc);
}

// Forward and backward of one batch (X: b x 8, T: b x 2); gradients into m->grads.
void forward_backward(mgpu_mlp *m, const float *X, const float *T, int b) {
  hipStream_t s = m->stream;
  const float *P = m->params;
  float *G = m->grads;
  const int H = kMlpHid, T256 = 256;
  gemm(s, b, H, kMlpIn, X, kMlpIn, 0, P + kMlpW1, H, 0, m->A1, H, 0);
  hipLaunchKernelGGL(k_bias_gelu, dim3(nblk((long long)b * H, T256)), dim3(T256), 0, s, b, H, m->A1, P + kMlpB1, m->H1);
  gemm(s, b, H, H, m->H1, H, 0, P + kMlpW2, H, 0, m->A2, H, 0);
  hipLaunchKernelGGL(k_bias_gelu, dim3(nblk((long long)b * H, T256)), dim3(T256), 0, s, b, H, m->A2, P + kMlpB2, m->H2);
  gemm(s, b, kMlpOut, H, m->H2, H, 0, P + kMlpW3, kMlpOut, 0, m->Y, kMlpOut, 0);
  hipLaunchKernelGGL(k_bias, dim3(nblk((long long)b * kMlpOut, T256)), dim3(T256), 0, s, b, kMlpOut, m->Y, P + kMlpB3);
  hipLaunchKernelGGL(k_mse_grad, dim3(nblk((long long)b * kMlpOut, T256)), dim3(T256), 0, s, b * kMlpOut, m->Y, T,
                     m->dY, m->loss);
  // layer 3
  gemm(s, H, kMlpOut, b, m->H2, H, 1, m->dY, kMlpOut, 0, G + kMlpW3, kMlpOut, 0);
  hipLaunchKernelGGL(k_colsum, dim3(kMlpOut), dim3(256), 0, s, b, kMlpOut, m->dY, G + kMlpB3);
  gemm(s, b, H, kMlpOut, m->dY, kMlpOut, 0, P + kMlpW3, kMlpOut, 1, m->dH, H, 0);
  hipLaunchKernelGGL(k_gelu_back, dim3(nblk((long long)b * H, T256)), dim3(T256), 0, s, b * H, m->dH, m->A2, m->dA);
  // layer 2
  gemm(s, H, H, b, m->H1, H, 1, m->dA, H, 0, G + kMlpW2, H, 0);
  hipLaunchKernelGGL(k_colsum, dim3(H), dim3(256), 0, s, b, H, m->dA, G + kMlpB2);
  gemm(s, b, H, H, m->dA, H, 0, P + kMlpW2, H, 1, m->dH, H, 0);
  hipLaunchKernelGGL(k_gelu_back, dim3(nblk((long long)b * H, T256)), dim3(T256), 0, s, b * H, m->dH, m->A1, m->dA);
  // layer 1
  gemm(s, kMlpIn, H, b, X, kMlpIn, 1, m->dA, H, 0, G + kMlpW1, H, 0);
  hipLaunchKernelGGL(k_colsum, dim3(H), dim3(256), 0, s, b, H, m->dA, G + kMlpB1);
}

}  // namespace

extern "C" {

int mgpu_mlp_create(int device, unsigned long long seed, mgpu_mlp **out) {
  if (!out) return engine_fail("mgpu_mlp_create: null argument");
  MLP_HIP_OK(hipSetDevice(device));
  mgpu_mlp *m = new mgpu_mlp();
  m->device = device;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    delete m;
    return engine_fail("mgpu_mlp_create: stream");
  }
  // Glorot-uniform weights from a seeded generator, zero biases
  std::vector<float> p(kMlpParams, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int fan_in, int fan_out) {
    const float a = std::sqrt(6.0f / (float)(fan_in + fan_out));
    std::uniform_real_distribution<float> U(-a, a);
    for (int k = 0; k < fan_in * fan_out; ++k) p[off + k] = U(rng);
  };
  fill(kMlpW1, kMlpIn, kMlpHid);
  fill(kMlpW2, kMlpHid, kMlpHid);
  fill(kMlpW3, kMlpHid, kMlpOut);
  float **bufs[] = {&m->params, &m->grads, &m->adam_m, &m->adam_v};
  for (auto b : bufs) {
    if (hipMalloc(b, sizeof(float) * kMlpParams) != hipSuccess) return engine_fail("mgpu_mlp_create: hipMalloc");
    hipMemset(*b, 0, sizeof(float) * kMlpParams);
  }
  MLP_HIP_OK(hipMalloc(&m->loss, sizeof(float)));
  MLP_HIP_OK(hipMalloc(&m->dev_pack, sizeof(MlpDev)));
  MLP_HIP_OK(hipMemcpy(m->params, p.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
  *out = m;
  return 0;
}

int mgpu_mlp_destroy(mgpu_mlp *m) {
  if (!m) return 0;
  if (m->stream) hipStreamSynchronize(m->stream);
  float *bufs[] = {m->params, m->grads, m->adam_m, m->adam_v, m->X, m->T, m->A1, m->H1, m->A2,
                   m->H2,     m->Y,     m->dY,     m->dH,     m->dA, m->loss};
  for (float *b : bufs)
    if (b) hipFree(b);
  if (m->dev_pack) hipFree(m->dev_pack);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

int mgpu_mlp_set_data(mgpu_mlp *m, const double *cons8, const double *w_p, long long n) {
  if (!m || !cons8 || !w_p || n <= 0) return engine_fail("mgpu_mlp_set_data: bad arguments");
  if (m->X) hipFree(m->X);
  if (m->T) hipFree(m->T);
  m->X = m->T = nullptr;
  double *dc = nullptr, *dw = nullptr;
  MLP_HIP_OK(hipMalloc(&m->X, sizeof(float) * 8 * n));
  MLP_HIP_OK(hipMalloc(&m->T, sizeof(float) * 2 * n));
  MLP_HIP_OK(hipMalloc(&dc, sizeof(double) * 8 * n));
  MLP_HIP_OK(hipMalloc(&dw, sizeof(double) * 2 * n));
  MLP_HIP_OK(hipMemcpy(dc, cons8, sizeof(double) * 8 * n, hipMemcpyHostToDevice));
  MLP_HIP_OK(hipMemcpy(dw, w_p, sizeof(double) * 2 * n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_make_data, dim3(nblk(n, 256)), dim3(256), 0, m->stream, n, dc, dw, m->X, m->T);
  MLP_HIP_OK(hipStreamSynchronize(m->stream));
  // output normalisation from the raw targets
  std::vector<float> t(2 * n);
  MLP_HIP_OK(hipMemcpy(t.data(), m->T, sizeof(float) * 2 * n, hipMemcpyDeviceToHost));
  for (int k = 0; k < 2; ++k) {
    double s = 0, s2 = 0;
    for (long long i = 0; i < n; ++i) {
      s += t[2 * i + k];
      s2 += (double)t[2 * i + k] * t[2 * i + k];
    }
    const double mean = s / n, var = std::max(s2 / n - mean * mean, 1e-12);
    m->ymean[k] = (float)mean;
    m->ystd[k] = (float)std::sqrt(var);
  }
  hipLaunchKernelGGL(k_normalise_targets, dim3(nblk(n, 256)), dim3(256), 0, m->stream, n, m->T, m->ymean[0],
                     m->ystd[0], m->ymean[1], m->ystd[1]);
  MLP_HIP_OK(hipStreamSynchronize(m->stream));
  hipFree(dc);
  hipFree(dw);
  m->n = n;
  return 0;
}

int mgpu_mlp_train(mgpu_mlp *m, int steps, int batch, float lr, float *loss_out) {
  if (!m || !m->X || steps <= 0 || batch <= 0 || batch > m->n) return engine_fail("mgpu_mlp_train: bad arguments");
  if (mlp_ensure_batch(m, batch)) return -1;
  float last = 0.f;
  const long long nb = m->n / batch;
  for (int s = 0; s < steps; ++s) {
    const long long off = (m->adam_t % nb) * batch;
    MLP_HIP_OK(hipMemsetAsync(m->loss, 0, sizeof(float), m->stream));
    forward_backward(m, m->X + 8 * off, m->T + 2 * off, batch);
    m->adam_t++;
    const float c1 = 1.0f - std::pow(0.9f, (float)m->adam_t), c2 = 1.0f - std::pow(0.999f, (float)m->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(nblk(kMlpParams, 256)), dim3(256), 0, m->stream, kMlpParams, m->params, m->grads,
                       m->adam_m, m->adam_v, lr, 0.9f, 0.999f, 1e-8f, c1, c2);
    if (s == steps - 1) MLP_HIP_OK(hipMemcpyAsync(&last, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  }
  MLP_HIP_OK(hipGetLastError());
  MLP_HIP_OK(hipStreamSynchronize(m->stream));
  if (loss_out) *loss_out = last;
  return 0;
}

int mgpu_mlp_gradients(mgpu_mlp *m, const float *x, const float *t, int n, float *grads, float *loss) {
  if (!m || !x || !t || !grads || n <= 0) return engine_fail("mgpu_mlp_gradients: bad arguments");
  if (mlp_ensure_batch(m, n)) return -1;
  float *dx = nullptr, *dt = nullptr;
  MLP_HIP_OK(hipMalloc(&dx, sizeof(float) * 8 * n));
  MLP_HIP_OK(hipMalloc(&dt, sizeof(float) * 2 * n));
  MLP_HIP_OK(hipMemcpy(dx, x, sizeof(float) * 8 * n, hipMemcpyHostToDevice));
  MLP_HIP_OK(hipMemcpy(dt, t, sizeof(float) * 2 * n, hipMemcpyHostToDevice));
  MLP_HIP_OK(hipMemsetAsync(m->loss, 0, sizeof(float), m->stream));
  forward_backward(m, dx, dt, n);
  MLP_HIP_OK(hipGetLastError());
  MLP_HIP_OK(hipMemcpyAsync(grads, m->grads, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m->stream));
  if (loss) MLP_HIP_OK(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP_OK(hipStreamSynchronize(m->stream));
  hipFree(dx);
  hipFree(dt);
  return 0;
}

int mgpu_mlp_get_params(mgpu_mlp *m, float *params, float *ymean_ystd) {
  if (!m || !params) return engine_fail("mgpu_mlp_get_params: bad arguments");
  MLP_HIP_OK(hipMemcpy(params, m->params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  if (ymean_ystd) {
    ymean_ystd[0] = m->ymean[0];
    ymean_ystd[1] = m->ymean[1];
    ymean_ystd[2] = m->ystd[0];
    ymean_ystd[3] = m->ystd[1];
  }
  return 0;
}

int mgpu_mlp_set_params(mgpu_mlp *m, const float *params, const float *ymean_ystd) {
  if (!m || !params) return engine_fail("mgpu_mlp_set_params: bad arguments");
  MLP_HIP_OK(hipMemcpy(m->params, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
  if (ymean_ystd) {
    m->ymean[0] = ymean_ystd[0];
    m->ymean[1] = ymean_ystd[1];
    m->ystd[0] = ymean_ystd[2];
    m->ystd[1] = ymean_ystd[3];
  }
  return 0;
}

int mgpu_mlp_predict(mgpu_mlp *m, const double *cons8, long long n, float *y) {
  if (!m || !cons8 || !y || n <= 0) return engine_fail("mgpu_mlp_predict: bad arguments");
  std::vector<float> p(kMlpParams);
  MLP_HIP_OK(hipMemcpy(p.data(), m->params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  MlpDev pack;
  mlp_pack(p.data(), m->ymean, m->ystd, &pack);
  MLP_HIP_OK(hipMemcpy(m->dev_pack, &pack, sizeof pack, hipMemcpyHostToDevice));
  double *dc = nullptr;
  float *dy = nullptr;
  MLP_HIP_OK(hipMalloc(&dc, sizeof(double) * 8 * n));
  MLP_HIP_OK(hipMalloc(&dy, sizeof(float) * 2 * n));
  MLP_HIP_OK(hipMemcpy(dc, cons8, sizeof(double) * 8 * n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mlp_predict, dim3(nblk(n, kThr)), dim3(kThr), 0, m->stream, n, dc, m->dev_pack, dy);
  MLP_HIP_OK(hipGetLastError());
  MLP_HIP_OK(hipMemcpyAsync(y, dy, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, m->stream));
  MLP_HIP_OK(hipStreamSynchronize(m->stream));
  hipFree(dc);
  hipFree(dy);
  return 0;
}

}  // extern "C"
