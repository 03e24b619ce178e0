// akx_surrogate.hip -- training set, training and evaluation of the
// constitutive surrogate (akx_mlp.hpp) on gfx950 matrix cores.
//
// The plain law it stands in for is the neo-Hookean energy/stress
// (material_neohookean_inline_impl.hh:46-65, 145-158) under the Marigo damage
// update (material_marigo_inline_impl.hh:67-76).  The damage update itself
// stays closed-form: it needs the point's own threshold Yd, which no network
// input carries, so the network returns the undamaged stress and the driving
// energy Y = W_NH, and the caller applies d_new = min(1, (Y - Yd) / Sd) when
// Y - Yd - Sd d > 0 and S = (1 - d_new) S_NH exactly as the plain law does.
//
// Training: minibatch Adam on the mean squared error of the scaled outputs,
// FP32 master weights, BF16 MFMA for every product of the forward and the
// backward pass (FP32 accumulation), deterministic reductions (fixed-order
// split-K partial sums, no floating-point atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "akx_internal.hpp"
#include "akx_mlp.hpp"

namespace akx {
using namespace mlp;

namespace {

// ---------------------------------------------------------------------------
// counter-based random numbers (SplitMix64 finaliser) for the path generator
// and the minibatch draws
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
struct Rng {
  uint64_t s;
  __device__ double uniform() {  // (0, 1)
    s = mix64(s);
    return (double((s >> 11)) + 0.5) * (1.0 / 9007199254740992.0);
  }
  __device__ double normal() {
    const double u1 = uniform(), u2 = uniform();
    return sqrt(-2. * log(u1)) * cos(6.283185307179586 * u2);
  }
};

// Neo-Hookean PK2 stress (Voigt) and energy from the Green-Lagrange strain:
// C = I + 2E, J = sqrt(det C), S = mu (I - C^-1) + lambda ln J C^-1,
// W = lambda/2 ln^2 J + mu (-ln J + (tr C - 3)/2).
__device__ inline void nh_green(const double* E, double lambda, double mu, double* S, double* W) {
  const double c00 = 1 + 2 * E[0], c11 = 1 + 2 * E[1], c22 = 1 + 2 * E[2];
  const double c12 = 2 * E[3], c02 = 2 * E[4], c01 = 2 * E[5];
  const double a00 = c11 * c22 - c12 * c12, a11 = c00 * c22 - c02 * c02,
               a22 = c00 * c11 - c01 * c01;
  const double a12 = c02 * c01 - c00 * c12, a02 = c01 * c12 - c11 * c02,
               a01 = c12 * c02 - c01 * c22;
  const double det = c00 * a00 + c01 * a01 + c02 * a02;
  const double lnJ = 0.5 * log(det);
  const double id = 1. / det;
  const double ci[6] = {a00 * id, a11 * id, a22 * id, a12 * id, a02 * id, a01 * id};
  const double f = lambda * lnJ - mu;
  S[0] = mu + f * ci[0];
  S[1] = mu + f * ci[1];
  S[2] = mu + f * ci[2];
  S[3] = f * ci[3];
  S[4] = f * ci[4];
  S[5] = f * ci[5];
  *W = 0.5 * lambda * lnJ * lnJ + mu * (-lnJ + 0.5 * (c00 + c11 + c22 - 3.));
}

__device__ inline void linear_part(const double* E, double lambda, double mu, double* S,
                                   double* W) {
  const double tr = E[0] + E[1] + E[2];
  for (int k = 0; k < 3; ++k) S[k] = lambda * tr + 2 * mu * E[k];
  for (int k = 3; k < 6; ++k) S[k] = 2 * mu * E[k];
  const double ee = E[0] * E[0] + E[1] * E[1] + E[2] * E[2] +
                    2 * (E[3] * E[3] + E[4] * E[4] + E[5] * E[5]);
  *W = 0.5 * lambda * tr * tr + mu * ee;
}

// ---------------------------------------------------------------------------
// training set: one thread per strain path.  Monotone paths drift along a
// fixed random direction, cyclic paths reverse it every `period` steps; both
// add Gaussian increments and reflect at |E_k| = eps.  The damage follows the
// Marigo update with a per-path Weibull threshold.
__global__ void __launch_bounds__(256) k_sur_gen(int64_t paths, int steps, double eps,
                                                 double inc_std, double lambda, double mu,
                                                 double Sd, double yd_scale, double yd_shape,
                                                 double sS, double sW, uint64_t seed, float* X,
                                                 float* T) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= paths) return;
  Rng r{mix64(seed ^ (uint64_t(p) * 0xD1B54A32D192ED03ull))};
  const bool cyclic = r.uniform() < 0.5;
  double dir[6], nrm = 0;
  for (int k = 0; k < 6; ++k) {
    dir[k] = r.normal();
    nrm += dir[k] * dir[k];
  }
  nrm = sqrt(nrm);
  for (int k = 0; k < 6; ++k) dir[k] /= nrm;
  const double drift = inc_std * (0.5 + 2.5 * r.uniform());
  const int period = 10 + int(40 * r.uniform());
  const double Yd = yd_scale * pow(-log(r.uniform()), 1. / yd_shape);
  double E[6] = {0, 0, 0, 0, 0, 0};
  double d = 0;
  for (int s = 0; s < steps; ++s) {
    const double sg = (cyclic && ((s / period) & 1)) ? -1. : 1.;
    for (int k = 0; k < 6; ++k) {
      double v = E[k] + sg * drift * dir[k] + inc_std * r.normal();
      if (v > eps) v = 2 * eps - v;
      if (v < -eps) v = -2 * eps - v;
      E[k] = fmin(eps, fmax(-eps, v));
    }
    double S[6], W, Sl[6], Wl;
    nh_green(E, lambda, mu, S, &W);
    linear_part(E, lambda, mu, Sl, &Wl);
    const int64_t q = p * steps + s;
    for (int k = 0; k < 6; ++k) X[q * 7 + k] = float(E[k] / eps);
    X[q * 7 + 6] = float(d);
    for (int k = 0; k < 6; ++k) T[q * 7 + k] = float((S[k] - Sl[k]) / sS);
    T[q * 7 + 6] = float((W - Wl) / sW);
    // material_marigo_inline_impl.hh:67-76 with Y = W_NH
    const double Fd = W - Yd - Sd * d;
    if (Fd > 0) d = (W - Yd) / Sd;
    d = fmin(d, 1.);
  }
}

// ---------------------------------------------------------------------------
// minibatch assembly: row b of the batch is sample idx(b) (drawn with
// replacement) or, for an explicit batch, sample b
__global__ void __launch_bounds__(256) k_sur_batch(const float* __restrict__ X,
                                                   const float* __restrict__ T, int64_t n,
                                                   int B, uint64_t seed, int64_t iter,
                                                   int draw, __bf16* X0, __bf16* X0T,
                                                   float* Tb) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int64_t idx = draw ? int64_t(mix64(seed ^ mix64(uint64_t(iter) * B + b)) % uint64_t(n)) : b;
  float x[7];
  for (int k = 0; k < 7; ++k) x[k] = X[idx * 7 + k];
  for (int k = 0; k < K0; ++k) {
    float v = 0.f;
    if (k < NIN) v = x[k];
    else if (k < 2 * NIN) v = x[k - NIN] - from_bf16(to_bf16(x[k - NIN]));
    const __bf16 bv = to_bf16(v);
    X0[int64_t(b) * K0 + k] = bv;
    X0T[int64_t(k) * B + b] = bv;
  }
  for (int k = 0; k < 7; ++k) Tb[int64_t(b) * 8 + k] = T[idx * 7 + k];
  Tb[int64_t(b) * 8 + 7] = 0.f;
}

// forward layer: Z^T = W H^T + b over a 32 x 32 tile per wave.
//   Hin  [B x KIN] bf16 (row = point), W [NP x KIN] bf16, rows >= nout zero
//   ACT:  Zt [nout x B] fp32, Hout [B x nout], HoutT [nout x B] bf16 (activation)
//   !ACT: Y [B x 8] fp32
template <int KIN, int NP, bool ACT>
__global__ void __launch_bounds__(256) k_sur_fwd(const __bf16* __restrict__ Hin,
                                                 const __bf16* __restrict__ W,
                                                 const float* __restrict__ bias, int nout, int B,
                                                 float* Zt, __bf16* Hout, __bf16* HoutT,
                                                 float* Y) {
  const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  constexpr int MB = NP / 32;
  if (wave >= int64_t(MB) * (B / 32)) return;  // uniform over the wave
  const int mb = int(wave % MB), jb = int(wave / MB);
  const int lane = threadIdx.x & 63, i = lane & 31, h = lane >> 5;
  f32x16 acc = f32x16{};
#pragma unroll
  for (int t = 0; t < KIN / 16; ++t) {
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(W + (32 * mb + i) * KIN + 16 * t + 8 * h);
    const bf16x8 b =
        *reinterpret_cast<const bf16x8*>(Hin + int64_t(32 * jb + i) * KIN + 16 * t + 8 * h);
    acc = mfma(a, b, acc);
  }
  const int64_t c = 32 * jb + i;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int r = 32 * mb + crow(v, h);
    if (r >= nout) continue;
    const float z = acc[v] + bias[r];
    if constexpr (ACT) {
      Zt[int64_t(r) * B + c] = z;
      const __bf16 hv = to_bf16(act(z));
      Hout[c * nout + r] = hv;
      HoutT[int64_t(r) * B + c] = hv;
    } else {
      Y[c * 8 + r] = z;
    }
  }
}

// output error: dY = 2 (Y - T) / (7 B) (bf16, both layouts), loss partial sums
__global__ void __launch_bounds__(256) k_sur_loss(const float* __restrict__ Y,
                                                  const float* __restrict__ T, int B,
                                                  __bf16* dY, __bf16* dYT, float* partial) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (b < B) {
    const float sc = 2.f / (7.f * float(B));
    for (int k = 0; k < 7; ++k) {
      const float diff = Y[int64_t(b) * 8 + k] - T[int64_t(b) * 8 + k];
      l += diff * diff;
      const __bf16 g = to_bf16(sc * diff);
      dY[int64_t(b) * 16 + k] = g;
      dYT[int64_t(k) * B + b] = g;
    }
  }
  __shared__ float sh[4];
  for (int o = 32; o > 0; o >>= 1) l += __shfl_down(l, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// backward through a layer's weights and the previous activation:
//   dH^T = W^T dZ^T (tile rows: input features 0..127, columns: points),
//   dZprev = dH * a'(Zprev), stored bf16 in both layouts.
//   WT [128 x KC] bf16, dZ [B x KC] bf16, Zt_prev [128 x B] fp32
template <int KC>
__global__ void __launch_bounds__(256) k_sur_bwd_data(const __bf16* __restrict__ WT,
                                                      const __bf16* __restrict__ dZ,
                                                      const float* __restrict__ Zt_prev, int B,
                                                      __bf16* dZp, __bf16* dZpT) {
  const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  if (wave >= int64_t(4) * (B / 32)) return;
  const int mb = int(wave % 4), jb = int(wave / 4);
  const int lane = threadIdx.x & 63, i = lane & 31, h = lane >> 5;
  f32x16 acc = f32x16{};
#pragma unroll
  for (int t = 0; t < KC / 16; ++t) {
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(WT + (32 * mb + i) * KC + 16 * t + 8 * h);
    const bf16x8 b =
        *reinterpret_cast<const bf16x8*>(dZ + int64_t(32 * jb + i) * KC + 16 * t + 8 * h);
    acc = mfma(a, b, acc);
  }
  const int64_t c = 32 * jb + i;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int r = 32 * mb + crow(v, h);
    const float z = Zt_prev[int64_t(r) * B + c];
    const __bf16 g = to_bf16(acc[v] * act_grad(z));
    dZp[c * HID + r] = g;
    dZpT[int64_t(r) * B + c] = g;
  }
}

// weight and bias gradients, split over batch chunks:
//   partial[c][n][k] = sum_{b in chunk c} dZ^T[n][b] H^T[k][b]   (k < KC)
//   partial[c][n][KC] = sum_{b in chunk c} dZ^T[n][b]            (bias)
template <int NP, int KC>
__global__ void __launch_bounds__(256) k_sur_bwd_weight(const __bf16* __restrict__ dZT, int nout,
                                                        const __bf16* __restrict__ HT, int B,
                                                        int chunk, float* partial) {
  constexpr int MB = NP / 32, KB = (KC + 1 + 31) / 32;
  const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int nch = B / chunk;
  if (wave >= int64_t(MB) * KB * nch) return;
  const int mb = int(wave % MB), kb = int((wave / MB) % KB), ch = int(wave / (MB * KB));
  const int lane = threadIdx.x & 63, i = lane & 31, h = lane >> 5;
  const int n_a = 32 * mb + i, k_b = 32 * kb + i;
  const int64_t b0 = int64_t(ch) * chunk;
  bf16x8 one;
#pragma unroll
  for (int s = 0; s < 8; ++s) one[s] = to_bf16(1.f);
  const bf16x8 zero = bf16x8{};
  f32x16 acc = f32x16{};
  for (int t = 0; t < chunk / 16; ++t) {
    const int64_t bb = b0 + 16 * t + 8 * h;
    const bf16x8 a = n_a < nout ? *reinterpret_cast<const bf16x8*>(dZT + int64_t(n_a) * B + bb) : zero;
    const bf16x8 b = k_b < KC ? *reinterpret_cast<const bf16x8*>(HT + int64_t(k_b) * B + bb)
                              : (k_b == KC ? one : zero);
    acc = mfma(a, b, acc);
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int n = 32 * mb + crow(v, h);
    if (n < nout && k_b <= KC) partial[(int64_t(ch) * nout + n) * (KC + 1) + k_b] = acc[v];
  }
}

struct LayerDesc {
  int nout, nin, kc;     // kc: partial columns before the bias column (16 for layer 0)
  int off_w, off_b;
  const float* partial;  // [nch][nout][kc + 1]
  __bf16* Wbf;           // forward weights [np][kbf]
  int kbf;
  __bf16* WT;            // backward weights [nin][ktp] (nullptr for layer 0)
  int ktp;
};

// gradient of every parameter of one layer from the chunk partials (fixed
// order); then either an Adam step on the FP32 master copy (refreshing the
// BF16 copies) or a plain copy of the gradient to `grad_out`
__global__ void __launch_bounds__(256) k_sur_update(LayerDesc L, int nch, float* params,
                                                    float* m1, float* m2, float lr, float beta1,
                                                    float beta2, float eps, float bc1, float bc2,
                                                    float* grad_out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int nw = L.nout * L.nin;
  if (idx >= nw + L.nout) return;
  const bool is_w = idx < nw;
  const int n = is_w ? idx / L.nin : idx - nw;
  const int k = is_w ? idx % L.nin : L.kc;
  const int cols = L.kc + 1;
  float g = 0.f;
  for (int c = 0; c < nch; ++c) {
    const float* P = L.partial + (int64_t(c) * L.nout + n) * cols;
    g += P[k];
    if (is_w && L.kc == K0) g += P[k + NIN];  // layer 0: x_hi and x_lo share the weight
  }
  const int p = is_w ? L.off_w + idx : L.off_b + n;
  if (grad_out) {
    grad_out[p] = g;
    return;
  }
  const float a = beta1 * m1[p] + (1.f - beta1) * g;
  const float b = beta2 * m2[p] + (1.f - beta2) * g * g;
  m1[p] = a;
  m2[p] = b;
  const float w = params[p] - lr * (a / bc1) / (sqrtf(b / bc2) + eps);
  params[p] = w;
  if (is_w) {
    const __bf16 wb = to_bf16(w);
    L.Wbf[n * L.kbf + k] = wb;
    if (L.kc == K0) L.Wbf[n * L.kbf + k + NIN] = wb;
    if (L.WT) L.WT[k * L.ktp + n] = wb;
  }
}

// BF16 copies of the master weights (forward layout, transposed layout) and
// the permuted inference image, FP16 or (bf16 != 0) BF16
__global__ void __launch_bounds__(256) k_sur_refresh(const float* __restrict__ P, __bf16* W0,
                                                     __bf16* W1, __bf16* W2, __bf16* W3,
                                                     __bf16* W4, __bf16* T1, __bf16* T2,
                                                     __bf16* T3, __bf16* T4, f16x8* img,
                                                     float* imgb, int bf16) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  // forward / transposed copies
  if (idx < HID * K0) {
    const int n = idx / K0, k = idx % K0;
    const float w = k < NIN ? P[OFF_W0 + n * NIN + k] : k < 2 * NIN ? P[OFF_W0 + n * NIN + k - NIN] : 0.f;
    W0[idx] = to_bf16(w);
  }
  if (idx < HID * HID) {
    const int n = idx / HID, k = idx % HID;
    const __bf16 a = to_bf16(P[OFF_W1 + idx]), b = to_bf16(P[OFF_W2 + idx]),
                 c = to_bf16(P[OFF_W3 + idx]);
    W1[idx] = a; W2[idx] = b; W3[idx] = c;
    T1[k * HID + n] = a; T2[k * HID + n] = b; T3[k * HID + n] = c;
  }
  if (idx < 32 * HID) {
    const int n = idx / HID, k = idx % HID;
    const __bf16 w = to_bf16(n < NOUT ? P[OFF_W4 + n * HID + k] : 0.f);
    W4[idx] = w;
    if (n < 16) T4[k * 16 + n] = w;
  }
  // permuted image: one thread per (vector, slot)
  if (idx < PW_SIZE * 8) {
    const int vec = idx / 8, s = idx % 8, lane = vec % 64, i = lane & 31, h = lane >> 5;
    float w = 0.f;
    if (vec < PW_L1) {
      const int mb = vec / 64, k = 8 * h + s, row = 32 * mb + i;
      w = k < NIN ? P[OFF_W0 + row * NIN + k] : k < 2 * NIN ? P[OFF_W0 + row * NIN + k - NIN] : 0.f;
    } else if (vec < PW_L4) {
      const int l = (vec - PW_L1) / (4 * 8 * 64), rem = (vec - PW_L1) % (4 * 8 * 64);
      const int mb = rem / (8 * 64), t = (rem / 64) % 8, row = 32 * mb + i;
      const int off = l == 0 ? OFF_W1 : l == 1 ? OFF_W2 : OFF_W3;
      w = P[off + row * HID + chain_k(t, h, s)];
    } else {
      const int t = (vec - PW_L4) / 64;
      w = i < NOUT ? P[OFF_W4 + i * HID + chain_k(t, h, s)] : 0.f;
    }
    if (bf16)
      reinterpret_cast<Bf16::x8*>(img)[vec][s] = static_cast<__bf16>(w);
    else
      img[vec][s] = static_cast<_Float16>(w);
  }
  if (idx < PB_SIZE) {
    float b = 0.f;
    if (idx < 4 * HID) {
      const int l = idx / HID, n = idx % HID;
      const int off = l == 0 ? OFF_B0 : l == 1 ? OFF_B1 : l == 2 ? OFF_B2 : OFF_B3;
      b = P[off + n];
    } else if (idx - 4 * HID < NOUT) {
      b = P[OFF_B4 + idx - 4 * HID];
    }
    imgb[idx] = b;
  }
}

// evaluation of the network on n points (scaled inputs, sx floats per point
// -> outputs, sy floats per point); persistent blocks of 8 waves, the image
// staged in LDS once per block.  With 8 floats per point on both sides (the
// explicit step's buffers) the points move as two 16-byte words and the 8th
// input float, the strain pass's envelope flag, is copied to the 8th output
// float, where the force pass reads it with the outputs.
template <class F>
__global__ void __launch_bounds__(512) k_sur_infer(const typename F::x8* __restrict__ img,
                                                   const float* __restrict__ imgb, int64_t n,
                                                   const float* __restrict__ X, int sx, float* Y,
                                                   int sy) {
  __shared__ Image<F> L;
  stage_image<F>(L, img, imgb);
  __syncthreads();
  const int lane = threadIdx.x & 63, j = lane & 31, h = lane >> 5;
  const int64_t tiles = (n + 31) / 32;
  const int waves = blockDim.x >> 6;
  const int64_t stride = int64_t(gridDim.x) * waves;
  const bool vec8 = sx == 8 && sy == 8;
  int64_t tile = int64_t(blockIdx.x) * waves + (threadIdx.x >> 6);
  // the next tile's inputs are loaded while the current one is evaluated;
  // lanes past the end read point 0 (n > 0) and use zeros
  float xn[8];
  auto load = [&](int64_t tl, float* x) {
    const int64_t p = tl * 32 + j;
    const bool ok = tl < tiles && p < n;
    const float* src = X + (ok ? p * sx : 0);
    float v[8];
    if (vec8) {
      const float4 a = reinterpret_cast<const float4*>(src)[0];
      const float4 b = reinterpret_cast<const float4*>(src)[1];
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
      v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
      for (int k = 0; k < 7; ++k) v[k] = src[k];
      v[7] = 0.f;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = ok ? v[k] : 0.f;
  };
  load(tile, xn);
  for (; tile < tiles; tile += stride) {
    float x[8], y[7];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = xn[k];
    load(tile + stride, xn);
    eval_wave<F>(L, x, y);
    const int64_t p = tile * 32 + j;
    if (p < n && h == 0) {
      if (vec8) {
        float4* dst = reinterpret_cast<float4*>(Y + p * 8);
        dst[0] = make_float4(y[0], y[1], y[2], y[3]);
        dst[1] = make_float4(y[4], y[5], y[6], x[7]);
      } else {
#pragma unroll
        for (int k = 0; k < 7; ++k) Y[p * sy + k] = y[k];
      }
    }
  }
}

template <typename T>
hipError_t dalloc_n(T** p, size_t n) {
  *p = nullptr;
  return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T));
}
template <typename T>
void dfree_p(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

inline unsigned blocks_for_waves(int64_t waves) { return unsigned((waves + 3) / 4); }

}  // namespace

// ---------------------------------------------------------------------------
Surrogate::~Surrogate() { release(); }

void Surrogate::release_batch() {
  dfree_p(X0); dfree_p(X0T); dfree_p(Tb); dfree_p(Yb); dfree_p(dY); dfree_p(dYT);
  dfree_p(loss_part);
  for (int l = 0; l < 5; ++l) { dfree_p(H[l]); dfree_p(HT[l]); dfree_p(part[l]); }
  for (int l = 0; l < 4; ++l) { dfree_p(Zt[l]); dfree_p(dZ[l]); dfree_p(dZT[l]); }
  bcap = 0;
}

void Surrogate::release() {
  release_batch();
  dfree_p(params); dfree_p(m1); dfree_p(m2); dfree_p(grad);
  for (int l = 0; l < 5; ++l) { dfree_p(Wbf[l]); dfree_p(WT[l]); }
  dfree_p(img); dfree_p(imgb);
  dfree_p(X); dfree_p(T);
  dfree_p(fallbacks);
  dfree_p(io_x); dfree_p(io_y);
  dfree_p(pt_x); dfree_p(pt_y);
  nsamples = 0;
  io_cap = 0;
  pt_cap = 0;
}

hipError_t Surrogate::ensure_params() {
  if (params) return hipSuccess;
  hipError_t e;
  if ((e = dalloc_n(&params, NPARAM)) || (e = dalloc_n(&m1, NPARAM)) || (e = dalloc_n(&m2, NPARAM)) ||
      (e = dalloc_n(&grad, NPARAM)))
    return e;
  const int kbf[5] = {K0, HID, HID, HID, HID};
  const int rows[5] = {HID, HID, HID, HID, 32};
  for (int l = 0; l < 5; ++l)
    if ((e = dalloc_n(&Wbf[l], size_t(rows[l]) * kbf[l]))) return e;
  for (int l = 1; l < 5; ++l)
    if ((e = dalloc_n(&WT[l], size_t(HID) * (l == 4 ? 16 : HID)))) return e;
  if ((e = dalloc_n(&img, size_t(PW_SIZE) * 8)) || (e = dalloc_n(&imgb, PB_SIZE)) ||
      (e = dalloc_n(&fallbacks, 1)))
    return e;
  if ((e = hipMemset(fallbacks, 0, sizeof(unsigned long long)))) return e;
  // deterministic initialisation: U(-a, a), a = sqrt(3 / fan_in), zero biases
  std::vector<float> h(NPARAM, 0.f);
  std::mt19937_64 gen(0x5eed5eedull);
  auto fill = [&](int off, int nout, int nin) {
    const float a = std::sqrt(3.f / float(nin));
    std::uniform_real_distribution<float> U(-a, a);
    for (int k = 0; k < nout * nin; ++k) h[off + k] = U(gen);
  };
  fill(OFF_W0, HID, NIN);
  fill(OFF_W1, HID, HID);
  fill(OFF_W2, HID, HID);
  fill(OFF_W3, HID, HID);
  fill(OFF_W4, NOUT, HID);
  if ((e = hipMemcpy(params, h.data(), NPARAM * 4, hipMemcpyHostToDevice))) return e;
  if ((e = hipMemset(m1, 0, NPARAM * 4)) || (e = hipMemset(m2, 0, NPARAM * 4))) return e;
  for (int l = 1; l < 5; ++l)
    if ((e = hipMemset(WT[l], 0, size_t(HID) * (l == 4 ? 16 : HID) * 2))) return e;
  step = 0;
  return hipSuccess;
}

hipError_t Surrogate::refresh(hipStream_t s) {
  const int nthreads = std::max({HID * HID, PW_SIZE * 8});
  k_sur_refresh<<<(nthreads + 255) / 256, 256, 0, s>>>(params, Wbf[0], Wbf[1], Wbf[2], Wbf[3],
                                                       Wbf[4], WT[1], WT[2], WT[3], WT[4], img,
                                                       imgb, int(infer_bf16));
  return hipGetLastError();
}

hipError_t Surrogate::ensure_batch(int B) {
  if (B <= bcap) return hipSuccess;
  release_batch();
  hipError_t e;
  const size_t b = size_t(B);
  if ((e = dalloc_n(&X0, b * K0)) || (e = dalloc_n(&X0T, b * K0)) || (e = dalloc_n(&Tb, b * 8)) ||
      (e = dalloc_n(&Yb, b * 8)) || (e = dalloc_n(&dY, b * 16)) || (e = dalloc_n(&dYT, b * NOUT)) ||
      (e = dalloc_n(&loss_part, (b + 255) / 256)))
    return e;
  for (int l = 1; l < 5; ++l)
    if ((e = dalloc_n(&H[l], b * HID)) || (e = dalloc_n(&HT[l], b * HID))) return e;
  for (int l = 0; l < 4; ++l)
    if ((e = dalloc_n(&Zt[l], b * HID)) || (e = dalloc_n(&dZ[l], b * HID)) ||
        (e = dalloc_n(&dZT[l], b * HID)))
      return e;
  // chunk partials: at most B / 256 chunks of [nout][kc + 1]
  const size_t nch = b / 256;
  const int kc[5] = {K0, HID, HID, HID, HID}, no[5] = {HID, HID, HID, HID, NOUT};
  for (int l = 0; l < 5; ++l)
    if ((e = dalloc_n(&part[l], nch * no[l] * (kc[l] + 1)))) return e;
  if ((e = hipMemset(dY, 0, b * 16 * 2))) return e;  // columns 7..15 stay zero
  bcap = B;
  return hipSuccess;
}

// one forward + backward pass over the current batch; then either an Adam
// step (grad_only = false) or the gradient into `grad` (grad_only = true)
hipError_t Surrogate::pass(hipStream_t s, int B, bool grad_only, float lr) {
  const int chunk = (B % 1024 == 0) ? 1024 : 256;
  const int nch = B / chunk;
  const int64_t jbs = B / 32;
  k_sur_fwd<K0, HID, true><<<blocks_for_waves(4 * jbs), 256, 0, s>>>(
      X0, Wbf[0], params + OFF_B0, HID, B, Zt[0], H[1], HT[1], nullptr);
  k_sur_fwd<HID, HID, true><<<blocks_for_waves(4 * jbs), 256, 0, s>>>(
      H[1], Wbf[1], params + OFF_B1, HID, B, Zt[1], H[2], HT[2], nullptr);
  k_sur_fwd<HID, HID, true><<<blocks_for_waves(4 * jbs), 256, 0, s>>>(
      H[2], Wbf[2], params + OFF_B2, HID, B, Zt[2], H[3], HT[3], nullptr);
  k_sur_fwd<HID, HID, true><<<blocks_for_waves(4 * jbs), 256, 0, s>>>(
      H[3], Wbf[3], params + OFF_B3, HID, B, Zt[3], H[4], HT[4], nullptr);
  k_sur_fwd<HID, 32, false><<<blocks_for_waves(jbs), 256, 0, s>>>(
      H[4], Wbf[4], params + OFF_B4, NOUT, B, nullptr, nullptr, nullptr, Yb);
  k_sur_loss<<<(B + 255) / 256, 256, 0, s>>>(Yb, Tb, B, dY, dYT, loss_part);
  // layer 4
  k_sur_bwd_weight<32, HID><<<blocks_for_waves(int64_t(1) * 5 * nch), 256, 0, s>>>(
      dYT, NOUT, HT[4], B, chunk, part[4]);
  k_sur_bwd_data<16><<<blocks_for_waves(4 * jbs), 256, 0, s>>>(WT[4], dY, Zt[3], B, dZ[3], dZT[3]);
  // layers 3, 2, 1
  for (int l = 3; l >= 1; --l) {
    k_sur_bwd_weight<HID, HID><<<blocks_for_waves(int64_t(4) * 5 * nch), 256, 0, s>>>(
        dZT[l], HID, HT[l], B, chunk, part[l]);
    k_sur_bwd_data<HID><<<blocks_for_waves(4 * jbs), 256, 0, s>>>(WT[l], dZ[l], Zt[l - 1], B,
                                                                  dZ[l - 1], dZT[l - 1]);
  }
  // layer 0 (inputs X0^T, 16 columns)
  k_sur_bwd_weight<HID, K0><<<blocks_for_waves(int64_t(4) * 1 * nch), 256, 0, s>>>(
      dZT[0], HID, X0T, B, chunk, part[0]);
  hipError_t e = hipGetLastError();
  if (e) return e;
  const int nin[5] = {NIN, HID, HID, HID, HID}, nout[5] = {HID, HID, HID, HID, NOUT};
  const int kc[5] = {K0, HID, HID, HID, HID};
  const int offw[5] = {OFF_W0, OFF_W1, OFF_W2, OFF_W3, OFF_W4};
  const int offb[5] = {OFF_B0, OFF_B1, OFF_B2, OFF_B3, OFF_B4};
  const float b1 = 0.9f, b2 = 0.999f;
  if (!grad_only) ++step;
  const float bc1 = 1.f - std::pow(b1, float(std::max<int64_t>(step, 1)));
  const float bc2 = 1.f - std::pow(b2, float(std::max<int64_t>(step, 1)));
  for (int l = 0; l < 5; ++l) {
    LayerDesc L;
    L.nout = nout[l]; L.nin = nin[l]; L.kc = kc[l];
    L.off_w = offw[l]; L.off_b = offb[l];
    L.partial = part[l];
    L.Wbf = Wbf[l];
    L.kbf = l == 0 ? K0 : HID;
    L.WT = l == 0 ? nullptr : WT[l];
    L.ktp = l == 4 ? 16 : HID;
    const int cnt = L.nout * L.nin + L.nout;
    k_sur_update<<<(cnt + 255) / 256, 256, 0, s>>>(L, nch, params, m1, m2, lr, b1, b2, 1e-8f, bc1,
                                                   bc2, grad_only ? grad : nullptr);
  }
  return hipGetLastError();
}

hipError_t Surrogate::generate(hipStream_t s, int64_t paths, int steps, double inc_std,
                               double yd_scale, double yd_shape, uint64_t seed) {
  dfree_p(X);
  dfree_p(T);
  nsamples = 0;
  const int64_t n = paths * steps;
  hipError_t e;
  if ((e = dalloc_n(&X, size_t(n) * 7)) || (e = dalloc_n(&T, size_t(n) * 7))) return e;
  k_sur_gen<<<unsigned((paths + 255) / 256), 256, 0, s>>>(paths, steps, eps, inc_std, lambda, mu,
                                                          Sd, yd_scale, yd_shape, sS, sW, seed, X, T);
  if ((e = hipGetLastError())) return e;
  nsamples = n;
  return hipSuccess;
}

hipError_t Surrogate::train(hipStream_t s, int iterations, int B, double lr0, uint64_t seed,
                            double* final_loss) {
  hipError_t e;
  if ((e = ensure_params()) || (e = ensure_batch(B))) return e;
  if ((e = refresh(s))) return e;
  const int warm = std::min(200, iterations / 10 + 1);
  double loss = 0;
  for (int it = 0; it < iterations; ++it) {
    // linear warm-up, then cosine decay to lr0 / 100
    double lr = lr0;
    if (it < warm) lr = lr0 * double(it + 1) / warm;
    else lr = lr0 * (0.01 + 0.99 * 0.5 * (1 + std::cos(M_PI * double(it - warm) /
                                                      double(std::max(1, iterations - warm)))));
    k_sur_batch<<<(B + 255) / 256, 256, 0, s>>>(X, T, nsamples, B, seed, it, 1, X0, X0T, Tb);
    if ((e = pass(s, B, false, float(lr)))) return e;
  }
  if ((e = refresh(s))) return e;
  // loss of the last batch (after the final update, on the same batch)
  if ((e = batch_loss(s, B, &loss))) return e;
  if (final_loss) *final_loss = loss;
  return hipSuccess;
}

hipError_t Surrogate::batch_loss(hipStream_t s, int B, double* loss) {
  // forward only, reusing the pass kernels' first half
  const int64_t jbs = B / 32;
  k_sur_fwd<K0, HID, true><<<blocks_for_waves(4 * jbs), 256, 0, s>>>(
      X0, Wbf[0], params + OFF_B0, HID, B, Zt[0], H[1], HT[1], nullptr);
  for (int l = 1; l <= 3; ++l)
    k_sur_fwd<HID, HID, true><<<blocks_for_waves(4 * jbs), 256, 0, s>>>(
        H[l], Wbf[l], params + (l == 1 ? OFF_B1 : l == 2 ? OFF_B2 : OFF_B3), HID, B, Zt[l],
        H[l + 1], HT[l + 1], nullptr);
  k_sur_fwd<HID, 32, false><<<blocks_for_waves(jbs), 256, 0, s>>>(
      H[4], Wbf[4], params + OFF_B4, NOUT, B, nullptr, nullptr, nullptr, Yb);
  const int nb = (B + 255) / 256;
  k_sur_loss<<<nb, 256, 0, s>>>(Yb, Tb, B, dY, dYT, loss_part);
  hipError_t e = hipGetLastError();
  if (e) return e;
  std::vector<float> h(nb);
  if ((e = hipMemcpyAsync(h.data(), loss_part, nb * 4, hipMemcpyDeviceToHost, s)) ||
      (e = hipStreamSynchronize(s)))
    return e;
  double sum = 0;
  for (float v : h) sum += v;
  *loss = sum / (7.0 * B);
  return hipSuccess;
}

hipError_t Surrogate::gradients(hipStream_t s, int64_t n, const float* x_host, const float* t_host,
                                float* grad_host, double* loss) {
  hipError_t e;
  const int B = int(n);
  if ((e = ensure_params()) || (e = ensure_batch(B)) || (e = ensure_io(n))) return e;
  if ((e = refresh(s))) return e;
  float* tdev = io_y;
  if ((e = hipMemcpyAsync(io_x, x_host, size_t(n) * 7 * 4, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(tdev, t_host, size_t(n) * 7 * 4, hipMemcpyHostToDevice, s)))
    return e;
  k_sur_batch<<<(B + 255) / 256, 256, 0, s>>>(io_x, tdev, n, B, 0, 0, 0, X0, X0T, Tb);
  if ((e = pass(s, B, true, 0.f))) return e;
  if ((e = hipMemcpyAsync(grad_host, grad, NPARAM * 4, hipMemcpyDeviceToHost, s))) return e;
  // loss with the same (unchanged) parameters
  return batch_loss(s, B, loss);
}

hipError_t Surrogate::ensure_io(int64_t n) {
  if (n <= io_cap) return hipSuccess;
  dfree_p(io_x);
  dfree_p(io_y);
  hipError_t e;
  if ((e = dalloc_n(&io_x, size_t(n) * 7)) || (e = dalloc_n(&io_y, size_t(n) * 7))) return e;
  io_cap = n;
  return hipSuccess;
}

hipError_t Surrogate::infer(hipStream_t s, int64_t n, const float* x_host, float* y_host) {
  hipError_t e;
  if ((e = ensure_params()) || (e = ensure_io(n))) return e;
  if ((e = hipMemcpyAsync(io_x, x_host, size_t(n) * 7 * 4, hipMemcpyHostToDevice, s))) return e;
  if ((e = launch_sur_mlp(s, img, imgb, infer_bf16, n, io_x, 7, io_y, 7))) return e;
  if ((e = hipMemcpyAsync(y_host, io_y, size_t(n) * 7 * 4, hipMemcpyDeviceToHost, s))) return e;
  return hipStreamSynchronize(s);
}

hipError_t launch_sur_mlp(hipStream_t s, const void* img, const float* imgb, bool bf16,
                          int64_t n, const float* x, int sx, float* y, int sy) {
  (void)hipGetLastError();
  if (n <= 0) return hipSuccess;
  // one block of 8 waves per CU (the weight image fills most of the LDS):
  // two waves per SIMD, so one wave's activations overlap the other's MFMAs
  const int64_t tiles = (n + 31) / 32;
  const unsigned g = unsigned(std::min<int64_t>(256, (tiles + 7) / 8));
  if (bf16)
    k_sur_infer<Bf16><<<g, 512, 0, s>>>(static_cast<const Bf16::x8*>(img), imgb, n, x, sx, y,
                                        sy);
  else
    k_sur_infer<Fp16><<<g, 512, 0, s>>>(static_cast<const Fp16::x8*>(img), imgb, n, x, sx, y,
                                        sy);
  return hipGetLastError();
}

hipError_t Surrogate::ensure_points(int64_t n) {
  if (n <= pt_cap) return hipSuccess;
  dfree_p(pt_x);
  dfree_p(pt_y);
  hipError_t e;
  if ((e = dalloc_n(&pt_x, size_t(n) * 8)) || (e = dalloc_n(&pt_y, size_t(n) * 8))) return e;
  pt_cap = n;
  return hipSuccess;
}

hipError_t Surrogate::set_params(hipStream_t s, const float* host) {
  hipError_t e;
  if ((e = ensure_params())) return e;
  if ((e = hipMemcpyAsync(params, host, NPARAM * 4, hipMemcpyHostToDevice, s))) return e;
  if ((e = hipMemsetAsync(m1, 0, NPARAM * 4, s)) || (e = hipMemsetAsync(m2, 0, NPARAM * 4, s)))
    return e;
  step = 0;
  if ((e = refresh(s))) return e;
  return hipStreamSynchronize(s);
}

hipError_t Surrogate::get_params(hipStream_t s, float* host) {
  hipError_t e;
  if ((e = ensure_params())) return e;
  if ((e = hipMemcpyAsync(host, params, NPARAM * 4, hipMemcpyDeviceToHost, s))) return e;
  return hipStreamSynchronize(s);
}

}  // namespace akx
