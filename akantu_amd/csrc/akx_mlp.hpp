// akx_mlp.hpp -- the constitutive surrogate network on gfx950 matrix cores.
//
// Network: 7 -> 128 -> 128 -> 128 -> 128 -> 7, the smoothstep swish
//   a(z) = z S(z),  S = t^2 (3 - 2 t),  t = clamp(k z + 1/2, 0, 1),  k = 1/6
// after every hidden layer (k is the FP16 value next to 1/6, so that the
// FP16 inference evaluates the same function the training fits); FP32
// accumulation and biases; training in BF16 (weights, activations, errors),
// inference in FP16 (the default) or BF16 (weights, activations).  S is a
// cubic smoothstep from 0 at
// z = -1/(2k) to 1 at z = 1/(2k), with S(0) = 1/2 and S'(0) = 1/4 as the
// logistic sigmoid; a is continuously differentiable (S' vanishes at both
// ends) and needs no transcendental: one clamped fused multiply-add, three
// multiplies and one multiply-add per value, on packed FP16 in inference.
//   inputs  x = (E_xx, E_yy, E_zz, E_yz, E_xz, E_xy) / eps_max, damage
//           (E = Green-Lagrange strain, tensor components, Voigt order of
//            aka_voigthelper.cc:52-61)
//   outputs y = ((S_NH - S_lin) / sS  in the same Voigt order,
//                (W_NH - W_lin) / sW)
//           S_lin = lambda tr(E) I + 2 mu E, W_lin = lambda/2 tr(E)^2 + mu E:E,
//           S_NH, W_NH the neo-Hookean PK2 stress and energy
//           (material_neohookean_inline_impl.hh:46-65, 145-158).
// The input layer sees every input twice, as its 16-bit value and as the
// 16-bit value of the rounding residual, with the same weights: its 16-wide K
// slot (7 + 7 + 2 zeros) costs nothing extra and keeps the inputs at ~16 bits
// (BF16) or ~22 bits (FP16).
//
// MFMA v_mfma_f32_32x32x16_bf16 / _f16 (one wave, 64 lanes, same layout):
//   A operand (32 x 16): lane l supplies row l % 32 and 8 values of the
//   K half l / 32; B operand (16 x 32): lane l supplies column l % 32 and 8
//   values of the K half l / 32.  The product sums A-slot s of lane (i, h)
//   against B-slot s of lane (j, h), whatever K label the hardware gives the
//   slots, so any K order works as long as A and B use the same one.
//   C/D (32 x 32, 16 floats per lane): lane l holds column l % 32, register v
//   holds row crow(v, l / 32).
// Layers are chained without moving data: features are the M (row) index,
// points the N (column) index, so a layer's C registers are, after bias and
// activation, the next layer's B operand in the K order chain_k; the weights are
// stored pre-permuted to that order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace akx {
namespace mlp {

constexpr int NIN = 7, NOUT = 7, HID = 128, K0 = 16;
// FP32 parameters, row-major W[out][in], then the bias
constexpr int OFF_W0 = 0;
constexpr int OFF_B0 = OFF_W0 + HID * NIN;
constexpr int OFF_W1 = OFF_B0 + HID;
constexpr int OFF_B1 = OFF_W1 + HID * HID;
constexpr int OFF_W2 = OFF_B1 + HID;
constexpr int OFF_B2 = OFF_W2 + HID * HID;
constexpr int OFF_W3 = OFF_B2 + HID;
constexpr int OFF_B3 = OFF_W3 + HID * HID;
constexpr int OFF_W4 = OFF_B3 + HID;
constexpr int OFF_B4 = OFF_W4 + NOUT * HID;
constexpr int NPARAM = OFF_B4 + NOUT;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// permuted inference weight image (units of 8 16-bit weights = one lane's operand)
constexpr int PW_L0 = 0;                      // [mb 4][lane 64]
constexpr int PW_L1 = PW_L0 + 4 * 64;         // [mb 4][t 8][lane 64]
constexpr int PW_L2 = PW_L1 + 4 * 8 * 64;
constexpr int PW_L3 = PW_L2 + 4 * 8 * 64;
constexpr int PW_L4 = PW_L3 + 4 * 8 * 64;     // [t 8][lane 64]
constexpr int PW_SIZE = PW_L4 + 8 * 64;       // 6912 vectors = 110,592 bytes
constexpr int PB_SIZE = 4 * HID + 8;          // biases b0..b3, b4 padded to 8

__host__ __device__ constexpr int crow(int v, int h) { return (v >> 2) * 8 + h * 4 + (v & 3); }
// feature index carried by B slot s of K step t of lane half h in a chained layer
__host__ __device__ constexpr int chain_k(int t, int h, int s) {
  return 32 * (t >> 1) + crow(8 * (t & 1) + s, h);
}

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

constexpr float ACT_K = 0.1666259765625f;  // the FP16 value next to 1/6

// the activation and its derivative in FP32 (training)
__device__ __forceinline__ float act(float z) {
  const float t = fminf(fmaxf(ACT_K * z + 0.5f, 0.f), 1.f);
  return z * (t * t * (3.f - 2.f * t));
}
__device__ __forceinline__ float act_grad(float z) {
  const float u = ACT_K * z + 0.5f;
  const float t = fminf(fmaxf(u, 0.f), 1.f);
  const float dS = (u > 0.f && u < 1.f) ? 6.f * ACT_K * t * (1.f - t) : 0.f;
  return t * t * (3.f - 2.f * t) + z * dS;
}

__device__ __forceinline__ __bf16 to_bf16(float x) { return static_cast<__bf16>(x); }
__device__ __forceinline__ float from_bf16(__bf16 x) { return static_cast<float>(x); }

// ---------------------------------------------------------------------------
// Inference (k_sur_infer), in one of two 16-bit formats, FP32 accumulation and
// the bias as the MFMA's C operand of a layer's first K step in both:
//   Fp16: FP16 weights and activations on v_mfma_f32_32x32x16_f16, the
//         activation on packed FP16: per pair of values one v_cvt_pk_f16_f32
//         (to nearest), one v_pk_fma_f16 with the clamp bit for t, and four
//         v_pk_mul_f16 / v_pk_fma_f16;
//   Bf16: BF16 weights and activations on v_mfma_f32_32x32x16_bf16 (the
//         training kernels' format); gfx950 has no packed BF16 arithmetic, so
//         the activation runs in FP32, one value per instruction, and each
//         result is rounded once to BF16 (v_cvt_pk_bf16_f32, to nearest).
struct Fp16 {
  typedef _Float16 S;
  typedef _Float16 x8 __attribute__((ext_vector_type(8)));
  typedef _Float16 x4 __attribute__((ext_vector_type(4)));
  typedef _Float16 x2 __attribute__((ext_vector_type(2)));
  static __device__ __forceinline__ f32x16 mfma(x8 a, x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ x2 splat2(float v) {
    const S h = static_cast<S>(v);
    return x2{h, h};
  }
  // a(z) for the two accumulator values z0, z1: Z (S), with S = (t t)(3 - 2 t)
  static __device__ __forceinline__ x2 act2(float z0, float z1) {
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    const x2 Z = __builtin_convertvector(f32x2{z0, z1}, x2);
    x2 t = __builtin_elementwise_fma(Z, splat2(ACT_K), splat2(0.5f));
    t = __builtin_elementwise_min(__builtin_elementwise_max(t, splat2(0.f)), splat2(1.f));
    const x2 u = __builtin_elementwise_fma(t, splat2(-2.f), splat2(3.f));
    return Z * ((t * t) * u);
  }
};

struct Bf16 {
  typedef __bf16 S;
  typedef __bf16 x8 __attribute__((ext_vector_type(8)));
  typedef __bf16 x4 __attribute__((ext_vector_type(4)));
  typedef __bf16 x2 __attribute__((ext_vector_type(2)));
  static __device__ __forceinline__ f32x16 mfma(x8 a, x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
  // a(z0), a(z1) in FP32 (t = clamp(fma(z, k, 1/2)), u = fma(t, -2, 3),
  // a = z ((t t) u)), each rounded to BF16
  static __device__ __forceinline__ x2 act2(float z0, float z1) {
    const float t0 = fminf(fmaxf(__builtin_fmaf(z0, ACT_K, 0.5f), 0.f), 1.f);
    const float t1 = fminf(fmaxf(__builtin_fmaf(z1, ACT_K, 0.5f), 0.f), 1.f);
    const float a0 = z0 * ((t0 * t0) * __builtin_fmaf(t0, -2.f, 3.f));
    const float a1 = z1 * ((t1 * t1) * __builtin_fmaf(t1, -2.f, 3.f));
    return x2{static_cast<S>(a0), static_cast<S>(a1)};
  }
};

// the Fp16 operand vector: the unit in which the weight image is stored
typedef Fp16::x8 f16x8;

template <class F>
__device__ __forceinline__ typename F::x8 join4(typename F::x2 a, typename F::x2 b,
                                                typename F::x2 c, typename F::x2 d) {
  const typename F::x4 ab = __builtin_shufflevector(a, b, 0, 1, 2, 3);
  const typename F::x4 cd = __builtin_shufflevector(c, d, 0, 1, 2, 3);
  return __builtin_shufflevector(ab, cd, 0, 1, 2, 3, 4, 5, 6, 7);
}

// Weights (16-bit, permuted as the chained layout needs) and biases (FP32, in
// feature order) of one network, as staged in LDS by the inference kernel.
template <class F>
struct Image {
  typename F::x8 w[PW_SIZE];
  float b[PB_SIZE];
};

// The C operand holding layer `layer`'s bias for row block mb: register v of
// lane half h is feature 32 mb + crow(v, h); registers 4g..4g+3 are the four
// consecutive features 32 mb + 8 g + 4 h + 0..3 (one 16-byte LDS read).
__device__ __forceinline__ f32x16 bias_block(const float* B, int layer, int mb, int h) {
  f32x16 c;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const float4 bb = *reinterpret_cast<const float4*>(&B[layer * HID + 32 * mb + 8 * g + 4 * h]);
    c[4 * g + 0] = bb.x;
    c[4 * g + 1] = bb.y;
    c[4 * g + 2] = bb.z;
    c[4 * g + 3] = bb.w;
  }
  return c;
}

// g of the 16 values of one C row block: act slots of K steps 2 mb (registers
// 0-7) and 2 mb + 1 (registers 8-15) of the next layer (chain_k order)
template <class F>
__device__ __forceinline__ void activate_block(const f32x16& a, typename F::x8& lo,
                                               typename F::x8& hi) {
  lo = join4<F>(F::act2(a[0], a[1]), F::act2(a[2], a[3]), F::act2(a[4], a[5]),
                F::act2(a[6], a[7]));
  hi = join4<F>(F::act2(a[8], a[9]), F::act2(a[10], a[11]), F::act2(a[12], a[13]),
                F::act2(a[14], a[15]));
}

// Row block mb of a hidden layer: its 8 K steps chained on one accumulator,
// which starts from the bias.
template <class F>
__device__ __forceinline__ f32x16 hidden_block(const float* B, const typename F::x8* wl,
                                               const typename F::x8 (&in)[8], int layer, int mb,
                                               int h) {
  f32x16 c = bias_block(B, layer, mb, h);
#pragma unroll
  for (int t = 0; t < 8; ++t) c = F::mfma(wl[(mb * 8 + t) * 64], in[t], c);
  return c;
}

// One hidden layer, in -> out.  The row blocks run one after the other so
// that the activation of block mb (vector pipe) overlaps the MFMAs of block
// mb + 1 (matrix pipe) within the wave.
template <class F>
__device__ __forceinline__ void hidden_layer(const typename F::x8* W, const float* B, int layer,
                                             int lane, int h, const typename F::x8 (&in)[8],
                                             typename F::x8 (&out)[8]) {
  const typename F::x8* wl = W + PW_L1 + (layer - 1) * (PW_L2 - PW_L1) + lane;
  f32x16 a0 = hidden_block<F>(B, wl, in, layer, 0, h);
  f32x16 a1 = hidden_block<F>(B, wl, in, layer, 1, h);
  activate_block<F>(a0, out[0], out[1]);
  f32x16 a2 = hidden_block<F>(B, wl, in, layer, 2, h);
  activate_block<F>(a1, out[2], out[3]);
  f32x16 a3 = hidden_block<F>(B, wl, in, layer, 3, h);
  activate_block<F>(a2, out[4], out[5]);
  activate_block<F>(a3, out[6], out[7]);
}

// Evaluate the network for the 32 points of one wave: lanes j and j + 32 both
// carry point j's inputs x[0..6] and both receive its outputs y[0..6].  Every
// lane of the wave must call it (the MFMAs are wave-wide).
template <class F>
__device__ __forceinline__ void eval_wave(const Image<F>& L, const float* x, float* y) {
  typedef typename F::S S;
  typedef typename F::x8 V;
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5;
  // A zero offset that passes through an empty asm statement on every call:
  // the weight reads cannot be hoisted out of the caller's loop over tiles
  // (they would not fit in registers and would spill).
  int zero = 0;
  asm volatile("" : "+v"(zero));
  const V* W = L.w + zero;
  const float* B = L.b + zero;
  // input layer: K slot 8h + s holds x_hi[k] (k < 7), x_lo[k - 7] (7 <= k < 14), 0;
  // 16-bit round to nearest, x_lo = x - x_hi exactly in FP32
  S xh[NIN], xl[NIN];
#pragma unroll
  for (int k = 0; k < NIN; ++k) {
    xh[k] = static_cast<S>(x[k]);
    xl[k] = static_cast<S>(x[k] - static_cast<float>(xh[k]));
  }
  V bx;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const S first = s < NIN ? xh[s] : xl[s - NIN];                       // k = s
    const S second = s + 8 < 2 * NIN ? xl[s + 8 - NIN] : static_cast<S>(0.f);  // k = 8 + s
    bx[s] = h ? second : first;
  }
  V actA[8], actB[8];  // the activations, ping-pong between the layers
  {
    f32x16 a[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
      a[mb] = F::mfma(W[PW_L0 + mb * 64 + lane], bx, bias_block(B, 0, mb, h));
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) activate_block<F>(a[mb], actA[2 * mb], actA[2 * mb + 1]);
  }
  hidden_layer<F>(W, B, 1, lane, h, actA, actB);
  hidden_layer<F>(W, B, 2, lane, h, actB, actA);
  hidden_layer<F>(W, B, 3, lane, h, actA, actB);
  f32x16 o = f32x16{};
#pragma unroll
  for (int t = 0; t < 8; ++t) o = F::mfma(W[PW_L4 + t * 64 + lane], actB[t], o);
  // rows 0..3 sit in registers 0..3 of half 0, rows 4..7 in registers 0..3 of half 1
  float mine[4], other[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    mine[v] = o[v];
    other[v] = __shfl_xor(o[v], 32, 64);
  }
  const float* lo = h ? other : mine;
  const float* hi = h ? mine : other;
#pragma unroll
  for (int r = 0; r < 4; ++r) y[r] = lo[r] + B[4 * HID + r];
#pragma unroll
  for (int r = 4; r < NOUT; ++r) y[r] = hi[r - 4] + B[4 * HID + r];
}

// Cooperative copy of a permuted image (global) into LDS by one block.
template <class F>
__device__ __forceinline__ void stage_image(Image<F>& L, const typename F::x8* __restrict__ w,
                                            const float* __restrict__ b) {
  for (int i = threadIdx.x; i < PW_SIZE; i += blockDim.x) L.w[i] = w[i];
  for (int i = threadIdx.x; i < PB_SIZE; i += blockDim.x) L.b[i] = b[i];
}

}  // namespace mlp
}  // namespace akx
