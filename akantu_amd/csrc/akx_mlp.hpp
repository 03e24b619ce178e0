// akx_mlp.hpp -- the constitutive surrogate network on gfx950 matrix cores.
//
// Network: 7 -> 128 -> 128 -> 128 -> 128 -> 7, SiLU after every hidden layer,
// BF16 weights and activations, FP32 accumulation and biases.
//   inputs  x = (E_xx, E_yy, E_zz, E_yz, E_xz, E_xy) / eps_max, damage
//           (E = Green-Lagrange strain, tensor components, Voigt order of
//            aka_voigthelper.cc:52-61)
//   outputs y = ((S_NH - S_lin) / sS  in the same Voigt order,
//                (W_NH - W_lin) / sW)
//           S_lin = lambda tr(E) I + 2 mu E, W_lin = lambda/2 tr(E)^2 + mu E:E,
//           S_NH, W_NH the neo-Hookean PK2 stress and energy
//           (material_neohookean_inline_impl.hh:46-65, 145-158).
// The input layer sees every input twice, as its BF16 value and as the BF16
// value of the rounding residual, with the same weights: its 16-wide K slot
// (7 + 7 + 2 zeros) costs nothing extra and keeps the inputs at ~16 bits.
//
// MFMA v_mfma_f32_32x32x16_bf16 (one wave, 64 lanes):
//   A operand (32 x 16): lane l supplies row l % 32 and 8 values of the
//   K half l / 32; B operand (16 x 32): lane l supplies column l % 32 and 8
//   values of the K half l / 32.  The product sums A-slot s of lane (i, h)
//   against B-slot s of lane (j, h), whatever K label the hardware gives the
//   slots, so any K order works as long as A and B use the same one.
//   C/D (32 x 32, 16 floats per lane): lane l holds column l % 32, register v
//   holds row crow(v, l / 32).
// Layers are chained without moving data: features are the M (row) index,
// points the N (column) index, so a layer's C registers are, after bias and
// SiLU, the next layer's B operand in the K order chain_k; the weights are
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

// permuted BF16 weight image (units of bf16x8 = 8 weights = one lane's operand)
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

// sigmoid through the hardware exp2 and reciprocal (v_exp_f32, v_rcp_f32, about
// 1 ulp each): the result is rounded to BF16 right after, so an IEEE division
// (a ten-instruction sequence per value) would buy nothing
__device__ __forceinline__ float sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigmoid(z);
  return s * (1.f + z * (1.f - s));
}

__device__ __forceinline__ __bf16 to_bf16(float x) { return static_cast<__bf16>(x); }
__device__ __forceinline__ float from_bf16(__bf16 x) { return static_cast<float>(x); }

// Weights and biases of one network, as staged in LDS by the kernels that
// evaluate it.
struct Image {
  bf16x8 w[PW_SIZE];
  float b[PB_SIZE];
};

// Evaluate the network for the 32 points of one wave: lanes j and j + 32 both
// carry point j's inputs x[0..6] and both receive its outputs y[0..6].  Every
// lane of the wave must call it (the MFMAs are wave-wide).
__device__ __forceinline__ void eval_wave(const Image& L, const float* x, float* y) {
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5;
  // input layer: K slot 8h + s holds x_hi[k] (k < 7), x_lo[k - 7] (7 <= k < 14), 0
  bf16x8 bx;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int k = 8 * h + s;
    float v = 0.f;
    if (k < NIN) {
      v = x[k];
    } else if (k < 2 * NIN) {
      const float xv = x[k - NIN];
      v = xv - from_bf16(to_bf16(xv));
    }
    bx[s] = to_bf16(v);
  }
  f32x16 acc[4];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    acc[mb] = f32x16{};
    acc[mb] = mfma(L.w[PW_L0 + mb * 64 + lane], bx, acc[mb]);
  }
  bf16x8 act[8];
  auto activate = [&](int layer) {
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        // slots 4g..4g+3 carry 4 consecutive features: one 16-byte bias read
        const float4 bb =
            *reinterpret_cast<const float4*>(&L.b[layer * HID + chain_k(t, h, 4 * g)]);
        const float bv[4] = {bb.x, bb.y, bb.z, bb.w};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int s = 4 * g + r;
          const float z = acc[t >> 1][8 * (t & 1) + s] + bv[r];
          act[t][s] = to_bf16(silu(z));
        }
      }
  };
  activate(0);
  // One copy of the hidden-layer code (unrolling the three layers lets the
  // scheduler hoist every weight load and spill).  K steps outermost: the
  // four MFMAs of a step use independent accumulators; the LDS latency of
  // the weight reads is covered by the other wave of the SIMD.
#pragma unroll 1
  for (int layer = 1; layer <= 3; ++layer) {
    const bf16x8* wl = L.w + PW_L1 + (layer - 1) * (PW_L2 - PW_L1) + lane;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) acc[mb] = f32x16{};
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      bf16x8 w[4];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) w[mb] = wl[(mb * 8 + t) * 64];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) acc[mb] = mfma(w[mb], act[t], acc[mb]);
    }
    activate(layer);
  }
  f32x16 o = f32x16{};
  bf16x8 w4[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) w4[t] = L.w[PW_L4 + t * 64 + lane];
#pragma unroll
  for (int t = 0; t < 8; ++t) o = mfma(w4[t], act[t], o);
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
  for (int r = 0; r < 4; ++r) y[r] = lo[r] + L.b[4 * HID + r];
#pragma unroll
  for (int r = 4; r < NOUT; ++r) y[r] = hi[r - 4] + L.b[4 * HID + r];
}

// Cooperative copy of a permuted image (global) into LDS by one block.
__device__ __forceinline__ void stage_image(Image& L, const bf16x8* __restrict__ w,
                                            const float* __restrict__ b) {
  for (int i = threadIdx.x; i < PW_SIZE; i += blockDim.x) L.w[i] = w[i];
  for (int i = threadIdx.x; i < PB_SIZE; i += blockDim.x) L.b[i] = b[i];
}

}  // namespace mlp
}  // namespace akx
