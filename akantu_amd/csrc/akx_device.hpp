// akx_device.hpp -- device-side element tables and per-point arithmetic.
//
// Two families of per-point routines:
//  * *_ref (double): the reference's own operation order (non-BLAS loops of
//    src/common/aka_math_tmpl.hh, material laws of material_elastic_inline_impl.hh
//    and material_neohookean_inline_impl.hh, finite-strain residual of
//    material.cc:637-718 with aka_voigthelper_tmpl.hh).  Compiled with
//    -ffp-contract=off they reproduce the reference arithmetic bit for bit
//    (log() aside, whose last bit may differ from glibc's).
//  * *_f32 (float): the mixed-precision element math.  Displacements enter as
//    differences to the element's first node (exact in FP64, then rounded), and
//    the finite-strain law is evaluated in terms of small quantities (Green
//    strain, J - 1) so that FP32 keeps its relative accuracy at small strain.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace akx {

enum ElementType { TRI3 = 0, QUAD4 = 1, TET4 = 2, HEX8 = 3 };
enum MaterialKind { ELASTIC = 0, NEOHOOKEAN = 1, NH_DAMAGE = 2 };

template <int ET> struct Elem;
template <> struct Elem<TRI3> { static constexpr int d = 2, n = 3, Q = 1; };
template <> struct Elem<QUAD4> { static constexpr int d = 2, n = 4, Q = 4; };
template <> struct Elem<TET4> { static constexpr int d = 3, n = 4, Q = 1; };
template <> struct Elem<HEX8> { static constexpr int d = 3, n = 8, Q = 8; };

// Material parameters as the device sees them (one entry per material).
struct MatParams {
  double lambda, mu, rho, Sd;
  float lambda_f, mu_f, Sd_f, inv_Sd_f;
  int kind;
  int pad[3];
};

// Element reference data computed on the host in FP64 (natural coordinates,
// dN/dxi, N and weights of every quadrature point).
struct ElemTables {
  double dnds[8][24];  // [q][i + a*d]
  double N[8][8];      // [q][a]
  double w[8];
};

// FP32 copy of dN/dxi and the weights for the mixed kernels that rebuild B.
struct ElemTablesF {
  float dnds[8][24];
  float w[8];
};

// ---------------------------------------------------------------------------
// XCD-aware block index: consecutive logical blocks land on the same XCD (the
// hardware deals workgroups round-robin over the 8 XCDs), so neighbouring
// elements / nodes share one L2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t logical_block(int remap) {
  const int64_t b = blockIdx.x, nb = gridDim.x;
  if (!remap || nb < 8) return b;
  const int64_t q = nb >> 3, r = nb & 7, x = b & 7, l = b >> 3;
  return (x < r) ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
}

// ---------------------------------------------------------------------------
// reference-order FP64 helpers (column-major)
// ---------------------------------------------------------------------------
template <int im, int in, int ik>
__device__ __forceinline__ void matrix_matrix(const double* A, const double* B, double* C) {
#pragma unroll
  for (int j = 0; j < in; ++j)
#pragma unroll
    for (int i = 0; i < im; ++i) {
      double s = 0.;
#pragma unroll
      for (int l = 0; l < ik; ++l) s += A[i + l * im] * B[l + j * ik];
      C[i + j * im] = s;
    }
}
template <int im, int in, int ik>
__device__ __forceinline__ void matrix_matrixt(const double* A, const double* B, double* C) {
#pragma unroll
  for (int j = 0; j < in; ++j)
#pragma unroll
    for (int i = 0; i < im; ++i) {
      double s = 0.;
#pragma unroll
      for (int l = 0; l < ik; ++l) s += A[i + l * im] * B[j + l * in];
      C[i + j * im] = s;
    }
}
template <int im, int in, int ik>
__device__ __forceinline__ void matrixt_matrix(const double* A, const double* B, double* C) {
#pragma unroll
  for (int j = 0; j < in; ++j)
#pragma unroll
    for (int i = 0; i < im; ++i) {
      double s = 0.;
#pragma unroll
      for (int l = 0; l < ik; ++l) s += A[l + i * ik] * B[l + j * ik];
      C[i + j * im] = s;
    }
}

template <typename T> __device__ __forceinline__ T det2(const T* m) { return m[0] * m[3] - m[1] * m[2]; }
template <typename T> __device__ __forceinline__ T det3(const T* m) {
  return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
         m[6] * (m[1] * m[5] - m[4] * m[2]);
}
template <int d, typename T> __device__ __forceinline__ T detn(const T* m) {
  if constexpr (d == 2) return det2(m); else return det3(m);
}
template <int d, typename T> __device__ __forceinline__ void invn(const T* m, T* r) {
  if constexpr (d == 2) {
    T dd = det2(m);
    r[0] = m[3] / dd; r[1] = -m[1] / dd; r[2] = -m[2] / dd; r[3] = m[0] / dd;
  } else {
    T dd = det3(m);
    r[0] = (m[4] * m[8] - m[7] * m[5]) / dd;
    r[1] = (m[2] * m[7] - m[8] * m[1]) / dd;
    r[2] = (m[1] * m[5] - m[4] * m[2]) / dd;
    r[3] = (m[5] * m[6] - m[8] * m[3]) / dd;
    r[4] = (m[0] * m[8] - m[6] * m[2]) / dd;
    r[5] = (m[2] * m[3] - m[5] * m[0]) / dd;
    r[6] = (m[3] * m[7] - m[6] * m[4]) / dd;
    r[7] = (m[1] * m[6] - m[7] * m[0]) / dd;
    r[8] = (m[0] * m[4] - m[3] * m[1]) / dd;
  }
}

// MaterialElastic::computeStressOnQuad (material_elastic_inline_impl.hh:41-55)
template <int d>
__device__ __forceinline__ void elastic_stress_ref(const MatParams& m, const double* G, double* S) {
  double trace = 0.;
#pragma unroll
  for (int i = 0; i < d; ++i) trace += G[i + i * d];
#pragma unroll
  for (int i = 0; i < d; ++i)
#pragma unroll
    for (int j = 0; j < d; ++j) {
      double v = (i == j) ? m.lambda * trace : 0.;
      S[i + j * d] = v + m.mu * (G[i + j * d] + G[j + i * d]);
    }
}

// MaterialNeohookean::computeStressOnQuad (material_neohookean_inline_impl.hh:46-65)
template <int d>
__device__ __forceinline__ double neohookean_stress_ref(const MatParams& m, const double* G,
                                                        double* S) {
  double F[d * d], C[d * d], Ci[d * d];
#pragma unroll
  for (int i = 0; i < d; ++i)
#pragma unroll
    for (int j = 0; j < d; ++j) F[i + j * d] = (i == j) ? 1. + G[i + j * d] : G[i + j * d];
  matrixt_matrix<d, d, d>(F, F, C);
  double J = detn<d>(F);
  invn<d>(C, Ci);
  double lj = log(J);
#pragma unroll
  for (int i = 0; i < d; ++i)
#pragma unroll
    for (int j = 0; j < d; ++j) {
      double c = (m.lambda * lj - m.mu) * Ci[i + j * d];
      S[i + j * d] = (i == j) ? m.mu + c : c;
    }
  return J;
}

// neo-Hookean energy density (material_neohookean_inline_impl.hh:145-158)
template <int d>
__device__ __forceinline__ double neohookean_energy_ref(const MatParams& m, const double* G) {
  double F[d * d], C[d * d];
#pragma unroll
  for (int i = 0; i < d; ++i)
#pragma unroll
    for (int j = 0; j < d; ++j) F[i + j * d] = (i == j) ? 1. + G[i + j * d] : G[i + j * d];
  matrixt_matrix<d, d, d>(F, F, C);
  double J = detn<d>(F);
  double trC = 0.;
#pragma unroll
  for (int i = 0; i < d; ++i) trC += C[i + i * d];
  double lj = log(J);
  return 0.5 * m.lambda * (lj * lj) + m.mu * (-lj + 0.5 * (trC - d));
}

// Damage update of MaterialMarigo::computeDamageAndStressOnQuad
// (material_marigo_inline_impl.hh:67-76) with Y = neo-Hookean energy density.
template <int d>
__device__ __forceinline__ void nh_damage_stress_ref(const MatParams& m, const double* G, double* S,
                                                     double& dam, double Ydq) {
  neohookean_stress_ref<d>(m, G, S);
  double Y = neohookean_energy_ref<d>(m, G);
  double Fd = Y - Ydq - m.Sd * dam;
  if (Fd > 0) dam = (Y - Ydq) / m.Sd;
  dam = fmin(dam, 1.);
#pragma unroll
  for (int k = 0; k < d * d; ++k) S[k] *= 1 - dam;
}

// r = (B_L0 + B_L2(G))^T S_voigt: Material::assembleInternalForces<dim>
// (material.cc:689-706), VoigtHelper::transferBMatrixToSymVoigtBMatrix
// (aka_voigthelper_tmpl.hh:104-143), transferBMatrixToBL2 (:177-226), Voigt
// order xx, yy, zz, yz, xz, xy (aka_voigthelper.cc:52-61).  The entries of the
// dense 6 x 3n matrices are formed exactly as the reference forms them, then
// contracted with S in the order of Math::matrixt_matrix.
template <int d, int n>
__device__ __forceinline__ void finite_residual_ref(const double* B, const double* G,
                                                    const double* S, double* r) {
  auto Bm = [&](int i, int a) { return B[i + a * d]; };
  auto Gm = [&](int i, int j) { return G[i + j * d]; };
  if constexpr (d == 3) {
    const double Sv[6] = {S[0], S[4], S[8], S[1 + 2 * 3], S[0 + 2 * 3], S[0 + 1 * 3]};
#pragma unroll
    for (int a = 0; a < n; ++a)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double bt[6];
#pragma unroll
        for (int I = 0; I < 3; ++I) {
          double bl2 = Gm(k, I) * Bm(I, a);
          bt[I] = (k == I) ? Bm(I, a) + bl2 : 0. + bl2;
        }
        // I = 3 (yz): sym (k=1: dndz, k=2: dndy); BL2 += G(k,1) B(2,a); += G(k,2) B(1,a)
        {
          double bl2 = 0.;
          bl2 += Gm(k, 1) * Bm(2, a);
          bl2 += Gm(k, 2) * Bm(1, a);
          double sym = (k == 1) ? Bm(2, a) : (k == 2) ? Bm(1, a) : 0.;
          bt[3] = sym + bl2;
        }
        // I = 4 (xz): sym (k=0: dndz, k=2: dndx); BL2 += G(k,0) B(2,a); += G(k,2) B(0,a)
        {
          double bl2 = 0.;
          bl2 += Gm(k, 0) * Bm(2, a);
          bl2 += Gm(k, 2) * Bm(0, a);
          double sym = (k == 0) ? Bm(2, a) : (k == 2) ? Bm(0, a) : 0.;
          bt[4] = sym + bl2;
        }
        // I = 5 (xy): sym (k=0: dndy, k=1: dndx); BL2 += G(k,0) B(1,a); += G(k,1) B(0,a)
        {
          double bl2 = 0.;
          bl2 += Gm(k, 0) * Bm(1, a);
          bl2 += Gm(k, 1) * Bm(0, a);
          double sym = (k == 0) ? Bm(1, a) : (k == 1) ? Bm(0, a) : 0.;
          bt[5] = sym + bl2;
        }
        double s = 0.;
#pragma unroll
        for (int I = 0; I < 6; ++I) s += bt[I] * Sv[I];
        r[3 * a + k] = s;
      }
  } else {
    const double Sv[3] = {S[0], S[3], S[0 + 1 * 2]};
#pragma unroll
    for (int a = 0; a < n; ++a)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        double bt[3];
#pragma unroll
        for (int I = 0; I < 2; ++I) {
          double bl2 = Gm(k, I) * Bm(I, a);
          bt[I] = (k == I) ? Bm(I, a) + bl2 : 0. + bl2;
        }
        double bl2 = 0.;
        bl2 += Gm(k, 0) * Bm(1, a);
        bl2 += Gm(k, 1) * Bm(0, a);
        double sym = (k == 0) ? Bm(1, a) : Bm(0, a);
        bt[2] = sym + bl2;
        double s = 0.;
#pragma unroll
        for (int I = 0; I < 3; ++I) s += bt[I] * Sv[I];
        r[2 * a + k] = s;
      }
  }
}

// ---------------------------------------------------------------------------
// FP32 element math (mixed precision)
// ---------------------------------------------------------------------------
// x - log1p(x) for the small |x| of J - 1, accurate to a few ulp relative.
__device__ __forceinline__ float x_minus_log1p(float x) {
  if (fabsf(x) < 0.0625f) {
    // x - log(1 + x) = x^2 (1/2 - x/3 + x^2/4 - ...), Horner over k = 12..2;
    // the first dropped term is below 1e-13 of the sum for |x| < 1/16.
    float acc = 0.f;
#pragma unroll
    for (int k = 12; k >= 2; --k) acc = fmaf(acc, x, ((k & 1) ? -1.f : 1.f) / float(k));
    return acc * x * x;
  }
  return x - log1pf(x);
}

// log(1 + x): for |x| < 1/16 as 2 atanh(s), s = x / (2 + x), |s| < 1/31, by
// its series to s^7 (first dropped term below 1e-11 relative; s carries the
// reciprocal's ~1 ulp); else log1pf.
__device__ __forceinline__ float log1p_small_f32(float x) {
#pragma clang fp contract(fast)
  if (fabsf(x) < 0.0625f) {
    const float s = x * __builtin_amdgcn_rcpf(2.f + x), s2 = s * s;
    const float p = fmaf(fmaf(fmaf(s2, 1.f / 7.f, 1.f / 5.f), s2, 1.f / 3.f), s2, 1.f);
    return 2.f * s * p;
  }
  return log1pf(x);
}

// neo-Hookean second Piola-Kirchhoff stress from G in small-quantity form:
//   E = (G + G^T + G^T G)/2,  C = I + 2E,  x = J - 1 from the invariants of G,
//   S = C^{-1} (2 mu E + lambda log(1 + x) I)
// Algebraically equal to mu I + (lambda ln J - mu) C^{-1}.  Returns x = J - 1
// and, if W != nullptr, the energy density
//   W = lambda/2 ln^2 J + mu (tr E - ln J).
template <int d>
__device__ __forceinline__ float neohookean_stress_f32(float lambda, float mu, const float* G,
                                                       float* S, float* W) {
  float E[d * d], C[d * d], Ci[d * d];
#pragma unroll
  for (int i = 0; i < d; ++i)
#pragma unroll
    for (int j = 0; j < d; ++j) {
      float gtg = 0.f;
#pragma unroll
      for (int l = 0; l < d; ++l) gtg += G[l + i * d] * G[l + j * d];
      E[i + j * d] = 0.5f * (G[i + j * d] + G[j + i * d] + gtg);
      C[i + j * d] = 2.f * E[i + j * d] + (i == j ? 1.f : 0.f);
    }
  float x;
  if constexpr (d == 3) {
    float tr = G[0] + G[4] + G[8];
    float i2 = G[0] * G[4] + G[4] * G[8] + G[0] * G[8] - G[3] * G[1] - G[7] * G[5] - G[6] * G[2];
    x = tr + i2 + det3(G);
  } else {
    float tr = G[0] + G[3];
    x = tr + det2(G);
  }
  invn<d>(C, Ci);
  float lj = log1pf(x);
  float T[d * d];
#pragma unroll
  for (int k = 0; k < d * d; ++k) T[k] = 2.f * mu * E[k];
#pragma unroll
  for (int i = 0; i < d; ++i) T[i + i * d] += lambda * lj;
  // S = Ci * T (both symmetric)
#pragma unroll
  for (int i = 0; i < d; ++i)
#pragma unroll
    for (int j = 0; j < d; ++j) {
      float s = 0.f;
#pragma unroll
      for (int l = 0; l < d; ++l) s += Ci[i + l * d] * T[l + j * d];
      S[i + j * d] = s;
    }
  if (W) {
    // tr E - ln J = (tr E - x) + (x - ln(1 + x));  tr E - x = |G|^2/2 - i2 - det G
    float g2 = 0.f;
#pragma unroll
    for (int k = 0; k < d * d; ++k) g2 += G[k] * G[k];
    float rest;
    if constexpr (d == 3) {
      float i2 = G[0] * G[4] + G[4] * G[8] + G[0] * G[8] - G[3] * G[1] - G[7] * G[5] - G[6] * G[2];
      rest = 0.5f * g2 - i2 - det3(G);
    } else {
      rest = 0.5f * g2 - det2(G);
    }
    *W = 0.5f * lambda * lj * lj + mu * (rest + x_minus_log1p(x));
  }
  return x;
}

// The same law in 3D on symmetric storage, for the kernels where the stress
// evaluation dominates: E, C^{-1} and S kept as their 6 distinct entries
// (xx, yy, zz, yz, xz, xy), C^{-1} through the adjugate and the hardware
// reciprocal.  S (C^{-1} and T = 2 mu E + lambda ln J I commute, so S is
// symmetric) is returned multiplied by `scale`.  Returns x = J - 1; W as
// neohookean_stress_f32.
__device__ __forceinline__ float neohookean_stress_sym3_f32(float lambda, float mu, const float* G,
                                                             float scale, float* S6, float* W) {
#pragma clang fp contract(fast)
  auto gtg = [&](int i, int j) {
    return G[3 * i] * G[3 * j] + G[3 * i + 1] * G[3 * j + 1] + G[3 * i + 2] * G[3 * j + 2];
  };
  float E[6];
  E[0] = G[0] + 0.5f * gtg(0, 0);
  E[1] = G[4] + 0.5f * gtg(1, 1);
  E[2] = G[8] + 0.5f * gtg(2, 2);
  E[3] = 0.5f * (G[7] + G[5] + gtg(1, 2));
  E[4] = 0.5f * (G[6] + G[2] + gtg(0, 2));
  E[5] = 0.5f * (G[3] + G[1] + gtg(0, 1));
  const float c00 = 1.f + 2.f * E[0], c11 = 1.f + 2.f * E[1], c22 = 1.f + 2.f * E[2];
  const float c12 = 2.f * E[3], c02 = 2.f * E[4], c01 = 2.f * E[5];
  float Ci[6];
  Ci[0] = c11 * c22 - c12 * c12;
  Ci[1] = c00 * c22 - c02 * c02;
  Ci[2] = c00 * c11 - c01 * c01;
  Ci[3] = c01 * c02 - c00 * c12;
  Ci[4] = c01 * c12 - c11 * c02;
  Ci[5] = c02 * c12 - c22 * c01;
  const float detc = c00 * Ci[0] + c01 * Ci[5] + c02 * Ci[4];
  const float rc = __builtin_amdgcn_rcpf(detc);
  const float tr = G[0] + G[4] + G[8];
  const float i2 = G[0] * G[4] + G[4] * G[8] + G[0] * G[8] - G[3] * G[1] - G[7] * G[5] - G[6] * G[2];
  const float dg = det3(G);
  const float x = tr + i2 + dg;
  const float lj = log1p_small_f32(x);
  const float tm = 2.f * mu, tl = lambda * lj;
  const float T0 = tm * E[0] + tl, T1 = tm * E[1] + tl, T2 = tm * E[2] + tl;
  const float T3 = tm * E[3], T4 = tm * E[4], T5 = tm * E[5];
  // S = Ci T, rows of the symmetric matrices as (xx xy xz / xy yy yz / xz yz zz)
  const float sc = scale * rc;
  S6[0] = sc * (Ci[0] * T0 + Ci[5] * T5 + Ci[4] * T4);
  S6[1] = sc * (Ci[5] * T5 + Ci[1] * T1 + Ci[3] * T3);
  S6[2] = sc * (Ci[4] * T4 + Ci[3] * T3 + Ci[2] * T2);
  S6[3] = sc * (Ci[5] * T4 + Ci[1] * T3 + Ci[3] * T2);
  S6[4] = sc * (Ci[0] * T4 + Ci[5] * T3 + Ci[4] * T2);
  S6[5] = sc * (Ci[0] * T5 + Ci[5] * T1 + Ci[4] * T3);
  if (W) {
    float g2 = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) g2 += G[k] * G[k];
    const float rest = 0.5f * g2 - i2 - dg;
    *W = 0.5f * lambda * lj * lj + mu * (rest + x_minus_log1p(x));
  }
  return x;
}

}  // namespace akx
