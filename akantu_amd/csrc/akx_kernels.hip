// akx_kernels.hip -- HIP kernels of the explicit central-difference step (gfx950).
//
// Per step two kernels run:
//   element kernel  (one thread per element): gather u of the element's nodes,
//     per quadrature point grad u -> stress -> B^T sigma, integrate with
//     det(J) w, write the element's n*d force "slots" (AoS per element);
//   node kernel     (one thread per node): gather the node's slots through the
//     node -> slot CSR in assembly order (no atomics: fixed summation order),
//     residual + lumped solve + corrector of this step, then Dirichlet
//     increments + predictor of the next step.
// Both are bandwidth-bound (a few flops per byte); the tables are SoA so that
// consecutive threads read consecutive addresses.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "akx_internal.hpp"
#include "akx_mlp.hpp"

namespace akx {

// Shape derivatives and det(J) w of one point from the element's reference
// coordinates: the arithmetic of k_precompute (element_class_tmpl.hh:282-355,
// integrator_gauss_inline_impl.hh:208-252), so the values are bit-identical to
// the tables it writes.
template <int ET>
__device__ __forceinline__ double point_shapes_ref(const ElemTables& tab, int q, const double* xe,
                                                   double* B) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n;
  double J[d * d], Ji[d * d];
  matrix_matrixt<d, d, n>(tab.dnds[q], xe, J);
  invn<d>(J, Ji);
  matrix_matrix<d, n, d>(Ji, tab.dnds[q], B);
  return detn<d>(J) * tab.w[q];
}

// ===========================================================================
// precompute: B = J^{-1} dN/dxi and det(J) w
//   ShapeLagrange::computeShapeDerivativesOnIntegrationPoints
//   (shape_lagrange_inline_impl.hh:204-252, :62-82; element_class_tmpl.hh:282-355),
//   IntegratorGauss::computeJacobiansOnIntegrationPoints + weights
//   (integrator_gauss_inline_impl.hh:208-252, 396-423)
// ===========================================================================
template <int ET>
__global__ void __launch_bounds__(256) k_precompute(int64_t E, const double* __restrict__ X,
                                                    const uint32_t* __restrict__ conn, ElemTables tab,
                                                    double* __restrict__ B64, double* __restrict__ J64,
                                                    float* __restrict__ B32, float* __restrict__ J32,
                                                    int* bad) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, Q = T::Q, dn = d * n;
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double xe[dn];
#pragma unroll
  for (int a = 0; a < n; ++a) {
    const int64_t node = conn[a * E + e];
#pragma unroll
    for (int i = 0; i < d; ++i) xe[i + a * d] = X[node * d + i];
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    double B[dn];
    const double jw = point_shapes_ref<ET>(tab, q, xe, B);
#pragma unroll
    for (int k = 0; k < dn; ++k) {
      B64[(int64_t(q) * dn + k) * E + e] = B[k];
      B32[(int64_t(q) * dn + k) * E + e] = float(B[k]);
    }
    J64[int64_t(q) * E + e] = jw;
    J32[int64_t(q) * E + e] = float(jw);
    if (jw < 0) *bad = 1;
  }
}

// ===========================================================================
// lumped mass by row sum (SolidMechanicsModel::assembleMassLumped,
// solid_mechanics_model_mass.cc:58-102 -> assembleLumpedRowSum,
// fe_engine_template_tmpl_field.hh:165-199, computeNtb
// shape_lagrange_inline_impl.hh:495-533): per node, in element-ascending order,
//   m += sum_q (rho N_a(q)) J_q
// ===========================================================================
template <int ET>
__global__ void __launch_bounds__(256) k_mass(int64_t N, int64_t E, const uint32_t* __restrict__ off,
                                              const uint32_t* __restrict__ slot, ElemTables tab,
                                              const double* __restrict__ J64,
                                              const uint16_t* __restrict__ elem_mat,
                                              const MatParams* __restrict__ mats,
                                              double* __restrict__ mass) {
  using T = Elem<ET>;
  constexpr int n = T::n, Q = T::Q;
  const int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (node >= N) return;
  double local = 0.;
  for (uint32_t p = off[node]; p < off[node + 1]; ++p) {
    const uint32_t s = slot[p];
    const int64_t e = s / n;
    const int a = s % n;
    const double rho = mats[elem_mat ? elem_mat[e] : 0].rho;
    double m = 0.;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      double ntb = 0.;
      ntb += rho * tab.N[q][a];
      m += ntb * J64[int64_t(q) * E + e];
    }
    local += m;
  }
  double r = 0.;
  r += local;
  mass[node] = r;
}

// ===========================================================================
// element force slots.  Two layouts:
//   element-ordered (slot_pos == nullptr): element e's n*d values at
//     slots[e*n*d ...], written as 16-byte vectors; the node kernel gathers
//     them through the CSR slot ids;
//   node-ordered: node a of element e goes to CSR position slot_pos[a][e], so
//     that every node finds its contributions contiguous, in assembly order.
// ===========================================================================
template <int d, int n, typename T>
__device__ __forceinline__ void store_slots(const ElemArgs& A, int64_t e, const T* fe) {
  constexpr int dn = d * n;
  if (A.slot_pos) {
    T* __restrict__ out = static_cast<T*>(A.slots);
#pragma unroll
    for (int a = 0; a < n; ++a) {
      const int64_t p = A.slot_pos[a * A.E + e];
#pragma unroll
      for (int i = 0; i < d; ++i) out[p * d + i] = fe[i + a * d];
    }
    return;
  }
  T* __restrict__ out = static_cast<T*>(A.slots) + e * dn;
  constexpr int per = 16 / sizeof(T);  // values per 16-byte store
  if constexpr (dn % per == 0) {
    using V = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    V* __restrict__ vo = reinterpret_cast<V*>(out);
#pragma unroll
    for (int k = 0; k < dn / per; ++k) {
      V v;
      if constexpr (sizeof(T) == 8) {
        v.x = fe[2 * k]; v.y = fe[2 * k + 1];
      } else {
        v.x = fe[4 * k]; v.y = fe[4 * k + 1]; v.z = fe[4 * k + 2]; v.w = fe[4 * k + 3];
      }
      vo[k] = v;
    }
  } else {
#pragma unroll
    for (int k = 0; k < dn; ++k) out[k] = fe[k];
  }
}

// ===========================================================================
// element kernel, FP64, reference operation order
// ===========================================================================
// RECOMPUTE = true: B and det(J) w rebuilt per point from the coordinates
// (point_shapes_ref) instead of read from the FP64 tables: E*Q*(d*n+1)*8
// bytes of table reads become one more gather of the nodes' coordinates.
// Element forces in the reference's FP64 operation order from the element's
// nodal displacements ue and (RECOMPUTE) coordinates xe, d x n column-major.
// Shared by the element kernel and the fused FP64 step.
template <int ET, int KIND, bool RECOMPUTE>
__device__ __forceinline__ void element_forces_fp64(const ElemArgs& A, int64_t e,
                                                    const MatParams& m, const double* ue,
                                                    const double* xe, double* fe) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, Q = T::Q, dn = d * n;
  const int64_t E = A.E;
  const double* __restrict__ Bt = static_cast<const double*>(A.B);
  const double* __restrict__ Jt = static_cast<const double*>(A.J);
#pragma unroll
  for (int k = 0; k < dn; ++k) fe[k] = 0.;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    double Bq[dn];
    double jq_rc = 0.;
    if constexpr (RECOMPUTE) {
      jq_rc = point_shapes_ref<ET>(*A.tab64, q, xe, Bq);
    } else {
#pragma unroll
      for (int k = 0; k < dn; ++k) Bq[k] = Bt[(int64_t(q) * dn + k) * E + e];
    }
    double G[d * d], S[d * d], R[dn];
    // grad u = u_e B^T (shape_functions_inline_impl.hh:388, Matrix::mul<false, true>)
    matrix_matrixt<d, d, n>(ue, Bq, G);
    if constexpr (KIND == ELASTIC) {
      elastic_stress_ref<d>(m, G, S);
      // B^T sigma computed as sigma B (shape_lagrange_inline_impl.hh:421)
      matrix_matrix<d, n, d>(S, Bq, R);
    } else {
      if constexpr (KIND == NEOHOOKEAN) {
        neohookean_stress_ref<d>(m, G, S);
      } else {
        const int64_t eq = e * Q + q;
        double dam = A.damage[eq];
        nh_damage_stress_ref<d>(m, G, S, dam, A.Yd[eq]);
        A.damage[eq] = dam;
      }
      finite_residual_ref<d, n>(Bq, G, S, R);
    }
    // IntegratorGauss::integrate (integrator_gauss_inline_impl.hh:446-452)
    const double jq = RECOMPUTE ? jq_rc : Jt[int64_t(q) * E + e];
#pragma unroll
    for (int k = 0; k < dn; ++k) fe[k] += R[k] * jq;
  }
}

template <int ET, int KIND, bool RECOMPUTE>
__global__ void __launch_bounds__(256) k_element_fp64(ElemArgs A) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, dn = d * n;
  const int64_t t = logical_block(A.xcd_remap) * blockDim.x + threadIdx.x;
  if (t >= A.count) return;
  const int64_t E = A.E;
  const int64_t e = A.filter ? int64_t(A.filter[t]) : t;
  const MatParams m = A.mats[A.elem_mat ? A.elem_mat[e] : 0];
  // FEEngine::extractNodalToElementField (fe_engine_inline_impl.hh:104-143)
  double ue[dn];
  double xe[RECOMPUTE ? dn : 1];
#pragma unroll
  for (int a = 0; a < n; ++a) {
    const int64_t node = A.conn[a * E + e];
#pragma unroll
    for (int i = 0; i < d; ++i) {
      ue[i + a * d] = A.u[node * d + i];
      if constexpr (RECOMPUTE) xe[i + a * d] = A.X[node * d + i];
    }
  }
  double fe[dn];
  element_forces_fp64<ET, KIND, RECOMPUTE>(A, e, m, ue, xe, fe);
  store_slots<d, n>(A, e, fe);
}

// ===========================================================================
// element kernel, FP32 element math (mixed precision)
//   RECOMPUTE = false: B and det(J) w read from the FP32 tables;
//   RECOMPUTE = true:  both rebuilt per point from the element's coordinates
//     (differences to node 0 formed in FP64, then rounded), which replaces the
//     E*Q*(d*n+1) table reads by one more gather of the (cache-resident) nodes.
// ===========================================================================
// Element forces of the mixed path from the element's displacement and
// coordinate differences to node 0 (du, dX; dX unused when the FP32 tables are
// read).  Shared by the gather kernel and the LDS-patch kernel.
// In-place 8-point Walsh-Hadamard butterflies over the corner index c (bit 0:
// xi sign, bit 1: eta, bit 2: zeta; bit set = +1).  Analysis: v[w] becomes
// sum_c prod_{k in w} s_k(c) v[c].  Synthesis: v[c] becomes
// sum_w prod_{k in w} s_k(c) v[w].
template <bool SYNTH>
__device__ __forceinline__ void walsh8(float (&v)[8][3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if (c & (1 << k)) continue;
      const int h = c | (1 << k);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const float lo = v[c][i], hi = v[h][i];
        v[c][i] = SYNTH ? lo - hi : lo + hi;
        v[h][i] = SYNTH ? lo + hi : hi - lo;
      }
    }
}

// Gauss points of the 8-point hexahedron rule in the reference order (xi
// fastest; gauss_integration_tmpl.hh:131-160): xi, eta, zeta = +-1/sqrt(3),
// then eta zeta, xi zeta, xi eta (FP32 products), padded to 8 floats.
constexpr float HEX8_G = 0.577350269189625764509f;
#define AKX_HEX8_POINT(q)                                                        \
  {((q)&1 ? HEX8_G : -HEX8_G), ((q)&2 ? HEX8_G : -HEX8_G), ((q)&4 ? HEX8_G : -HEX8_G), \
   ((q)&2 ? HEX8_G : -HEX8_G) * ((q)&4 ? HEX8_G : -HEX8_G),                      \
   ((q)&1 ? HEX8_G : -HEX8_G) * ((q)&4 ? HEX8_G : -HEX8_G),                      \
   ((q)&1 ? HEX8_G : -HEX8_G) * ((q)&2 ? HEX8_G : -HEX8_G), 0.f, 0.f}
__constant__ float k_hex8_points[8][8] = {AKX_HEX8_POINT(0), AKX_HEX8_POINT(1), AKX_HEX8_POINT(2),
                                          AKX_HEX8_POINT(3), AKX_HEX8_POINT(4), AKX_HEX8_POINT(5),
                                          AKX_HEX8_POINT(6), AKX_HEX8_POINT(7)};
#undef AKX_HEX8_POINT

// Mixed-precision forces of one 8-node hexahedron with the geometry rebuilt
// at every point, in the Walsh basis of the trilinear shape functions.  With
// node a at corner (s1, s2, s3) (element_class_hexahedron_8_inline_impl.hh
// node order), 8 dN_a/dxi = s1 (1 + s2 eta)(1 + s3 zeta) and cyclically, so
//   8 dx/dxi   = h1 + eta h3 + zeta h5 + eta zeta h7
//   8 dx/deta  = h2 + xi  h3 + zeta h6 + xi zeta h7
//   8 dx/dzeta = h4 + xi  h5 + eta h6  + xi eta h7
// with h_w the Walsh coefficients of the node values (same for u).  For
// J' = 8 J:  grad u = (H' adj J'^T) / det J',  and with Qm = P adj J' (w / 512)
// the element forces are f_a = sum_l Qm(:, l) 8 dN_a/dxi_l, accumulated as
// Walsh coefficients over the points and synthesised once.  No B matrix and
// no dN/dxi table: 2-3x fewer instructions than the B form.  Points in the
// reference order (xi fastest; gauss_integration_tmpl.hh:131-160), weight 1.
template <int KIND>
__device__ __forceinline__ void hex8_mixed_forces(const ElemArgs& A, int64_t e, const MatParams& m,
                                                  const float* du, const float* dX, float* fe) {
#pragma clang fp contract(fast)
  constexpr int corner[8] = {0, 1, 3, 2, 4, 5, 7, 6};  // node a -> corner bits
  float hx[8][3], hu[8][3], F[8][3];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      hx[corner[a]][i] = dX[i + a * 3];
      hu[corner[a]][i] = du[i + a * 3];
      F[a][i] = 0.f;
    }
  walsh8<false>(hx);
  walsh8<false>(hu);
#pragma unroll 1
  for (int q = 0; q < 8; ++q) {
    // the point's coordinates and their products, uniform over the wave:
    // scalar loads from the constant table instead of vector selects
    const float* c = k_hex8_points[q];
    const float xi = c[0], eta = c[1], zeta = c[2], ez = c[3], xz = c[4], xe = c[5];
    // J'[l + 3 j] = 8 dx_j / dxi_l,  H'[i + 3 l] = 8 du_i / dxi_l
    float J[9], H[9];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      J[0 + 3 * j] = hx[1][j] + eta * hx[3][j] + zeta * hx[5][j] + ez * hx[7][j];
      J[1 + 3 * j] = hx[2][j] + xi * hx[3][j] + zeta * hx[6][j] + xz * hx[7][j];
      J[2 + 3 * j] = hx[4][j] + xi * hx[5][j] + eta * hx[6][j] + xe * hx[7][j];
      H[j + 0] = hu[1][j] + eta * hu[3][j] + zeta * hu[5][j] + ez * hu[7][j];
      H[j + 3] = hu[2][j] + xi * hu[3][j] + zeta * hu[6][j] + xz * hu[7][j];
      H[j + 6] = hu[4][j] + xi * hu[5][j] + eta * hu[6][j] + xe * hu[7][j];
    }
    float Ad[9];
    Ad[0] = J[4] * J[8] - J[7] * J[5];
    Ad[1] = J[2] * J[7] - J[8] * J[1];
    Ad[2] = J[1] * J[5] - J[4] * J[2];
    Ad[3] = J[5] * J[6] - J[8] * J[3];
    Ad[4] = J[0] * J[8] - J[6] * J[2];
    Ad[5] = J[2] * J[3] - J[5] * J[0];
    Ad[6] = J[3] * J[7] - J[6] * J[4];
    Ad[7] = J[1] * J[6] - J[7] * J[0];
    Ad[8] = J[0] * J[4] - J[3] * J[1];
    const float det = J[0] * Ad[0] + J[3] * Ad[1] + J[6] * Ad[2];
    const float idet = __builtin_amdgcn_rcpf(det);
    // G(i, j) = sum_l H'(i, l) Ji'(j, l),  Ji' = Ad / det
    float G[9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int i = 0; i < 3; ++i)
        G[i + 3 * j] = idet * (H[i] * Ad[j] + H[i + 3] * Ad[j + 3] + H[i + 6] * Ad[j + 6]);
    constexpr float wq = 1.f / 512.f;  // Gauss weight 1, J' = 8 J
    float P[9];
    if constexpr (KIND == ELASTIC) {
      const float tr = G[0] + G[4] + G[8];
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          float s = m.mu_f * (G[i + 3 * j] + G[j + 3 * i]);
          if (i == j) s += m.lambda_f * tr;
          P[i + 3 * j] = s * wq;
        }
    } else {
      float scale = wq, W;
      float S6[6];
      if constexpr (KIND == NH_DAMAGE) {
        // Marigo update (material_marigo_inline_impl.hh:67-76) with Y = W
        neohookean_stress_sym3_f32(m.lambda_f, m.mu_f, G, 1.f, S6, &W);
        const int64_t eq = e * 8 + q;
        double dam = A.damage[eq];
        const double Fd = double(W) - A.Yd[eq] - m.Sd * dam;
        if (Fd > 0) dam = (double(W) - A.Yd[eq]) / m.Sd;
        dam = fmin(dam, 1.);
        A.damage[eq] = dam;
        scale *= float(1. - dam);
#pragma unroll
        for (int k = 0; k < 6; ++k) S6[k] *= scale;
      } else {
        neohookean_stress_sym3_f32(m.lambda_f, m.mu_f, G, scale, S6, nullptr);
      }
      // P = F S = S + G S, S from its 6 entries
      const float S[9] = {S6[0], S6[5], S6[4], S6[5], S6[1], S6[3], S6[4], S6[3], S6[2]};
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i)
          P[i + 3 * j] = S[i + 3 * j] + G[i] * S[3 * j] + G[i + 3] * S[3 * j + 1] +
                         G[i + 6] * S[3 * j + 2];
    }
    // Qm(i, l) = sum_j P(i, j) Ad(j, l), then the Walsh coefficients of the
    // forces: 8 dN_a/dxi_l expanded in prod s_k
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float q0 = P[i] * Ad[0] + P[i + 3] * Ad[1] + P[i + 6] * Ad[2];
      const float q1 = P[i] * Ad[3] + P[i + 3] * Ad[4] + P[i + 6] * Ad[5];
      const float q2 = P[i] * Ad[6] + P[i + 3] * Ad[7] + P[i + 6] * Ad[8];
      // one fused multiply-add per term (a sum of products added at once
      // would be formed first and cost one more instruction)
      F[1][i] += q0;
      F[2][i] += q1;
      F[4][i] += q2;
      F[3][i] += eta * q0;
      F[3][i] += xi * q1;
      F[5][i] += zeta * q0;
      F[5][i] += xi * q2;
      F[6][i] += zeta * q1;
      F[6][i] += eta * q2;
      F[7][i] += ez * q0;
      F[7][i] += xz * q1;
      F[7][i] += xe * q2;
    }
  }
  walsh8<true>(F);
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) fe[i + a * 3] = F[corner[a]][i];
}

// Kind of the mixed element kernels for neo-Hookean + damage through the
// constitutive surrogate (tetrahedron_4): the point's neo-Hookean stress and
// energy come from the network's outputs (sur_y) when its strain is inside
// the envelope (flag in sur_x), from the plain FP32 law otherwise.
constexpr int NH_DAMAGE_SUR = 3;

// Green-Lagrange strain E = (G + G^T + G^T G) / 2, Voigt xx yy zz yz xz xy
__device__ __forceinline__ void green_voigt(const float* G, float* Ev) {
#pragma clang fp contract(fast)
  const int vi[6] = {0, 1, 2, 1, 0, 0}, vj[6] = {0, 1, 2, 2, 2, 1};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int a = vi[k], b = vj[k];
    float gtg = 0.f;
#pragma unroll
    for (int l = 0; l < 3; ++l) gtg += G[l + a * 3] * G[l + b * 3];
    Ev[k] = 0.5f * (G[a + b * 3] + G[b + a * 3] + gtg);
  }
}

template <int ET, int KIND, bool RECOMPUTE>
__device__ __forceinline__ void mixed_element_forces(const ElemArgs& A, int64_t e,
                                                     const MatParams& m, const float* du,
                                                     const float* dX, float* fe) {
#pragma clang fp contract(fast)
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, Q = T::Q, dn = d * n;
  constexpr bool simplex = (ET == TRI3 || ET == TET4);
  const int64_t E = A.E;
  if constexpr (ET == HEX8 && RECOMPUTE) {
    hex8_mixed_forces<KIND>(A, e, m, du, dX, fe);
    return;
  }
  const float* __restrict__ Bt = static_cast<const float*>(A.B);
  const float* __restrict__ Jt = static_cast<const float*>(A.J);
#pragma unroll
  for (int k = 0; k < dn; ++k) fe[k] = 0.f;
  if constexpr (simplex && RECOMPUTE && KIND == ELASTIC) {
    // Linear simplex, one point, linear law: with A = adj(J),
    // B_a = A(:, a-1) / det for a > 0, grad u = G' / det where
    // G' = sum_a du_a A(:, a-1)^T, and det(J) w cancels one 1/det:
    //   f_a = (w / det) sigma(G') A(:, a-1),   f_0 = -(f_1 + ... + f_{n-1}).
    float Ad[d * d], det;
    if constexpr (d == 2) {
      // J(i,j) = dX[j + (i+1) d]
      const float j00 = dX[2], j10 = dX[4], j01 = dX[3], j11 = dX[5];
      Ad[0] = j11; Ad[1] = -j10; Ad[2] = -j01; Ad[3] = j00;
      det = j00 * Ad[0] + j01 * Ad[1];
    } else {
      float J[9];
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) J[i + j * 3] = dX[j + (i + 1) * 3];
      Ad[0] = J[4] * J[8] - J[7] * J[5];
      Ad[1] = J[2] * J[7] - J[8] * J[1];
      Ad[2] = J[1] * J[5] - J[4] * J[2];
      Ad[3] = J[5] * J[6] - J[8] * J[3];
      Ad[4] = J[0] * J[8] - J[6] * J[2];
      Ad[5] = J[2] * J[3] - J[5] * J[0];
      Ad[6] = J[3] * J[7] - J[6] * J[4];
      Ad[7] = J[1] * J[6] - J[7] * J[0];
      Ad[8] = J[0] * J[4] - J[3] * J[1];
      det = J[0] * Ad[0] + J[3] * Ad[1] + J[6] * Ad[2];
    }
    const float s = A.tabf.w[0] * __builtin_amdgcn_rcpf(det);
    float G[d * d];
#pragma unroll
    for (int j = 0; j < d; ++j)
#pragma unroll
      for (int i = 0; i < d; ++i) {
        float g = 0.f;
#pragma unroll
        for (int a = 1; a < n; ++a) g += du[i + a * d] * Ad[j + (a - 1) * d];
        G[i + j * d] = g;
      }
    float tr = 0.f;
#pragma unroll
    for (int i = 0; i < d; ++i) tr += G[i + i * d];
    float P[d * d];
#pragma unroll
    for (int j = 0; j < d; ++j)
#pragma unroll
      for (int i = 0; i <= j; ++i) {
        float v = m.mu_f * (G[i + j * d] + G[j + i * d]);
        if (i == j) v += m.lambda_f * tr;
        P[i + j * d] = P[j + i * d] = v * s;
      }
#pragma unroll
    for (int a = 1; a < n; ++a)
#pragma unroll
      for (int i = 0; i < d; ++i) {
        float f = 0.f;
#pragma unroll
        for (int j = 0; j < d; ++j) f += P[i + j * d] * Ad[j + (a - 1) * d];
        fe[i + a * d] = f;
        fe[i] -= f;
      }
    return;
  }
  // With the damage law the quadrature loop stays rolled (unrolled, its
  // 8-point hexahedron needed 448 VGPRs + 192 AGPRs: one wave per SIMD) and
  // non-simplex elements rebuilding the geometry never form B = Ji dN/dxi:
  //   G(i,j) = sum_l H(i,l) Ji(j,l),    H(i,l) = sum_a du_a,i dN_a/dxi_l;
  //   f_a,i += sum_l Qm(i,l) dN_a/dxi_l, Qm(i,l) = sum_j P(i,j) Ji(j,l),
  // which holds fewer values live (18 spilled values instead of 39 under the
  // 128-VGPR cap).  For the other laws the B form is the faster one (320 vs
  // 335 us on the 128^3 neo-Hookean hexahedron mesh), though it needs more
  // multiply-adds.
  constexpr bool damage = KIND == NH_DAMAGE || KIND == NH_DAMAGE_SUR;
  static_assert(KIND != NH_DAMAGE_SUR || ET == TET4, "the surrogate is for tetrahedron_4");
  constexpr bool on_the_fly = RECOMPUTE && !simplex && damage;
  constexpr int QUNROLL = damage ? 1 : Q;
#pragma unroll QUNROLL
  for (int q = 0; q < Q; ++q) {
    float Bq[on_the_fly ? 1 : dn];
    float Jiq[on_the_fly ? d * d : 1];
    float jq;
    if constexpr (!RECOMPUTE) {
#pragma unroll
      for (int k = 0; k < dn; ++k) Bq[k] = Bt[(int64_t(q) * dn + k) * E + e];
      jq = Jt[int64_t(q) * E + e];
    } else {
      // J(i,j) = sum_a dN_a/dxi_i x_a,j (column-major), with x_0 = 0
      float J[d * d];
#pragma unroll
      for (int j = 0; j < d; ++j)
#pragma unroll
        for (int i = 0; i < d; ++i) {
          if constexpr (simplex) {
            J[i + j * d] = dX[j + (i + 1) * d];
          } else {
            float s = 0.f;
#pragma unroll
            for (int a = 1; a < n; ++a) s += A.tabf.dnds[q][i + a * d] * dX[j + a * d];
            J[i + j * d] = s;
          }
        }
      // inverse through the adjugate: det from its first column, 1/det by
      // the hardware reciprocal (about 1 ulp; this is the FP32 path)
      float Ad[d * d], det;
      if constexpr (d == 2) {
        Ad[0] = J[3]; Ad[1] = -J[1]; Ad[2] = -J[2]; Ad[3] = J[0];
        det = J[0] * Ad[0] + J[2] * Ad[1];
      } else {
        Ad[0] = J[4] * J[8] - J[7] * J[5];
        Ad[1] = J[2] * J[7] - J[8] * J[1];
        Ad[2] = J[1] * J[5] - J[4] * J[2];
        Ad[3] = J[5] * J[6] - J[8] * J[3];
        Ad[4] = J[0] * J[8] - J[6] * J[2];
        Ad[5] = J[2] * J[3] - J[5] * J[0];
        Ad[6] = J[3] * J[7] - J[6] * J[4];
        Ad[7] = J[1] * J[6] - J[7] * J[0];
        Ad[8] = J[0] * J[4] - J[3] * J[1];
        det = J[0] * Ad[0] + J[3] * Ad[1] + J[6] * Ad[2];
      }
      const float idet = __builtin_amdgcn_rcpf(det);
      float Ji[d * d];
#pragma unroll
      for (int k = 0; k < d * d; ++k) Ji[k] = Ad[k] * idet;
      if constexpr (on_the_fly) {
#pragma unroll
        for (int k = 0; k < d * d; ++k) Jiq[k] = Ji[k];
      } else {
        // B(i,a) = sum_l Ji(i,l) dN_a/dxi_l
#pragma unroll
        for (int a = 0; a < n; ++a)
#pragma unroll
          for (int i = 0; i < d; ++i) {
            float s = 0.f;
            if constexpr (simplex) {
              if (a == 0) {
#pragma unroll
                for (int l = 0; l < d; ++l) s -= Ji[i + l * d];
              } else {
                s = Ji[i + (a - 1) * d];
              }
            } else {
#pragma unroll
              for (int l = 0; l < d; ++l) s += Ji[i + l * d] * A.tabf.dnds[q][l + a * d];
            }
            Bq[i + a * d] = s;
          }
      }
      jq = det * A.tabf.w[q];
    }
    // column a of B: B(i,a) = sum_l Ji(i,l) dN_a/dxi_l, or the stored one
    float G[d * d];
#pragma unroll
    for (int k = 0; k < d * d; ++k) G[k] = 0.f;
    if constexpr (on_the_fly) {
      float H[d * d];
#pragma unroll
      for (int k = 0; k < d * d; ++k) H[k] = 0.f;
#pragma unroll
      for (int a = 1; a < n; ++a)
#pragma unroll
        for (int l = 0; l < d; ++l) {
          const float dn_al = A.tabf.dnds[q][l + a * d];
#pragma unroll
          for (int i = 0; i < d; ++i) H[i + l * d] += du[i + a * d] * dn_al;
        }
#pragma unroll
      for (int j = 0; j < d; ++j)
#pragma unroll
        for (int l = 0; l < d; ++l)
#pragma unroll
          for (int i = 0; i < d; ++i) G[i + j * d] += H[i + l * d] * Jiq[j + l * d];
    } else {
#pragma unroll
      for (int a = 1; a < n; ++a)
#pragma unroll
        for (int j = 0; j < d; ++j)
#pragma unroll
          for (int i = 0; i < d; ++i) G[i + j * d] += du[i + a * d] * Bq[j + a * d];
    }
    float P[d * d];  // first Piola-Kirchhoff (Cauchy at small strain) times det(J) w
    if constexpr (KIND == ELASTIC) {
      float tr = 0.f;
#pragma unroll
      for (int i = 0; i < d; ++i) tr += G[i + i * d];
#pragma unroll
      for (int i = 0; i < d; ++i)
#pragma unroll
        for (int j = 0; j < d; ++j) {
          float s = m.mu_f * (G[i + j * d] + G[j + i * d]);
          if (i == j) s += m.lambda_f * tr;
          P[i + j * d] = s * jq;
        }
    } else if constexpr (d == 3) {
      float S6[6], W;
      float scale = jq;
      if constexpr (KIND == NH_DAMAGE_SUR) {
        // the surrogate's point index is the element's (one point, no filter);
        // the network's outputs, and in their 8th float the envelope flag
        const float4* yv = reinterpret_cast<const float4*>(A.sur_y + e * 8);
        const float4 y0 = yv[0], y1 = yv[1];
        if (y1.w != 0.f) {
          // S = S_lin + sS y[0..5], W = W_lin + sW y[6] (Voigt xx yy zz yz xz xy)
          float Ev[6];
          green_voigt(G, Ev);
          const float y[7] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z};
          const float tr = Ev[0] + Ev[1] + Ev[2];
          const float ee = Ev[0] * Ev[0] + Ev[1] * Ev[1] + Ev[2] * Ev[2] +
                           2.f * (Ev[3] * Ev[3] + Ev[4] * Ev[4] + Ev[5] * Ev[5]);
#pragma unroll
          for (int k = 0; k < 6; ++k)
            S6[k] = 2.f * m.mu_f * Ev[k] + (k < 3 ? m.lambda_f * tr : 0.f) + A.sur_sS * y[k];
          W = 0.5f * m.lambda_f * tr * tr + m.mu_f * ee + A.sur_sW * y[6];
        } else {
          neohookean_stress_sym3_f32(m.lambda_f, m.mu_f, G, 1.f, S6, &W);
        }
      } else {
        neohookean_stress_sym3_f32(m.lambda_f, m.mu_f, G, 1.f, S6, KIND == NH_DAMAGE ? &W : nullptr);
      }
      if constexpr (damage) {
        const int64_t eq = e * Q + q;
        double dam = A.damage[eq];
        const double Fd = double(W) - A.Yd[eq] - m.Sd * dam;
        if (Fd > 0) dam = (double(W) - A.Yd[eq]) / m.Sd;
        dam = fmin(dam, 1.);
        A.damage[eq] = dam;
        scale *= float(1. - dam);
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) S6[k] *= scale;
      // P = F S = S + G S, S from its 6 entries
      const float S[9] = {S6[0], S6[5], S6[4], S6[5], S6[1], S6[3], S6[4], S6[3], S6[2]};
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i)
          P[i + 3 * j] = S[i + 3 * j] + G[i] * S[3 * j] + G[i + 3] * S[3 * j + 1] +
                         G[i + 6] * S[3 * j + 2];
    } else {
      float S[d * d];
      float W;
      neohookean_stress_f32<d>(m.lambda_f, m.mu_f, G, S, KIND == NH_DAMAGE ? &W : nullptr);
      float scale = jq;
      if constexpr (KIND == NH_DAMAGE) {
        const int64_t eq = e * Q + q;
        double dam = A.damage[eq];
        const double Fd = double(W) - A.Yd[eq] - m.Sd * dam;
        if (Fd > 0) dam = (double(W) - A.Yd[eq]) / m.Sd;
        dam = fmin(dam, 1.);
        A.damage[eq] = dam;
        scale *= float(1. - dam);
      }
      // P = F S = S + G S
#pragma unroll
      for (int i = 0; i < d; ++i)
#pragma unroll
        for (int j = 0; j < d; ++j) {
          float s = S[i + j * d];
#pragma unroll
          for (int l = 0; l < d; ++l) s += G[i + l * d] * S[l + j * d];
          P[i + j * d] = s * scale;
        }
    }
    // f_a,i += sum_j P_ij dN_a/dX_j
    if constexpr (on_the_fly) {
      float Qm[d * d];
#pragma unroll
      for (int l = 0; l < d; ++l)
#pragma unroll
        for (int i = 0; i < d; ++i) {
          float s = 0.f;
#pragma unroll
          for (int j = 0; j < d; ++j) s += P[i + j * d] * Jiq[j + l * d];
          Qm[i + l * d] = s;
        }
#pragma unroll
      for (int a = 0; a < n; ++a)
#pragma unroll
        for (int i = 0; i < d; ++i) {
          float s = fe[i + a * d];
#pragma unroll
          for (int l = 0; l < d; ++l) s += Qm[i + l * d] * A.tabf.dnds[q][l + a * d];
          fe[i + a * d] = s;
        }
    } else {
#pragma unroll
      for (int a = 0; a < n; ++a)
#pragma unroll
        for (int i = 0; i < d; ++i) {
          float s = fe[i + a * d];
#pragma unroll
          for (int j = 0; j < d; ++j) s += P[i + j * d] * Bq[j + a * d];
          fe[i + a * d] = s;
        }
    }
  }
}

template <int ET, int KIND, bool RECOMPUTE>
__global__ void __launch_bounds__(256) k_element_mixed(ElemArgs A) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, dn = d * n;
  const int64_t t = logical_block(A.xcd_remap) * blockDim.x + threadIdx.x;
  if (t >= A.count) return;
  const int64_t E = A.E;
  const int64_t e = A.filter ? int64_t(A.filter[t]) : t;
  const MatParams m = A.mats[A.elem_mat ? A.elem_mat[e] : 0];
  // displacement differences to node 0, formed in FP64 then rounded: sum_a B_a = 0
  // (partition of unity), so grad u = sum_{a>0} (u_a - u_0) B_a^T exactly.
  int64_t nodes[n];
#pragma unroll
  for (int a = 0; a < n; ++a) nodes[a] = A.conn[a * E + e];
  double u0[d];
#pragma unroll
  for (int i = 0; i < d; ++i) u0[i] = A.u[nodes[0] * d + i];
  float du[dn];
#pragma unroll
  for (int i = 0; i < d; ++i) du[i] = 0.f;
#pragma unroll
  for (int a = 1; a < n; ++a)
#pragma unroll
    for (int i = 0; i < d; ++i) du[i + a * d] = float(A.u[nodes[a] * d + i] - u0[i]);
  float dX[RECOMPUTE ? dn : 1];
  if constexpr (RECOMPUTE) {
    double x0[d];
#pragma unroll
    for (int i = 0; i < d; ++i) x0[i] = A.X[nodes[0] * d + i];
#pragma unroll
    for (int i = 0; i < d; ++i) dX[i] = 0.f;
#pragma unroll
    for (int a = 1; a < n; ++a)
#pragma unroll
      for (int i = 0; i < d; ++i) dX[i + a * d] = float(A.X[nodes[a] * d + i] - x0[i]);
  }
  float fe[dn];
  mixed_element_forces<ET, KIND, RECOMPUTE>(A, e, m, du, dX, fe);
  store_slots<d, n>(A, e, fe);
}

// ===========================================================================
// element step, mixed precision, neo-Hookean + damage through the
// constitutive surrogate (tetrahedron_4), in three launches:
//   k_sur_strain  one thread per element: Green strain of the point, network
//                 inputs and the envelope flag (|E_k| <= eps_max, damage in
//                 [0, 1]) to sur_x;
//   network       persistent MFMA blocks, sur_x -> sur_y (akx_surrogate.hip);
//   forces        S_NH, W_NH from the network (or the plain FP32 law outside
//                 the envelope, counted by k_sur_strain), Marigo update,
//                 B^T (F S) det(J) w: with the LDS patches, the patch element
//                 kernel of kind NH_DAMAGE_SUR (per-patch partial sums or
//                 contiguous runs, as for the plain law); without them
//                 k_sur_forces, one thread per element, into the force slots.
// The strain gathers run at full occupancy; the network kernel streams its
// inputs.
// ===========================================================================
struct SurPoint {
  float Bq[12], jq, G[9];
};

// Geometry and displacement gradient of tetrahedron_4 element e (same
// arithmetic as k_element_mixed<TET4, KIND, true>).
__device__ __forceinline__ void sur_point(const ElemArgs& A, int64_t e, SurPoint& P) {
#pragma clang fp contract(fast)
  constexpr int d = 3, n = 4, dn = 12;
  const int64_t E = A.E;
  int64_t nodes[n];
#pragma unroll
  for (int a = 0; a < n; ++a) nodes[a] = A.conn[a * E + e];
  float du[dn], dX[dn];
  double u0[d], x0[d];
#pragma unroll
  for (int i = 0; i < d; ++i) {
    u0[i] = A.u[nodes[0] * d + i];
    x0[i] = A.X[nodes[0] * d + i];
    du[i] = 0.f;
    dX[i] = 0.f;
  }
#pragma unroll
  for (int a = 1; a < n; ++a)
#pragma unroll
    for (int i = 0; i < d; ++i) {
      du[i + a * d] = float(A.u[nodes[a] * d + i] - u0[i]);
      dX[i + a * d] = float(A.X[nodes[a] * d + i] - x0[i]);
    }
  float Jm[9];
#pragma unroll
  for (int jj = 0; jj < d; ++jj)
#pragma unroll
    for (int i = 0; i < d; ++i) Jm[i + jj * d] = dX[jj + (i + 1) * d];
  const float det = det3(Jm);
  const float idet = 1.f / det;
  float Ji[9];
  Ji[0] = (Jm[4] * Jm[8] - Jm[7] * Jm[5]) * idet;
  Ji[1] = (Jm[2] * Jm[7] - Jm[8] * Jm[1]) * idet;
  Ji[2] = (Jm[1] * Jm[5] - Jm[4] * Jm[2]) * idet;
  Ji[3] = (Jm[5] * Jm[6] - Jm[8] * Jm[3]) * idet;
  Ji[4] = (Jm[0] * Jm[8] - Jm[6] * Jm[2]) * idet;
  Ji[5] = (Jm[2] * Jm[3] - Jm[5] * Jm[0]) * idet;
  Ji[6] = (Jm[3] * Jm[7] - Jm[6] * Jm[4]) * idet;
  Ji[7] = (Jm[1] * Jm[6] - Jm[7] * Jm[0]) * idet;
  Ji[8] = (Jm[0] * Jm[4] - Jm[3] * Jm[1]) * idet;
#pragma unroll
  for (int i = 0; i < d; ++i) {
    float sm = 0.f;
#pragma unroll
    for (int l = 0; l < d; ++l) sm -= Ji[i + l * d];
    P.Bq[i] = sm;
#pragma unroll
    for (int a = 1; a < n; ++a) P.Bq[i + a * d] = Ji[i + (a - 1) * d];
  }
  P.jq = det * A.tabf.w[0];
#pragma unroll
  for (int jj = 0; jj < d; ++jj)
#pragma unroll
    for (int i = 0; i < d; ++i) {
      float sm = 0.f;
#pragma unroll
      for (int a = 1; a < n; ++a) sm += du[i + a * d] * P.Bq[jj + a * d];
      P.G[i + jj * d] = sm;
    }
}

// Network inputs of point t (scaled Green strain, damage) and its envelope
// flag to sur_x; returns the flag.
__device__ __forceinline__ bool sur_strain_point(const ElemArgs& A, int64_t t) {
  const int64_t e = A.filter ? int64_t(A.filter[t]) : t;
  SurPoint P;
  sur_point(A, e, P);
  float Ev[6];
  green_voigt(P.G, Ev);
  const double dam = A.damage[e];
  bool inside = dam >= 0. && dam <= 1.;
  const float ieps = 1.f / A.sur_eps;
  float x[8];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    inside = inside && fabsf(Ev[k]) <= A.sur_eps;
    x[k] = Ev[k] * ieps;
  }
  x[6] = float(dam);
  x[7] = inside ? 1.f : 0.f;
  float4* o = reinterpret_cast<float4*>(A.sur_x + t * 8);
  o[0] = make_float4(x[0], x[1], x[2], x[3]);
  o[1] = make_float4(x[4], x[5], x[6], x[7]);
  return inside;
}

__global__ void __launch_bounds__(256) k_sur_strain(ElemArgs A) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool live = t < A.count;
  // points outside the envelope take the plain law: count them, one atomic
  // per wave
  bool outside = false;
  if (live) outside = !sur_strain_point(A, t);
  const unsigned long long wave_fb = __popcll(__ballot(outside));
  if ((threadIdx.x & 63) == 0 && wave_fb) atomicAdd(A.sur_fallbacks, wave_fb);
}


__global__ void __launch_bounds__(256) k_sur_forces(ElemArgs A) {
#pragma clang fp contract(fast)
  constexpr int d = 3, n = 4, dn = 12;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < A.count) {
    const int64_t e = A.filter ? int64_t(A.filter[t]) : t;
    const MatParams m = A.mats[A.elem_mat ? A.elem_mat[e] : 0];
    SurPoint P;
    sur_point(A, e, P);
    const float* G = P.G;
    const bool inside = A.sur_y[t * 8 + 7] != 0.f;  // the flag, carried by the network
    const int vi[6] = {0, 1, 2, 1, 0, 0}, vj[6] = {0, 1, 2, 2, 2, 1};
    float S[9], W;
    if (inside) {
      float Ev[6];
      green_voigt(G, Ev);
      const float4* yv = reinterpret_cast<const float4*>(A.sur_y + t * 8);
      const float4 y0 = yv[0], y1 = yv[1];
      const float y[7] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z};
      const float tr = Ev[0] + Ev[1] + Ev[2];
      const float ee = Ev[0] * Ev[0] + Ev[1] * Ev[1] + Ev[2] * Ev[2] +
                       2.f * (Ev[3] * Ev[3] + Ev[4] * Ev[4] + Ev[5] * Ev[5]);
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const float sv = 2.f * m.mu_f * Ev[k] + (k < 3 ? m.lambda_f * tr : 0.f) + A.sur_sS * y[k];
        S[vi[k] + vj[k] * d] = sv;
        S[vj[k] + vi[k] * d] = sv;
      }
      W = 0.5f * m.lambda_f * tr * tr + m.mu_f * ee + A.sur_sW * y[6];
    } else {
      neohookean_stress_f32<d>(m.lambda_f, m.mu_f, G, S, &W);
    }
    // Marigo update (material_marigo_inline_impl.hh:67-76) with Y = W
    double dam = A.damage[e];
    const double Fd = double(W) - A.Yd[e] - m.Sd * dam;
    if (Fd > 0) dam = (double(W) - A.Yd[e]) / m.Sd;
    dam = fmin(dam, 1.);
    A.damage[e] = dam;
    const float scale = P.jq * float(1. - dam);
    float Pk[9];
#pragma unroll
    for (int i = 0; i < d; ++i)
#pragma unroll
      for (int jj = 0; jj < d; ++jj) {
        float sm = S[i + jj * d];
#pragma unroll
        for (int l = 0; l < d; ++l) sm += G[i + l * d] * S[l + jj * d];
        Pk[i + jj * d] = sm * scale;
      }
    float fe[dn];
#pragma unroll
    for (int a = 0; a < n; ++a)
#pragma unroll
      for (int i = 0; i < d; ++i) {
        float sm = 0.f;
#pragma unroll
        for (int jj = 0; jj < d; ++jj) sm += Pk[i + jj * d] * P.Bq[jj + a * d];
        fe[i + a * d] = sm;
      }
    store_slots<d, n>(A, e, fe);
  }
}

// Kind of the patch element kernel that only writes the surrogate's inputs:
// the Green strain of each tetrahedron_4 point from the patch's staged node
// differences (no forces).
constexpr int SUR_STRAIN = 4;

// Network inputs of element e (scaled Green strain, damage) and its envelope
// flag to sur_x, from its node differences to node 0 (du, dX, FP32); the
// gradient with the arithmetic of mixed_element_forces<TET4, ..., true>.
// Returns the flag.
__device__ __forceinline__ bool sur_strain_from_differences(const ElemArgs& A, int64_t e,
                                                            const float* du, const float* dX) {
#pragma clang fp contract(fast)
  constexpr int d = 3, n = 4;
  float J[9];
#pragma unroll
  for (int j = 0; j < d; ++j)
#pragma unroll
    for (int i = 0; i < d; ++i) J[i + j * d] = dX[j + (i + 1) * d];
  float Ad[9];
  Ad[0] = J[4] * J[8] - J[7] * J[5];
  Ad[1] = J[2] * J[7] - J[8] * J[1];
  Ad[2] = J[1] * J[5] - J[4] * J[2];
  Ad[3] = J[5] * J[6] - J[8] * J[3];
  Ad[4] = J[0] * J[8] - J[6] * J[2];
  Ad[5] = J[2] * J[3] - J[5] * J[0];
  Ad[6] = J[3] * J[7] - J[6] * J[4];
  Ad[7] = J[1] * J[6] - J[7] * J[0];
  Ad[8] = J[0] * J[4] - J[3] * J[1];
  const float det = J[0] * Ad[0] + J[3] * Ad[1] + J[6] * Ad[2];
  const float idet = __builtin_amdgcn_rcpf(det);
  float G[9];
#pragma unroll
  for (int j = 0; j < d; ++j)
#pragma unroll
    for (int i = 0; i < d; ++i) {
      float g = 0.f;
#pragma unroll
      for (int a = 1; a < n; ++a) g += du[i + a * d] * (Ad[j + (a - 1) * d] * idet);
      G[i + j * d] = g;
    }
  float Ev[6];
  green_voigt(G, Ev);
  const double dam = A.damage[e];
  bool inside = dam >= 0. && dam <= 1.;
  const float ieps = 1.f / A.sur_eps;
  float x[8];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    inside = inside && fabsf(Ev[k]) <= A.sur_eps;
    x[k] = Ev[k] * ieps;
  }
  x[6] = float(dam);
  x[7] = inside ? 1.f : 0.f;
  float4* o = reinterpret_cast<float4*>(A.sur_x + e * 8);
  o[0] = make_float4(x[0], x[1], x[2], x[3]);
  o[1] = make_float4(x[4], x[5], x[6], x[7]);
  return inside;
}

// ===========================================================================
// element kernel, mixed precision, LDS-staged element patches
//   One block = one patch of up to PATCH consecutive (Morton-ordered)
//   elements.  The block
//     1. loads the patch's distinct nodes (u and X in FP64, the start of the
//        node's run of slots) once, coalesced, into LDS;
//     2. computes every element from LDS through 16-bit patch-local node ids
//        (same arithmetic as k_element_mixed<..., true>);
//     3. stages the n*d forces of each element in LDS at the element's place
//        in the patch's per-node runs and writes the runs out contiguously,
//        or (A.partials) sums each node's run in FP64 and writes one partial
//        per patch node, which the node kernel gathers through a short
//        node -> patch-entry CSR.
//   Valid with the mixed-path CSR, which lists each node's slots by element
//   id: the patch's contributions to one node are then one contiguous run.
// ===========================================================================
// At least 4 waves per SIMD: the 8-point hexahedron with finite strain would
// otherwise take 151 VGPRs (3 waves); capped at 128 it spills a few values and
// runs 6 % faster.  The other instantiations are far below the cap.
template <int ET, int KIND>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_element_patch(ElemArgs A) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, dn = d * n;
  extern __shared__ __align__(16) unsigned char smem[];
  const int64_t b = logical_block(A.xcd_remap);
  const int64_t e0 = b * int64_t(blockDim.x);
  if (e0 >= A.count) return;  // uniform over the block
  const int ne = int(min<int64_t>(blockDim.x, A.count - e0));
  // per-patch lists are padded to a fixed stride S (the largest patch), so
  // that no load waits for the patch's offset; padding repeats a valid node
  const int S = A.patch_max_nodes;
  const int64_t k0 = b * int64_t(S);
  // LDS: per node (u - u_ref | X - X_ref) in FP32, run start (u32), run base
  // (u16, S + 1), then the forces (FP32, d per contribution)
  float* s_ux = reinterpret_cast<float*>(smem);
  uint32_t* s_rs = reinterpret_cast<uint32_t*>(s_ux + size_t(S) * 2 * d);
  uint16_t* s_rb = reinterpret_cast<uint16_t*>(s_rs + S);
  float* s_f = reinterpret_cast<float*>(smem + patch_force_offset(S, d));
  // the element's own indices do not depend on the staged nodes: issue their
  // loads first so that both latency chains overlap
  const int t = threadIdx.x;
  const int64_t e = e0 + min(t, ne - 1);
  int ln[n], rk[n];
  if (A.lcr) {
    // packed: patch-local node | rank << 8, one 8- or 16-byte load per element
    constexpr int np = n == 3 ? 4 : n;
    uint32_t w[np / 2];
    if constexpr (np == 4) {
      const uint2 v = reinterpret_cast<const uint2*>(A.lcr)[e];
      w[0] = v.x; w[1] = v.y;
    } else {
      static_assert(np == 8, "packed patch indices: 4 or 8 per element");
      const uint4 v = reinterpret_cast<const uint4*>(A.lcr)[e];
      w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    }
#pragma unroll
    for (int a = 0; a < n; ++a) {
      const uint32_t h = (w[a >> 1] >> (16 * (a & 1))) & 0xffffu;
      ln[a] = int(h & 0xffu);
      rk[a] = int(h >> 8);
    }
  } else {
#pragma unroll
    for (int a = 0; a < n; ++a) {
      ln[a] = A.lconn[a * A.E + e];
      rk[a] = A.lrank[a * A.E + e];
    }
  }
  const int mat = A.elem_mat ? A.elem_mat[e] : 0;
  const uint32_t m_nodes = A.pnode_cnt[b];
  // node values relative to the patch's first node, formed in FP64 and then
  // rounded: the element differences below stay accurate in FP32
  const int64_t ref = A.pnode_ids[k0];
  double uref[d], xref[d];
#pragma unroll
  for (int i = 0; i < d; ++i) {
    uref[i] = A.u[ref * d + i];
    xref[i] = A.X[ref * d + i];
  }
  for (int k = threadIdx.x; k < S; k += blockDim.x) {
    const int64_t node = A.pnode_ids[k0 + k];
    float v[2 * d];
#pragma unroll
    for (int i = 0; i < d; ++i) {
      v[i] = float(A.u[node * d + i] - uref[i]);
      v[d + i] = float(A.X[node * d + i] - xref[i]);
    }
    float2* dst = reinterpret_cast<float2*>(s_ux + size_t(k) * 2 * d);
#pragma unroll
    for (int j = 0; j < d; ++j) dst[j] = make_float2(v[2 * j], v[2 * j + 1]);
    if constexpr (KIND != SUR_STRAIN) {
      // run start (layout 2) or the partial's node-major position (layout 5)
      if (!A.partials) s_rs[k] = A.run_start[k0 + k];
      else if (A.part_pos) s_rs[k] = A.part_pos[k0 + k];
      s_rb[k] = A.run_base[b * int64_t(S + 1) + k];
    }
  }
  if (KIND != SUR_STRAIN && threadIdx.x == 0) s_rb[S] = A.run_base[b * int64_t(S + 1) + S];
  __syncthreads();
  bool outside = false;  // surrogate strain pass: point outside the envelope
  if (t < ne) {
    const MatParams m = A.mats[mat];
    auto node_ux = [&](int k, float* v) {
      const float2* src = reinterpret_cast<const float2*>(s_ux + size_t(k) * 2 * d);
#pragma unroll
      for (int j = 0; j < d; ++j) {
        const float2 w = src[j];
        v[2 * j] = w.x;
        v[2 * j + 1] = w.y;
      }
    };
    float v0[2 * d];
    node_ux(ln[0], v0);
    float du[dn], dX[dn];
#pragma unroll
    for (int i = 0; i < d; ++i) { du[i] = 0.f; dX[i] = 0.f; }
#pragma unroll
    for (int a = 1; a < n; ++a) {
      float v[2 * d];
      node_ux(ln[a], v);
#pragma unroll
      for (int i = 0; i < d; ++i) {
        du[i + a * d] = v[i] - v0[i];
        dX[i + a * d] = v[d + i] - v0[d + i];
      }
    }
    if constexpr (KIND == SUR_STRAIN) {
      outside = !sur_strain_from_differences(A, e, du, dX);
    } else {
      float fe[dn];
      mixed_element_forces<ET, KIND, true>(A, e, m, du, dX, fe);
#pragma unroll
      for (int a = 0; a < n; ++a) {
        const int q = s_rb[ln[a]] + rk[a];
#pragma unroll
        for (int i = 0; i < d; ++i) s_f[q * d + i] = fe[i + a * d];
      }
    }
  }
  if constexpr (KIND == SUR_STRAIN) {
    // the network's inputs are written: count the points outside the
    // envelope, one atomic per wave; no forces in this pass
    const unsigned long long wave_fb = __popcll(__ballot(outside));
    if ((threadIdx.x & 63) == 0 && wave_fb) atomicAdd(A.sur_fallbacks, wave_fb);
    return;
  }
  __syncthreads();
  if (A.partials) {
    // one sum per patch node, its run summed in element order in FP64 and
    // stored in FP64 (partials 1) or rounded once to FP32 (partials 2)
    for (uint32_t k = threadIdx.x; k < m_nodes; k += blockDim.x) {
      double f[d];
#pragma unroll
      for (int i = 0; i < d; ++i) f[i] = 0.;
      // chunks of RUNCHUNK entries: a chunk's LDS reads are issued together.
      // Full chunks add unmasked; the last, partial chunk reads into the
      // force area's RUNCHUNK - 1 entries of padding and its entries past
      // the run add +0, so the sum is the sequential one.  Patch-local nodes
      // are numbered by decreasing run length, so the lanes of a wavefront
      // run about the same number of chunks.
      const int q1 = s_rb[k + 1];
      int q = s_rb[k];
      for (; q + RUNCHUNK <= q1; q += RUNCHUNK) {
        float v[RUNCHUNK * d];
#pragma unroll
        for (int j = 0; j < RUNCHUNK * d; ++j) v[j] = s_f[q * d + j];
#pragma unroll
        for (int j = 0; j < RUNCHUNK; ++j)
#pragma unroll
          for (int i = 0; i < d; ++i) f[i] += double(v[j * d + i]);
      }
      if (q < q1) {
        float v[RUNCHUNK * d];
#pragma unroll
        for (int j = 0; j < RUNCHUNK * d; ++j) v[j] = s_f[q * d + j];
#pragma unroll
        for (int j = 0; j < RUNCHUNK; ++j)
#pragma unroll
          for (int i = 0; i < d; ++i) f[i] += double(q + j < q1 ? v[j * d + i] : 0.f);
      }
      const int64_t at = A.part_pos ? int64_t(s_rs[k]) : k0 + k;
      if (A.partials == 2) {
        float* __restrict__ part = static_cast<float*>(A.slots);
#pragma unroll
        for (int i = 0; i < d; ++i) part[at * d + i] = float(f[i]);
      } else {
        double* __restrict__ part = static_cast<double*>(A.slots);
#pragma unroll
        for (int i = 0; i < d; ++i) part[at * d + i] = f[i];
      }
    }
    return;
  }
  // contiguous runs: output q belongs to the node k with rb[k] <= q < rb[k+1]
  float* __restrict__ out = static_cast<float*>(A.slots);
  const int total = ne * n;
  for (int q = threadIdx.x; q < total; q += blockDim.x) {
    int lo = 0, hi = int(m_nodes);  // s_rb[lo] <= q < s_rb[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (s_rb[mid] <= q) lo = mid; else hi = mid;
    }
    const int64_t p = int64_t(s_rs[lo]) + (q - s_rb[lo]);
#pragma unroll
    for (int i = 0; i < d; ++i) out[p * d + i] = s_f[q * d + i];
  }
}

// ===========================================================================
// node kernel: force gather (CSR, assembly order) + residual + lumped solve +
// corrector of step n, then Dirichlet increments + predictor of step n+1.
//   residual: r = 0; r += 1 f_ext; r += 1 f_int; r += M (-1 a)
//     (TimeStepSolver::assembleResidual, solid_mechanics_model.cc:279-293,
//      dof_manager_default.cc:271-287)
//   solve:    delta = 1 (r / M)           (non_linear_solver_lumped.cc:85-98)
//   correct:  u += 0 delta; v += (dt/2) delta; a += delta   (newmark-beta.cc:176-210)
//   predict:  u += dt v + (dt/2)(dt a); v += dt a           (newmark-beta.cc:60-89)
// f_int[node] = sum over the node's slots of -fe (DOFManager::
// assembleElementalArrayLocalArray scale -1, dof_manager.cc:113-134).
// FP64 slots: plain sum in slot order (bit-identical to the sequential
// scatter).  FP32 slots: FP64 Neumaier-compensated sum.  The mixed path
// updates in the position form (node_position_form, NODE_POSITION).
// ===========================================================================
// One node's residual, lumped solve and corrector (NODE_CORRECT), with the
// assembled internal force fi (sum of -fe) and the node's own state u, v, a
// (d values each), external force fext (nullptr: none) and blocked flags b.
// Also writes f_int (NODE_STORE_FINT) and the increment (NODE_STORE_INC).
// Shared by the node kernel and the fused FP64 step.
template <int D, bool HAS_FEXT>
__device__ __forceinline__ void node_correct(const NodeArgs& A, int64_t node, double dt,
                                             const double* fi, const double* fext,
                                             const uint8_t* b, double* u, double* v, double* a) {
  const int mode = A.mode;
  const bool correct = mode & NODE_CORRECT;
  const double M = A.mass[node];
  const double dcoef = 0.5 * dt;            // beta dt
  const double ecoef = 0. * 0.5 * dt * dt;  // alpha beta dt^2
#pragma unroll
  for (int i = 0; i < D; ++i) {
    if (mode & NODE_STORE_FINT) A.fint_out[node * D + i] = fi[i];
    if (!correct) continue;
    double r = 0.;
    if constexpr (HAS_FEXT) r += fext[i];
    r += fi[i];
    double cache = 0.;
    cache += -1. * a[i];
    r += M * cache;
    if (!(b[i] & 1)) {
      const double delta = r / M;
      u[i] += ecoef * delta;
      v[i] += dcoef * delta;
      a[i] += delta;
    }
    if (mode & NODE_STORE_INC) {
      double inc = u[i];
      inc -= A.prev[node * D + i];
      A.inc[node * D + i] = inc;
    }
  }
}

// Dirichlet increments, previous displacement and predictor (NODE_PREDICT)
// of one node.
template <int D>
__device__ __forceinline__ void node_predict(const NodeArgs& A, int64_t node, double dt,
                                             const uint8_t* b, double* u, double* v, double* a) {
  const int mode = A.mode;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double ui = u[i];
    if (b[i] & 2) ui += A.bc_inc[i];
    if (mode & NODE_SAVE_PREV) A.prev[node * D + i] = ui;
    if (!(b[i] & 1)) {
      const double vi = v[i], acc = a[i];
      const double dt_a_n = dt * acc;
      ui += dt * vi + 0.5 * dt * dt_a_n;
      v[i] = vi + dt_a_n;
    }
    u[i] = ui;
  }
}

// The mixed path's node update in the position form (NODE_POSITION; see
// NodeMode): the first predictor (NODE_PREDICT alone) from (u, v, a), the
// steps between (NODE_CORRECT | NODE_PREDICT) from u and the previous
// displacement up, the last corrector (NODE_CORRECT alone) returning v and
// a.  Blocked degrees of freedom keep v and a and move by their Dirichlet
// increment only, as in node_correct / node_predict.  fi: assembled internal
// force (unused by the first predictor).
template <int D, bool HAS_FEXT>
__device__ __forceinline__ void node_position_form(const NodeArgs& A, int64_t node, double dt,
                                                   const double* fi, const double* fext,
                                                   const uint8_t* b, const double* up,
                                                   double* u, double* v, double* a) {
  const int mode = A.mode;
  const bool correct = mode & NODE_CORRECT, predict = mode & NODE_PREDICT;
  const bool mid = correct && predict;
  // between the steps: dt^2 / M; the last corrector: M; the first predictor
  // writes dt^2 / M for the steps after it
  const double M = correct && !mid ? A.mass[node] : 1.;
  const double k = mid ? A.pcoef[node] : 0.;
  if (predict && !correct) A.pcoef[node] = dt * dt / A.mass[node];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    if (mode & NODE_STORE_FINT) A.fint_out[node * D + i] = fi[i];
    const bool free_dof = !(b[i] & 1);
    double r = 0.;
    if (correct) {
      if constexpr (HAS_FEXT) r += fext[i];
      r += fi[i];
    }
    if (predict) {
      double ui = u[i];
      if (b[i] & 2) ui += A.bc_inc[i];
      if (mode & NODE_SAVE_PREV) A.prev[node * D + i] = ui;
      if (free_dof) {
        if (correct) ui += (u[i] - up[i]) + r * k;
        else ui += dt * (v[i] + 0.5 * dt * a[i]);
      }
      u[i] = ui;
    } else if (correct) {
      const double an = r / M;
      if (free_dof) {
        v[i] = (u[i] - up[i]) / dt + 0.5 * dt * an;
        a[i] = an;
      }
      if (mode & NODE_STORE_INC) {
        double inc = u[i];
        inc -= A.prev[node * D + i];
        A.inc[node * D + i] = inc;
      }
    }
  }
}

// SRC: what the node gathers -- 0 FP64 element slots, 1 FP32 element slots,
// 2 FP64 per-patch partial sums, 3 FP32 per-patch partial sums (FP32 values
// are summed in FP64 with Neumaier compensation).
template <int D, int SRC, bool HAS_FEXT>
__global__ void __launch_bounds__(256) k_node(NodeArgs A) {
  using TS = typename std::conditional<SRC == 1 || SRC == 3, float, double>::type;
  // The block's nodal state is staged through LDS: global loads and stores of
  // u, v, a, f_ext and the blocked flags (u and the previous displacement
  // between the position form's steps) are contiguous runs of doubles (fully
  // coalesced, each array written once), while each thread works on one node.
  // The runs move as 16-byte pairs of doubles and 4-byte groups of flags:
  // the block's first entry n0 D is a multiple of 64 (block sizes are
  // multiples of 64) and the arrays are device allocations.
  constexpr int NBMAX = 256;
  __shared__ __align__(16) double s_u[NBMAX * D], s_v[NBMAX * D], s_a[NBMAX * D],
      s_f[HAS_FEXT ? NBMAX * D : 2];
  __shared__ __align__(16) uint8_t s_b[NBMAX * D];
  const int nb = blockDim.x;
  const int64_t n0 = logical_block(A.xcd_remap) * nb;
  if (n0 >= A.N) return;  // uniform over the block
  const int cnt = int(min<int64_t>(nb, A.N - n0));
  const int64_t k0 = n0 * D;
  const int nk = cnt * D;
  const int mode = A.mode;
  const bool correct = mode & NODE_CORRECT, predict = mode & NODE_PREDICT;
  const bool position = mode & NODE_POSITION;
  // the position form between two steps touches neither v nor a: s_v holds
  // the previous displacement there
  const bool pos_mid = position && correct && predict;
  const TS* __restrict__ slots = static_cast<const TS*>(A.slots);
  const bool gather = mode & (NODE_CORRECT | NODE_GATHER);
  const bool staged = A.node_ordered && A.stage_lds > 0 && gather;
  const int ln = threadIdx.x;
  const int64_t node = n0 + ln;
  // Per-patch partial sums read from global memory (a few per node): the
  // node's range and its first group of 4 partials (through the entry list
  // when they are not node-ordered) are loaded before the block's state is
  // staged, so that this dependent chain overlaps the staging loads.  Element
  // slots (SRC 0, 1: long runs) are read after the staging in batches.
  const bool early = SRC >= 2 && gather && !staged && ln < cnt;
  uint32_t p0 = 0, p1 = 0;
  int c0 = 0;
  TS x0[4 * D];
  if (early) {
    p0 = A.csr_off[node];
    p1 = A.csr_off[node + 1];
    c0 = int(min<uint32_t>(4u, p1 - p0));
    if (c0 > 0) {
      uint32_t sj[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t q = j < c0 ? p0 + j : p0;
        sj[j] = A.node_ordered ? q : A.csr_slot[q];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < D; ++i) x0[j * D + i] = slots[int64_t(sj[j]) * D + i];
    }
  }
  {
    const int nk2 = nk >> 1, nk4 = nk >> 2;
    const double* vsrc = pos_mid ? A.u_prev : A.v;
    for (int t = threadIdx.x; t < nk2; t += nb) {
      reinterpret_cast<double2*>(s_u)[t] = reinterpret_cast<const double2*>(A.u + k0)[t];
      reinterpret_cast<double2*>(s_v)[t] = reinterpret_cast<const double2*>(vsrc + k0)[t];
      if (!pos_mid)
        reinterpret_cast<double2*>(s_a)[t] = reinterpret_cast<const double2*>(A.a + k0)[t];
      if constexpr (HAS_FEXT)
        reinterpret_cast<double2*>(s_f)[t] = reinterpret_cast<const double2*>(A.fext + k0)[t];
    }
    for (int t = threadIdx.x; t < nk4; t += nb)
      reinterpret_cast<uint32_t*>(s_b)[t] = reinterpret_cast<const uint32_t*>(A.blocked + k0)[t];
    // the last block's odd tails
    const int t2 = 2 * nk2 + int(threadIdx.x);
    if (t2 < nk) {
      s_u[t2] = A.u[k0 + t2];
      s_v[t2] = vsrc[k0 + t2];
      if (!pos_mid) s_a[t2] = A.a[k0 + t2];
      if constexpr (HAS_FEXT) s_f[t2] = A.fext[k0 + t2];
    }
    const int t4 = 4 * nk4 + int(threadIdx.x);
    if (t4 < nk) s_b[t4] = A.blocked[k0 + t4];
  }
  // Node-ordered slots with A.stage_lds > 0: the block's slots are one
  // contiguous range [off[n0], off[n0 + cnt]); copy it into LDS with coalesced
  // loads, then every thread reads its own run from there.
  extern __shared__ __align__(16) unsigned char dyn[];
  TS* st = reinterpret_cast<TS*>(dyn);
  uint32_t base = 0;
  if (staged) {
    const uint32_t q0 = A.csr_off[n0], q1 = A.csr_off[n0 + cnt];
    const int64_t len = int64_t(q1 - q0) * D;
    for (int64_t t = threadIdx.x; t < len; t += nb) st[t] = slots[int64_t(q0) * D + t];
    base = q0;
  }
  __syncthreads();
  if (ln < cnt) {
    // one time step per node when independent bodies advance together
    // the position form between two steps needs no time step (dt^2 / M)
    const double dt = pos_mid ? 0. : A.dt_node ? A.dt_node[node] : A.dt;
    if (gather) {
      double f[D], c[D];
#pragma unroll
      for (int i = 0; i < D; ++i) { f[i] = 0.; c[i] = 0.; }
      if (!early) {
        p0 = A.csr_off[node];
        p1 = A.csr_off[node + 1];
      }
      auto add = [&](int i, TS raw) {
        const double x = -1. * double(raw);
        if constexpr (sizeof(TS) == 8) {
          f[i] += x;
        } else {
          const double tsum = f[i] + x;
          c[i] += (fabs(f[i]) >= fabs(x)) ? (f[i] - tsum) + x : (x - tsum) + f[i];
          f[i] = tsum;
        }
      };
      // Batches of CHUNK entries: all loads of a batch are issued before the
      // (in-order) accumulation, so the gathers overlap instead of serialising.
      constexpr int CHUNK = 8;
      uint32_t p = p0;
      if (staged) {
        // the node's run, read from the block's LDS copy (starts at csr_off[n0])
        for (uint32_t q = p0 - base; q < p1 - base; ++q)
#pragma unroll
          for (int i = 0; i < D; ++i) add(i, st[q * D + i]);
        p = p1;
      } else {
        // the first (up to) 4 contributions, already loaded, are added first
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (j < c0)
#pragma unroll
            for (int i = 0; i < D; ++i) add(i, x0[j * D + i]);
        p += uint32_t(c0);
      }
      if (!staged && A.node_ordered) {
        // the node's contributions are contiguous: slots[p*D ...] for p in [p0, p1)
        // long runs of element slots in batches of CHUNK; the per-patch
        // partials (SRC 2, 3: a few per node) skip this loop and its registers
        if constexpr (SRC < 2) {
          for (; p + CHUNK <= p1; p += CHUNK) {
            TS x[CHUNK * D];
#pragma unroll
            for (int j = 0; j < CHUNK * D; ++j) x[j] = slots[int64_t(p) * D + j];
#pragma unroll
            for (int j = 0; j < CHUNK; ++j)
#pragma unroll
              for (int i = 0; i < D; ++i) add(i, x[j * D + i]);
          }
        }
        // tail in groups of 4, all loads of a group issued before the adds
        // (entries past the run reload the group's first entry, unused):
        // a node's few per-patch partials cost one round trip, not one each
        for (; p < p1; p += 4) {
          const int cnt4 = int(min<uint32_t>(4u, p1 - p));
          TS x[4 * D];
#pragma unroll
          for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < D; ++i) x[j * D + i] = slots[int64_t(j < cnt4 ? p + j : p) * D + i];
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (j < cnt4)
#pragma unroll
              for (int i = 0; i < D; ++i) add(i, x[j * D + i]);
        }
      }
      if constexpr (SRC < 2) {
        for (; p + CHUNK <= p1; p += CHUNK) {
          uint32_t s[CHUNK];
#pragma unroll
          for (int j = 0; j < CHUNK; ++j) s[j] = A.csr_slot[p + j];
          TS x[CHUNK][D];
#pragma unroll
          for (int j = 0; j < CHUNK; ++j)
#pragma unroll
            for (int i = 0; i < D; ++i) x[j][i] = slots[int64_t(s[j]) * D + i];
#pragma unroll
          for (int j = 0; j < CHUNK; ++j)
#pragma unroll
            for (int i = 0; i < D; ++i) add(i, x[j][i]);
        }
      }
      // tail in groups of 4 with all loads of a group issued first (the short
      // per-patch partial lists end up here entirely)
      for (; p < p1; p += 4) {
        const int cnt4 = int(min<uint32_t>(4u, p1 - p));
        uint32_t s[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) s[j] = A.csr_slot[j < cnt4 ? p + j : p];
        TS x[4][D];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int i = 0; i < D; ++i) x[j][i] = slots[int64_t(s[j]) * D + i];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (j < cnt4)
#pragma unroll
            for (int i = 0; i < D; ++i) add(i, x[j][i]);
      }
      double fi[D];
#pragma unroll
      for (int i = 0; i < D; ++i) {
        fi[i] = f[i];
        if constexpr (sizeof(TS) != 8) fi[i] = f[i] + c[i];
      }
      if (position) {
        // previous displacement: staged in s_v between two steps, read here
        // by the last corrector (once per run)
        double upl[D];
#pragma unroll
        for (int i = 0; i < D; ++i) upl[i] = pos_mid ? s_v[ln * D + i] : A.u_prev[node * D + i];
        node_position_form<D, HAS_FEXT>(A, node, dt, fi, HAS_FEXT ? &s_f[ln * D] : nullptr,
                                        &s_b[ln * D], upl, &s_u[ln * D], &s_v[ln * D],
                                        &s_a[ln * D]);
      } else {
        node_correct<D, HAS_FEXT>(A, node, dt, fi, HAS_FEXT ? &s_f[ln * D] : nullptr,
                                  &s_b[ln * D], &s_u[ln * D], &s_v[ln * D], &s_a[ln * D]);
      }
    }
    if (predict && !position) {
      node_predict<D>(A, node, dt, &s_b[ln * D], &s_u[ln * D], &s_v[ln * D], &s_a[ln * D]);
    } else if (predict && !gather) {
      // the position form's first predictor
      const double fz[D] = {};
      node_position_form<D, HAS_FEXT>(A, node, dt, fz, HAS_FEXT ? &s_f[ln * D] : nullptr,
                                      &s_b[ln * D], fz, &s_u[ln * D], &s_v[ln * D], &s_a[ln * D]);
    }
  }
  __syncthreads();
  if (correct || predict) {
    // u to u_out (the position form's other buffer) or in place; v and a
    // unless the position form leaves them (first predictor, between steps)
    double* __restrict__ uo = A.u_out ? A.u_out : A.u;
    const int nk2 = nk >> 1;
    const bool wv = !position || (correct && !predict);
    const bool wa = correct && !pos_mid;
    for (int t = threadIdx.x; t < nk2; t += nb) {
      reinterpret_cast<double2*>(uo + k0)[t] = reinterpret_cast<const double2*>(s_u)[t];
      if (wv) reinterpret_cast<double2*>(A.v + k0)[t] = reinterpret_cast<const double2*>(s_v)[t];
      if (wa) reinterpret_cast<double2*>(A.a + k0)[t] = reinterpret_cast<const double2*>(s_a)[t];
    }
    const int t2 = 2 * nk2 + int(threadIdx.x);
    if (t2 < nk) {
      uo[k0 + t2] = s_u[t2];
      if (wv) A.v[k0 + t2] = s_v[t2];
      if (wa) A.a[k0 + t2] = s_a[t2];
    }
  }
}

// ===========================================================================
// fused FP64 step over node blocks (stateless laws, slot layout 6)
//   Block b owns FUSED_NB consecutive nodes.  The elements touching them are
//   listed per block in the reference's assembly order; the block walks that
//   list in chunks of FUSED_CH elements: every thread computes one element's
//   forces (element_forces_fp64, the element kernel's arithmetic), writes the
//   contributions to the block's own nodes into LDS at their place in the
//   node's run for this chunk, and each node thread adds its run in order --
//   the chunks follow the assembly order, so every node's sum is the same
//   sequence of additions as the slot path's.  Then each node thread runs the
//   node kernel's update on its node.  Elements on block boundaries are
//   computed by every block they touch; no force slot goes through HBM.
//   u is read from NodeArgs::u and written to FusedArgs::u_out: other blocks
//   still read the old displacements while this one updates its nodes.
// ===========================================================================
template <int ET, int KIND, bool RECOMPUTE, bool HAS_FEXT>
__global__ void __launch_bounds__(256) k_step_fused(ElemArgs EA, NodeArgs NA, FusedArgs F) {
  static_assert(KIND != NH_DAMAGE, "the fused step recomputes elements: stateless laws only");
  static_assert(FUSED_NB <= 256 && FUSED_CH == 256, "one node and one element per thread of 256");
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, dn = d * n;
  extern __shared__ __align__(16) unsigned char smem[];
  double* s_f = reinterpret_cast<double*>(smem);  // FUSED_CH * n contributions, d each
  const int64_t b = logical_block(NA.xcd_remap);
  if (b >= F.nblk) return;  // uniform over the block
  const int t = threadIdx.x;
  const int64_t n0 = b * FUSED_NB;
  const int cnt = int(min<int64_t>(FUSED_NB, NA.N - n0));
  const bool owner = t < cnt;
  const int64_t node = n0 + t;
  const int64_t E = EA.E;
  const double* __restrict__ u_in = NA.u;
  double f[d];
#pragma unroll
  for (int i = 0; i < d; ++i) f[i] = 0.;
  const uint32_t j0 = F.el_off[b], j1 = F.el_off[b + 1];
  const uint32_t c0 = F.ch_off[b];
  uint32_t c = 0;
  for (uint32_t jc = j0; jc < j1; jc += FUSED_CH, ++c) {
    const uint32_t j = jc + uint32_t(t);
    if (j < j1) {
      const int64_t e = F.el[j];
      const MatParams m = EA.mats[EA.elem_mat ? EA.elem_mat[e] : 0];
      double ue[dn];
      double xe[RECOMPUTE ? dn : 1];
#pragma unroll
      for (int a = 0; a < n; ++a) {
        const int64_t nd = EA.conn[a * E + e];
#pragma unroll
        for (int i = 0; i < d; ++i) {
          ue[i + a * d] = u_in[nd * d + i];
          if constexpr (RECOMPUTE) xe[i + a * d] = EA.X[nd * d + i];
        }
      }
      double fe[dn];
      element_forces_fp64<ET, KIND, RECOMPUTE>(EA, e, m, ue, xe, fe);
#pragma unroll
      for (int a = 0; a < n; ++a) {
        const uint32_t ls = F.lslot[size_t(j) * n + a];
        if (ls != 0xFFFFu)
#pragma unroll
          for (int i = 0; i < d; ++i) s_f[ls * d + i] = fe[i + a * d];
      }
    }
    __syncthreads();
    if (owner) {
      // this chunk's run of the node, in assembly order (the node kernel's
      // sum: f += -1 * contribution)
      const uint16_t* cb = F.cbase + size_t(c0 + c) * (FUSED_NB + 1);
      const int q0 = cb[t], q1 = cb[t + 1];
      for (int q = q0; q < q1; ++q)
#pragma unroll
        for (int i = 0; i < d; ++i) {
          const double x = -1. * s_f[q * d + i];
          f[i] += x;
        }
    }
    __syncthreads();
  }
  if (!owner) return;
  const double dt = NA.dt_node ? NA.dt_node[node] : NA.dt;
  double u[d], v[d], acc[d], fx[d];
  uint8_t bl[d];
#pragma unroll
  for (int i = 0; i < d; ++i) {
    u[i] = u_in[node * d + i];
    v[i] = NA.v[node * d + i];
    acc[i] = NA.a[node * d + i];
    bl[i] = NA.blocked[node * d + i];
    fx[i] = HAS_FEXT ? NA.fext[node * d + i] : 0.;
  }
  if (NA.mode & (NODE_CORRECT | NODE_GATHER))
    node_correct<d, HAS_FEXT>(NA, node, dt, f, fx, bl, u, v, acc);
  if (NA.mode & NODE_PREDICT) node_predict<d>(NA, node, dt, bl, u, v, acc);
#pragma unroll
  for (int i = 0; i < d; ++i) {
    F.u_out[node * d + i] = u[i];
    NA.v[node * d + i] = v[i];
    NA.a[node * d + i] = acc[i];
  }
}

// ===========================================================================
// stable time step: min_e inradius(X + u) / c_p   (solid_mechanics_model.cc:495-533)
// ===========================================================================
__device__ __forceinline__ double dist2d(const double* x, const double* y) {
  return sqrt((y[0] - x[0]) * (y[0] - x[0]) + (y[1] - x[1]) * (y[1] - x[1]));
}
__device__ __forceinline__ double dist3d(const double* x, const double* y) {
  return sqrt((y[0] - x[0]) * (y[0] - x[0]) + (y[1] - x[1]) * (y[1] - x[1]) +
              (y[2] - x[2]) * (y[2] - x[2]));
}
// aka_math_tmpl.hh:568-587
__device__ double triangle_inradius(const double* c1, const double* c2, const double* c3) {
  double a = dist2d(c1, c2), b = dist2d(c2, c3), c = dist2d(c1, c3);
  double s = (a + b + c) * 0.5;
  return sqrt((s - a) * (s - b) * (s - c) / s);
}
// aka_math_tmpl.hh:596-679
__device__ double tet_inradius(const double* c1, const double* c2, const double* c3,
                               const double* c4) {
  double l12 = dist3d(c1, c2), l13 = dist3d(c1, c3), l14 = dist3d(c1, c4);
  double l23 = dist3d(c2, c3), l24 = dist3d(c2, c4), l34 = dist3d(c3, c4);
  double s1 = (l12 + l23 + l13) * 0.5; s1 = sqrt(s1 * (s1 - l12) * (s1 - l23) * (s1 - l13));
  double s2 = (l12 + l24 + l14) * 0.5; s2 = sqrt(s2 * (s2 - l12) * (s2 - l24) * (s2 - l14));
  double s3 = (l23 + l34 + l24) * 0.5; s3 = sqrt(s3 * (s3 - l23) * (s3 - l34) * (s3 - l24));
  double s4 = (l13 + l34 + l14) * 0.5; s4 = sqrt(s4 * (s4 - l13) * (s4 - l34) * (s4 - l14));
  double xx[9], vol;
  auto put = [&](const double* p, const double* q, const double* r) {
    xx[0] = p[0]; xx[1] = p[1]; xx[2] = p[2];
    xx[3] = q[0]; xx[4] = q[1]; xx[5] = q[2];
    xx[6] = r[0]; xx[7] = r[1]; xx[8] = r[2];
  };
  put(c2, c3, c4); vol = det3(xx);
  put(c1, c3, c4); vol -= det3(xx);
  put(c1, c2, c4); vol += det3(xx);
  put(c1, c2, c3); vol -= det3(xx);
  vol /= 6;
  return 3 * vol / (s1 + s2 + s3 + s4);
}

template <int ET>
__device__ double element_h(const double* x) {
  if constexpr (ET == TRI3) return 2. * triangle_inradius(x, x + 2, x + 4);
  if constexpr (ET == TET4) return 2. * tet_inradius(x, x + 3, x + 6, x + 9);
  if constexpr (ET == QUAD4) {
    // element_class_quadrangle_4_inline_impl.hh:131-154
    double a = dist2d(x + 0, x + 2), b = dist2d(x + 2, x + 4), c = dist2d(x + 4, x + 6),
           d = dist2d(x + 6, x + 0);
    return fmin(fmin(a, b), fmin(c, d));
  }
  if constexpr (ET == HEX8) {
    // element_class_hexahedron_8_inline_impl.hh:218-246
    const int ed[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                           {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
    double p = dist3d(x, x + 3);
    for (int k = 1; k < 12; ++k) p = fmin(p, dist3d(x + 3 * ed[k][0], x + 3 * ed[k][1]));
    return p;
  }
  return 0.;
}

template <typename Op>
__device__ __forceinline__ double block_reduce(double v, Op op, double init) {
  __shared__ double sh[32];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_down(v, o, 64));
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  const int nw = (blockDim.x + 63) >> 6;
  v = (int(threadIdx.x) < nw) ? sh[threadIdx.x] : init;
  if (wid == 0) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_down(v, o, 64));
  }
  return v;
}

struct OpMin { __device__ double operator()(double a, double b) const { return fmin(a, b); } };
struct OpSum { __device__ double operator()(double a, double b) const { return a + b; } };

template <int ET>
__global__ void __launch_bounds__(256) k_stable_dt(int64_t E, const double* __restrict__ X,
                                                   const double* __restrict__ u,
                                                   const uint32_t* __restrict__ conn,
                                                   const uint16_t* __restrict__ elem_mat,
                                                   const MatParams* __restrict__ mats,
                                                   double* partial) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n;
  double best = 1.7976931348623157e308;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < E;
       e += int64_t(gridDim.x) * blockDim.x) {
    double x[d * n];
#pragma unroll
    for (int a = 0; a < n; ++a) {
      const int64_t node = conn[a * E + e];
#pragma unroll
      for (int i = 0; i < d; ++i) {
        double p = X[node * d + i];
        p += u[node * d + i];
        x[i + a * d] = p;
      }
    }
    const MatParams m = mats[elem_mat ? elem_mat[e] : 0];
    const double c = sqrt((m.lambda + 2 * m.mu) / m.rho);
    best = fmin(best, element_h<ET>(x) / c);
  }
  best = block_reduce(best, OpMin(), 1.7976931348623157e308);
  if (threadIdx.x == 0) partial[blockIdx.x] = best;
}

// Stable time step of each material's elements (one independent body per
// material when bodies advance together): out[m] = min over the elements of
// material m of h_e / c_p(m), as the IEEE bit pattern of the double.  For
// positive doubles the bit patterns order like the values, so an unsigned
// 64-bit atomic minimum gives an exact, order-independent result.
template <int ET>
__global__ void __launch_bounds__(256) k_stable_dt_material(
    int64_t E, const double* __restrict__ X, const double* __restrict__ u,
    const uint32_t* __restrict__ conn, const uint16_t* __restrict__ elem_mat,
    const MatParams* __restrict__ mats, unsigned long long* out) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n;
  for (int64_t e0 = int64_t(blockIdx.x) * blockDim.x; e0 < E; e0 += int64_t(gridDim.x) * blockDim.x) {
    const int64_t e = e0 + threadIdx.x;
    const bool live = e < E;
    int mat = -1;
    double dte = 1.7976931348623157e308;
    if (live) {
      double x[d * n];
#pragma unroll
      for (int a = 0; a < n; ++a) {
        const int64_t node = conn[a * E + e];
#pragma unroll
        for (int i = 0; i < d; ++i) {
          double p = X[node * d + i];
          p += u[node * d + i];
          x[i + a * d] = p;
        }
      }
      mat = elem_mat ? elem_mat[e] : 0;
      const MatParams m = mats[mat];
      dte = element_h<ET>(x) / sqrt((m.lambda + 2 * m.mu) / m.rho);
    }
    // one atomic per wavefront when its live lanes share a material
    const int m0 = __shfl(mat, 0, 64);
    const bool uniform = __all(!live || mat == m0);
    if (uniform) {
      double v = dte;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
      if ((threadIdx.x & 63) == 0 && m0 >= 0)
        atomicMin(out + m0, static_cast<unsigned long long>(__double_as_longlong(v)));
    } else if (live) {
      atomicMin(out + mat, static_cast<unsigned long long>(__double_as_longlong(dte)));
    }
  }
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <int IS_MIN>
__global__ void __launch_bounds__(256) k_finish(const double* partial, int nb, double* out) {
  double v = IS_MIN ? 1.7976931348623157e308 : 0.;
  for (int i = threadIdx.x; i < nb; i += blockDim.x)
    v = IS_MIN ? fmin(v, partial[i]) : v + partial[i];
  if (IS_MIN) v = block_reduce(v, OpMin(), 1.7976931348623157e308);
  else v = block_reduce(v, OpSum(), 0.);
  if (threadIdx.x == 0) *out = v;
}

// ===========================================================================
// energies (solid_mechanics_model.cc:536-582, 615-662; material.cc:743-757)
// ===========================================================================
template <int D>
__global__ void __launch_bounds__(256) k_kinetic(int64_t N, const double* __restrict__ v,
                                                 const double* __restrict__ mass, double* partial) {
  double acc = 0.;
  for (int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; node < N;
       node += int64_t(gridDim.x) * blockDim.x) {
    const double m = mass[node];
    double mv2 = 0.;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const double vi = v[node * D + i];
      if (m > 2.220446049250313e-16) mv2 += vi * vi * m;
    }
    acc += mv2;
  }
  acc = block_reduce(acc, OpSum(), 0.);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void __launch_bounds__(256) k_external_work(int64_t ndof, const double* __restrict__ v,
                                                       const double* __restrict__ fext,
                                                       const double* __restrict__ fint,
                                                       const uint8_t* __restrict__ blocked,
                                                       double* partial) {
  double acc = 0.;
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < ndof;
       k += int64_t(gridDim.x) * blockDim.x) {
    if (blocked[k] & 1) acc -= fint[k] * v[k];
    else acc += fext[k] * v[k];
  }
  acc = block_reduce(acc, OpSum(), 0.);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

template <int ET, int KIND>
__global__ void __launch_bounds__(256) k_potential(int64_t count, const uint32_t* filter, int64_t E,
                                                   const double* __restrict__ u,
                                                   const uint32_t* __restrict__ conn,
                                                   const double* __restrict__ B64,
                                                   const double* __restrict__ J64,
                                                   const uint16_t* __restrict__ elem_mat,
                                                   const MatParams* __restrict__ mats,
                                                   const double* __restrict__ damage,
                                                   double* partial) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, Q = T::Q, dn = d * n;
  double acc = 0.;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < count;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t e = filter ? int64_t(filter[t]) : t;
    const MatParams m = mats[elem_mat ? elem_mat[e] : 0];
    double ue[dn];
#pragma unroll
    for (int a = 0; a < n; ++a) {
      const int64_t node = conn[a * E + e];
#pragma unroll
      for (int i = 0; i < d; ++i) ue[i + a * d] = u[node * d + i];
    }
    double se = 0.;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      double Bq[dn], G[d * d];
#pragma unroll
      for (int k = 0; k < dn; ++k) Bq[k] = B64[(int64_t(q) * dn + k) * E + e];
      matrix_matrixt<d, d, n>(ue, Bq, G);
      double w;
      if constexpr (KIND == ELASTIC) {
        double S[d * d];
        elastic_stress_ref<d>(m, G, S);
        double dd = 0.;
#pragma unroll
        for (int k = 0; k < d * d; ++k) dd += S[k] * G[k];
        w = .5 * dd;
      } else {
        w = neohookean_energy_ref<d>(m, G);
        if constexpr (KIND == NH_DAMAGE) w *= (1 - damage[e * Q + q]);
      }
      se += w * J64[int64_t(q) * E + e];
    }
    acc += se;
  }
  acc = block_reduce(acc, OpSum(), 0.);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// per-point grad u and stress on the current displacement (FP64, reference order;
// the damage state is read, not advanced)
template <int ET, int KIND>
__global__ void __launch_bounds__(256) k_point_fields(ElemArgs A, double* __restrict__ gradu,
                                                      double* __restrict__ stress) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, Q = T::Q, dn = d * n;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= A.count) return;
  const int64_t E = A.E;
  const int64_t e = A.filter ? int64_t(A.filter[t]) : t;
  const MatParams m = A.mats[A.elem_mat ? A.elem_mat[e] : 0];
  const double* __restrict__ Bt = static_cast<const double*>(A.B);
  double ue[dn];
#pragma unroll
  for (int a = 0; a < n; ++a) {
    const int64_t node = A.conn[a * E + e];
#pragma unroll
    for (int i = 0; i < d; ++i) ue[i + a * d] = A.u[node * d + i];
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    double Bq[dn], G[d * d], S[d * d];
#pragma unroll
    for (int k = 0; k < dn; ++k) Bq[k] = Bt[(int64_t(q) * dn + k) * E + e];
    matrix_matrixt<d, d, n>(ue, Bq, G);
    if constexpr (KIND == ELASTIC) {
      elastic_stress_ref<d>(m, G, S);
    } else {
      neohookean_stress_ref<d>(m, G, S);
      if constexpr (KIND == NH_DAMAGE) {
        const double dam = A.damage[e * Q + q];
#pragma unroll
        for (int k = 0; k < d * d; ++k) S[k] *= 1 - dam;
      }
    }
    const int64_t eq = e * Q + q;
#pragma unroll
    for (int k = 0; k < d * d; ++k) {
      gradu[eq * d * d + k] = G[k];
      stress[eq * d * d + k] = S[k];
    }
  }
}

__global__ void k_add(int64_t n, const double* x, const double* y, double* out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    double r = x[i];
    r += y[i];
    out[i] = r;
  }
}

// ===========================================================================
// launchers
// ===========================================================================
// Clear any error left by an earlier runtime call so that the status returned
// after a launch is the launch's own.
static inline void clear_error() { (void)hipGetLastError(); }

static inline unsigned grid_for(int64_t n, int block) {
  return unsigned((n + block - 1) / block);
}

#define AKX_ET_SWITCH(et, F)       \
  switch (et) {                    \
  case TRI3: F(TRI3); break;       \
  case QUAD4: F(QUAD4); break;     \
  case TET4: F(TET4); break;       \
  case HEX8: F(HEX8); break;       \
  default: return hipErrorInvalidValue; \
  }

hipError_t launch_precompute(int et, hipStream_t s, int64_t E, int64_t, const double* X,
                             const uint32_t* conn, const ElemTables& tab, double* B64, double* J64,
                             float* B32, float* J32, int* bad) {
  clear_error();
  if (E == 0) return hipSuccess;
#define F(ET) k_precompute<ET><<<grid_for(E, 256), 256, 0, s>>>(E, X, conn, tab, B64, J64, B32, J32, bad)
  AKX_ET_SWITCH(et, F)
#undef F
  return hipGetLastError();
}

hipError_t launch_mass(int et, hipStream_t s, int64_t N, int64_t E, const uint32_t* off,
                       const uint32_t* slot, const ElemTables& tab, const double* J64,
                       const uint16_t* elem_mat, const MatParams* mats, double* mass) {
  clear_error();
#define F(ET) k_mass<ET><<<grid_for(N, 256), 256, 0, s>>>(N, E, off, slot, tab, J64, elem_mat, mats, mass)
  AKX_ET_SWITCH(et, F)
#undef F
  return hipGetLastError();
}

template <int ET>
static hipError_t launch_element_et(int kind, int precision, hipStream_t s, const ElemArgs& A,
                                    int block) {
  const unsigned g = grid_for(A.count, block);
  if (precision == 0) {
    switch (kind) {
    case ELASTIC:
      if (A.recompute) k_element_fp64<ET, ELASTIC, true><<<g, block, 0, s>>>(A);
      else k_element_fp64<ET, ELASTIC, false><<<g, block, 0, s>>>(A);
      break;
    case NEOHOOKEAN:
      if (A.recompute) k_element_fp64<ET, NEOHOOKEAN, true><<<g, block, 0, s>>>(A);
      else k_element_fp64<ET, NEOHOOKEAN, false><<<g, block, 0, s>>>(A);
      break;
    case NH_DAMAGE:
      if (A.recompute) k_element_fp64<ET, NH_DAMAGE, true><<<g, block, 0, s>>>(A);
      else k_element_fp64<ET, NH_DAMAGE, false><<<g, block, 0, s>>>(A);
      break;
    default: return hipErrorInvalidValue;
    }
  } else if (A.sur_img) {
    if constexpr (ET == TET4) {
      if (kind != NH_DAMAGE) return hipErrorInvalidValue;
      const unsigned ge = grid_for(A.count, 256);
      if (A.patch)
        k_element_patch<TET4, SUR_STRAIN><<<unsigned(A.patch_blocks), 256, A.patch_lds, s>>>(A);
      else
        k_sur_strain<<<ge, 256, 0, s>>>(A);
      hipError_t err = hipGetLastError();
      if (err != hipSuccess) return err;
      if (A.net_ev0) (void)hipEventRecord(A.net_ev0, s);
      err = launch_sur_mlp(s, A.sur_img, A.sur_b, A.sur_bf16, A.count, A.sur_x, 8, A.sur_y, 8);
      if (err != hipSuccess) return err;
      if (A.net_ev1) (void)hipEventRecord(A.net_ev1, s);
      if (A.patch)
        k_element_patch<TET4, NH_DAMAGE_SUR><<<unsigned(A.patch_blocks), 256, A.patch_lds, s>>>(A);
      else
        k_sur_forces<<<ge, 256, 0, s>>>(A);
    } else {
      return hipErrorInvalidValue;
    }
  } else if (A.patch) {
    // one block per patch of 256 elements (the patches are built for that size)
    const unsigned gp = unsigned(A.patch_blocks);
    switch (kind) {
    case ELASTIC: k_element_patch<ET, ELASTIC><<<gp, 256, A.patch_lds, s>>>(A); break;
    case NEOHOOKEAN: k_element_patch<ET, NEOHOOKEAN><<<gp, 256, A.patch_lds, s>>>(A); break;
    case NH_DAMAGE: k_element_patch<ET, NH_DAMAGE><<<gp, 256, A.patch_lds, s>>>(A); break;
    default: return hipErrorInvalidValue;
    }
  } else {
    switch (kind) {
    case ELASTIC:
      if (A.recompute) k_element_mixed<ET, ELASTIC, true><<<g, block, 0, s>>>(A);
      else k_element_mixed<ET, ELASTIC, false><<<g, block, 0, s>>>(A);
      break;
    case NEOHOOKEAN:
      if (A.recompute) k_element_mixed<ET, NEOHOOKEAN, true><<<g, block, 0, s>>>(A);
      else k_element_mixed<ET, NEOHOOKEAN, false><<<g, block, 0, s>>>(A);
      break;
    case NH_DAMAGE:
      if (A.recompute) k_element_mixed<ET, NH_DAMAGE, true><<<g, block, 0, s>>>(A);
      else k_element_mixed<ET, NH_DAMAGE, false><<<g, block, 0, s>>>(A);
      break;
    default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

hipError_t launch_element(int et, int kind, int precision, hipStream_t s, const ElemArgs& A,
                          int block) {
  clear_error();
  if (A.count == 0) return hipSuccess;
#define F(ET) return launch_element_et<ET>(kind, precision, s, A, block)
  AKX_ET_SWITCH(et, F)
#undef F
  return hipSuccess;
}

template <int D>
static hipError_t launch_node_d(int precision, bool has_fext, hipStream_t s, const NodeArgs& A,
                                int block) {
  const unsigned g = grid_for(A.N, block);
  if (precision == 0) {
    if (A.partials == 2) {
      if (has_fext) k_node<D, 3, true><<<g, block, A.stage_lds, s>>>(A);
      else k_node<D, 3, false><<<g, block, A.stage_lds, s>>>(A);
    } else if (A.partials) {
      if (has_fext) k_node<D, 2, true><<<g, block, A.stage_lds, s>>>(A);
      else k_node<D, 2, false><<<g, block, A.stage_lds, s>>>(A);
    } else {
      if (has_fext) k_node<D, 0, true><<<g, block, A.stage_lds, s>>>(A);
      else k_node<D, 0, false><<<g, block, A.stage_lds, s>>>(A);
    }
  } else {
    if (has_fext) k_node<D, 1, true><<<g, block, A.stage_lds, s>>>(A);
    else k_node<D, 1, false><<<g, block, A.stage_lds, s>>>(A);
  }
  return hipGetLastError();
}

// The same fused step on the mixed path: element forces in FP32 from node
// differences formed in FP64 (mixed_element_forces, the patch kernel's
// arithmetic), contributions in LDS in FP32, each node's run summed in FP64,
// the node update in the position form of the mixed path.  Its
// sums follow the FP64 path's assembly order (any fixed order will do here).
// RECORD (triangles): the element data come from F.rec.
template <int ET, int KIND, bool HAS_FEXT, bool RECORD>
__global__ void __launch_bounds__(256) k_step_fused_mixed(ElemArgs EA, NodeArgs NA, FusedArgs F) {
  static_assert(!RECORD || ET == TRI3, "per-evaluation records exist for triangles only");
  static_assert(KIND != NH_DAMAGE, "the fused step recomputes elements: stateless laws only");
  static_assert(FUSED_NB <= 256 && FUSED_CH == 256, "one node and one element per thread of 256");
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, dn = d * n;
  extern __shared__ __align__(16) unsigned char smem[];
  float* s_f = reinterpret_cast<float*>(smem);  // FUSED_CH * n contributions, d each
  const int64_t b = logical_block(NA.xcd_remap);
  if (b >= F.nblk) return;  // uniform over the block
  const int t = threadIdx.x;
  const int64_t n0 = b * FUSED_NB;
  const int cnt = int(min<int64_t>(FUSED_NB, NA.N - n0));
  const bool owner = t < cnt;
  const int64_t node = n0 + t;
  const int64_t E = EA.E;
  const double* __restrict__ u_in = NA.u;
  double f[d];
#pragma unroll
  for (int i = 0; i < d; ++i) f[i] = 0.;
  const uint32_t j0 = F.el_off[b], j1 = F.el_off[b + 1];
  const uint32_t c0 = F.ch_off[b];
  uint32_t c = 0;
  for (uint32_t jc = j0; jc < j1; jc += FUSED_CH, ++c) {
    const uint32_t j = jc + uint32_t(t);
    if (j < j1) {
      // Triangles with records read one 32-byte record per evaluation, in
      // the block's evaluation order: the coordinate differences (the same
      // values k_fused_geometry formed from X), the node ids, the LDS slots
      // and the material, streamed instead of an element id and its
      // dependent gathers of connectivity, geometry and material; both forms
      // give the same forces bit for bit.  Larger elements share their
      // nodes among more elements and gather X (cached) faster than they
      // would read a record.
      constexpr int G = RECORD ? 4 : 1;
      static_assert(FUSED_CH * 3 < int(FUSED_REC_SLOT2_NONE), "LDS slots fit the records");
      int64_t e = 0;  // element id (not needed with a record)
      int mat = 0;
      int64_t nd[n];
      uint32_t lsl[n];
      float gx[G];
      if constexpr (RECORD) {
        const uint4 r0 = F.rec[size_t(j) * 2], r1 = F.rec[size_t(j) * 2 + 1];
        constexpr uint32_t LOW = (1u << FUSED_REC_NODE_BITS) - 1u;
        gx[0] = __uint_as_float(r0.x);
        gx[1] = __uint_as_float(r0.y);
        gx[2] = __uint_as_float(r0.z);
        gx[3] = __uint_as_float(r0.w);
        nd[0] = r1.x & LOW;
        nd[1] = r1.y & LOW;
        nd[2] = r1.z & LOW;
        lsl[0] = r1.w & 0xFFFFu;
        lsl[1] = r1.w >> 16;
        const uint32_t s2 = (r1.x >> 23) | ((r1.y >> 23) & 1u) << 9;
        lsl[2] = s2 == FUSED_REC_SLOT2_NONE ? 0xFFFFu : s2;
        mat = int((r1.y >> 24) | (r1.z >> 23) << 8);
      } else {
        e = F.el[j];
        mat = EA.elem_mat ? EA.elem_mat[e] : 0;
#pragma unroll
        for (int a = 0; a < n; ++a) nd[a] = EA.conn[a * E + e];
#pragma unroll
        for (int a = 0; a < n; ++a) lsl[a] = F.lslot[size_t(j) * n + a];
      }
      const MatParams m = EA.mats[mat];
      // differences to node 0 formed in FP64, then rounded
      float du[dn], dX[dn];
#pragma unroll
      for (int i = 0; i < d; ++i) {
        const double u0 = u_in[nd[0] * d + i];
        const double x0 = RECORD ? 0. : EA.X[nd[0] * d + i];
        du[i] = 0.f;
        dX[i] = 0.f;
#pragma unroll
        for (int a = 1; a < n; ++a) {
          du[i + a * d] = float(u_in[nd[a] * d + i] - u0);
          if constexpr (RECORD) dX[i + a * d] = gx[(a - 1) * d + i];
          else dX[i + a * d] = float(EA.X[nd[a] * d + i] - x0);
        }
      }
      float fe[dn];
      mixed_element_forces<ET, KIND, true>(EA, e, m, du, dX, fe);
#pragma unroll
      for (int a = 0; a < n; ++a) {
        const uint32_t ls = lsl[a];
        if (ls != 0xFFFFu)
#pragma unroll
          for (int i = 0; i < d; ++i) s_f[ls * d + i] = fe[i + a * d];
      }
    }
    __syncthreads();
    if (owner) {
      const uint16_t* cb = F.cbase + size_t(c0 + c) * (FUSED_NB + 1);
      const int q0 = cb[t], q1 = cb[t + 1];
      for (int q = q0; q < q1; ++q)
#pragma unroll
        for (int i = 0; i < d; ++i) f[i] += -1. * double(s_f[q * d + i]);
    }
    __syncthreads();
  }
  if (!owner) return;
  const int mode = NA.mode;
  // the position form (every launch of this kernel corrects): u_out holds
  // the previous displacement until this node overwrites it; between the
  // steps of a run neither v nor a is read or written
  const bool mid = mode & NODE_PREDICT;
  const double dt = mid ? 0. : NA.dt_node ? NA.dt_node[node] : NA.dt;
  double u[d], up[d], v[d], acc[d], fx[d];
  uint8_t bl[d];
#pragma unroll
  for (int i = 0; i < d; ++i) {
    u[i] = u_in[node * d + i];
    up[i] = F.u_out[node * d + i];
    v[i] = mid ? 0. : NA.v[node * d + i];
    acc[i] = mid ? 0. : NA.a[node * d + i];
    bl[i] = NA.blocked[node * d + i];
    fx[i] = HAS_FEXT ? NA.fext[node * d + i] : 0.;
  }
  node_position_form<d, HAS_FEXT>(NA, node, dt, f, fx, bl, up, u, v, acc);
#pragma unroll
  for (int i = 0; i < d; ++i) {
    F.u_out[node * d + i] = u[i];
    if (!mid) {
      NA.v[node * d + i] = v[i];
      NA.a[node * d + i] = acc[i];
    }
  }
}

// Geometry records of the mixed fused step on triangles: per element,
// x_a - x_0 for a > 0 formed in FP64 and rounded to FP32 (the arithmetic the
// element kernels use on the fly), zero-padded to fused_geo_stride floats.
template <int ET>
__global__ void __launch_bounds__(256) k_fused_geometry(int64_t E, const double* __restrict__ X,
                                                        const uint32_t* __restrict__ conn,
                                                        float* __restrict__ geo) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, G = fused_geo_stride(n, d);
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int64_t n0 = conn[e];
  float r[G];
#pragma unroll
  for (int k = 0; k < G; ++k) r[k] = 0.f;
#pragma unroll
  for (int a = 1; a < n; ++a) {
    const int64_t na = conn[a * E + e];
#pragma unroll
    for (int i = 0; i < d; ++i) r[(a - 1) * d + i] = float(X[na * d + i] - X[n0 * d + i]);
  }
#pragma unroll
  for (int k = 0; k < G; ++k) geo[e * G + k] = r[k];
}

// Per-evaluation records of the fused mixed step on triangles (FusedArgs::rec,
// recx), one thread per evaluation j of the blocks' element lists.
__global__ void __launch_bounds__(256) k_fused_records(int64_t J, int64_t E,
                                                       const uint32_t* __restrict__ el,
                                                       const uint32_t* __restrict__ conn,
                                                       const float* __restrict__ geo,
                                                       const uint16_t* __restrict__ lslot,
                                                       const uint16_t* __restrict__ elem_mat,
                                                       uint4* __restrict__ rec) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= J) return;
  const int64_t e = el[j];
  const float4 g = reinterpret_cast<const float4*>(geo)[e];
  rec[j * 2] = make_uint4(__float_as_uint(g.x), __float_as_uint(g.y), __float_as_uint(g.z),
                          __float_as_uint(g.w));
  const uint32_t mat = elem_mat ? uint32_t(elem_mat[e]) : 0u;
  rec[j * 2 + 1] = fused_record_ids(conn[e], conn[E + e], conn[2 * E + e], lslot[j * 3],
                                    lslot[j * 3 + 1], lslot[j * 3 + 2], mat);
}

hipError_t launch_fused_records(hipStream_t s, int64_t J, int64_t E, const uint32_t* el,
                                const uint32_t* conn_soa, const float* geo, const uint16_t* lslot,
                                const uint16_t* elem_mat, uint4* rec) {
  clear_error();
  if (J == 0) return hipSuccess;
  k_fused_records<<<unsigned((J + 255) / 256), 256, 0, s>>>(J, E, el, conn_soa, geo, lslot,
                                                             elem_mat, rec);
  return hipGetLastError();
}

hipError_t launch_fused_geometry(int et, hipStream_t s, int64_t E, const double* X,
                                 const uint32_t* conn_soa, float* geo) {
  clear_error();
  if (E == 0) return hipSuccess;
  const unsigned g = unsigned((E + 255) / 256);
  if (et != TRI3) return hipErrorInvalidValue;
  k_fused_geometry<TRI3><<<g, 256, 0, s>>>(E, X, conn_soa, geo);
  return hipGetLastError();
}

template <int ET, int KIND>
static hipError_t launch_fused_et(bool recompute, bool has_fext, hipStream_t s,
                                  const ElemArgs& EA, const NodeArgs& NA, const FusedArgs& F,
                                  int lds) {
  const unsigned g = unsigned(F.nblk);
  if (recompute) {
    if (has_fext) k_step_fused<ET, KIND, true, true><<<g, 256, lds, s>>>(EA, NA, F);
    else k_step_fused<ET, KIND, true, false><<<g, 256, lds, s>>>(EA, NA, F);
  } else {
    if (has_fext) k_step_fused<ET, KIND, false, true><<<g, 256, lds, s>>>(EA, NA, F);
    else k_step_fused<ET, KIND, false, false><<<g, 256, lds, s>>>(EA, NA, F);
  }
  return hipGetLastError();
}

template <int ET, int KIND>
static hipError_t launch_fused_mixed_et(bool has_fext, hipStream_t s, const ElemArgs& EA,
                                        const NodeArgs& NA, const FusedArgs& F, int lds) {
  const unsigned g = unsigned(F.nblk);
  if constexpr (ET == TRI3) {
    if (F.rec) {
      if (has_fext) k_step_fused_mixed<ET, KIND, true, true><<<g, 256, lds, s>>>(EA, NA, F);
      else k_step_fused_mixed<ET, KIND, false, true><<<g, 256, lds, s>>>(EA, NA, F);
      return hipGetLastError();
    }
  }
  if (has_fext) k_step_fused_mixed<ET, KIND, true, false><<<g, 256, lds, s>>>(EA, NA, F);
  else k_step_fused_mixed<ET, KIND, false, false><<<g, 256, lds, s>>>(EA, NA, F);
  return hipGetLastError();
}

hipError_t launch_fused(int et, int kind, int precision, bool recompute, bool has_fext,
                        hipStream_t s, const ElemArgs& EA, const NodeArgs& NA, const FusedArgs& F,
                        int lds) {
  clear_error();
  if (F.nblk == 0) return hipSuccess;
  if (kind != ELASTIC && kind != NEOHOOKEAN) return hipErrorInvalidValue;
  if (precision == 1) {
    // FP32 contributions: half the LDS of the FP64 step
#define M_(ET)                                                                            \
  return kind == ELASTIC ? launch_fused_mixed_et<ET, ELASTIC>(has_fext, s, EA, NA, F, lds / 2) \
                         : launch_fused_mixed_et<ET, NEOHOOKEAN>(has_fext, s, EA, NA, F, lds / 2);
    switch (et) {
    case TRI3: M_(TRI3);
    case QUAD4: M_(QUAD4);
    case TET4: M_(TET4);
    default: return hipErrorInvalidValue;
    }
#undef M_
  }
  // elements of up to 4 nodes (the host builds no fused tables for others)
#define F_(ET)                                                                       \
  return kind == ELASTIC ? launch_fused_et<ET, ELASTIC>(recompute, has_fext, s, EA, NA, F, lds) \
                         : launch_fused_et<ET, NEOHOOKEAN>(recompute, has_fext, s, EA, NA, F, lds);
  switch (et) {
  case TRI3: F_(TRI3);
  case QUAD4: F_(QUAD4);
  case TET4: F_(TET4);
  default: return hipErrorInvalidValue;
  }
#undef F_
}

hipError_t launch_node(int d, int precision, bool has_fext, hipStream_t s, const NodeArgs& A,
                       int block) {
  clear_error();
  if (A.N == 0) return hipSuccess;
  if (d == 2) return launch_node_d<2>(precision, has_fext, s, A, block);
  if (d == 3) return launch_node_d<3>(precision, has_fext, s, A, block);
  return hipErrorInvalidValue;
}

hipError_t launch_stable_dt(int et, hipStream_t s, int64_t E, const double* X, const double* u,
                            const uint32_t* conn, const uint16_t* elem_mat, const MatParams* mats,
                            double* partial, int nbmax, double* out) {
  clear_error();
  const int nb = int(std::min<int64_t>(nbmax, std::max<int64_t>(1, grid_for(E, 256))));
#define F(ET) k_stable_dt<ET><<<nb, 256, 0, s>>>(E, X, u, conn, elem_mat, mats, partial)
  AKX_ET_SWITCH(et, F)
#undef F
  k_finish<1><<<1, 256, 0, s>>>(partial, nb, out);
  return hipGetLastError();
}

hipError_t launch_stable_dt_material(int et, hipStream_t s, int64_t E, const double* X,
                                     const double* u, const uint32_t* conn,
                                     const uint16_t* elem_mat, const MatParams* mats, int nmat,
                                     unsigned long long* out) {
  clear_error();
  k_fill_u64<<<grid_for(nmat, 256), 256, 0, s>>>(out, nmat, 0x7FEFFFFFFFFFFFFFull);
  if (E == 0) return hipGetLastError();
  const unsigned nb = unsigned(std::min<int64_t>(4096, grid_for(E, 256)));
#define F(ET) k_stable_dt_material<ET><<<nb, 256, 0, s>>>(E, X, u, conn, elem_mat, mats, out)
  AKX_ET_SWITCH(et, F)
#undef F
  return hipGetLastError();
}

hipError_t launch_energy(int kind_energy, int et, int mat_kind, hipStream_t s, int64_t N,
                         int64_t E, const double* u, const double* v, const double* mass,
                         const double* fext, const double* fint, const uint8_t* blocked,
                         const uint32_t* conn, const double* B64, const double* J64,
                         const uint16_t* elem_mat, const MatParams* mats, const double* damage,
                         const uint32_t* filter, int64_t count, double* partial, int nbmax,
                         double* out) {
  clear_error();
  const int d = (et == TET4 || et == HEX8) ? 3 : 2;
  int nb = 1;
  if (kind_energy == 0) {
    nb = int(std::min<int64_t>(nbmax, std::max<int64_t>(1, grid_for(N, 256))));
    if (d == 2) k_kinetic<2><<<nb, 256, 0, s>>>(N, v, mass, partial);
    else k_kinetic<3><<<nb, 256, 0, s>>>(N, v, mass, partial);
  } else if (kind_energy == 2) {
    nb = int(std::min<int64_t>(nbmax, std::max<int64_t>(1, grid_for(N * d, 256))));
    k_external_work<<<nb, 256, 0, s>>>(N * d, v, fext, fint, blocked, partial);
  } else {
    nb = int(std::min<int64_t>(nbmax, std::max<int64_t>(1, grid_for(count, 256))));
#define G(ET, K) k_potential<ET, K><<<nb, 256, 0, s>>>(count, filter, E, u, conn, B64, J64, elem_mat, mats, damage, partial)
#define F(ET)                                      \
  switch (mat_kind) {                              \
  case ELASTIC: G(ET, ELASTIC); break;             \
  case NEOHOOKEAN: G(ET, NEOHOOKEAN); break;       \
  case NH_DAMAGE: G(ET, NH_DAMAGE); break;         \
  default: return hipErrorInvalidValue;            \
  }
    AKX_ET_SWITCH(et, F)
#undef F
#undef G
  }
  k_finish<0><<<1, 256, 0, s>>>(partial, nb, out);
  return hipGetLastError();
}

hipError_t launch_point_fields(int et, int kind, hipStream_t s, const ElemArgs& A, double* gradu,
                               double* stress) {
  clear_error();
  if (A.count == 0) return hipSuccess;
  const unsigned g = grid_for(A.count, 256);
#define G(ET, K) k_point_fields<ET, K><<<g, 256, 0, s>>>(A, gradu, stress)
#define F(ET)                                      \
  switch (kind) {                                  \
  case ELASTIC: G(ET, ELASTIC); break;             \
  case NEOHOOKEAN: G(ET, NEOHOOKEAN); break;       \
  case NH_DAMAGE: G(ET, NH_DAMAGE); break;         \
  default: return hipErrorInvalidValue;            \
  }
  AKX_ET_SWITCH(et, F)
#undef F
#undef G
  return hipGetLastError();
}

// Snapshot of a per-row device array in the caller's numbering, for the
// asynchronous downloads: dst[r * dw + j] = src[inv[r] * sw + (sw == 1 ? 0 : j)]
// (rows = nodes or elements; inv = caller -> internal, nullptr = identity;
// sw = 1 broadcasts one value per row, as the lumped mass is reported per dof).
template <typename T>
__global__ void __launch_bounds__(256) k_snapshot(int64_t rows, int dw, int sw,
                                                  const uint32_t* __restrict__ inv,
                                                  const T* __restrict__ src, T* __restrict__ dst,
                                                  int mask) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= rows * dw) return;
  const int64_t r = k / dw;
  const int j = int(k - r * dw);
  const int64_t from = (inv ? int64_t(inv[r]) : r) * sw + (sw == 1 ? 0 : j);
  T v = src[from];
  if constexpr (sizeof(T) == 1) v = T(v & mask);
  dst[k] = v;
}

hipError_t launch_snapshot(hipStream_t s, int64_t rows, int dw, int sw, const uint32_t* inv,
                           const void* src, void* dst, int elem_bytes) {
  clear_error();
  const int64_t n = rows * dw;
  if (n == 0) return hipSuccess;
  if (elem_bytes == 8)
    k_snapshot<double><<<grid_for(n, 256), 256, 0, s>>>(rows, dw, sw, inv,
                                                        static_cast<const double*>(src),
                                                        static_cast<double*>(dst), 0);
  else
    k_snapshot<uint8_t><<<grid_for(n, 256), 256, 0, s>>>(rows, dw, sw, inv,
                                                         static_cast<const uint8_t*>(src),
                                                         static_cast<uint8_t*>(dst), 1);
  return hipGetLastError();
}

hipError_t launch_add(hipStream_t s, int64_t n, const double* x, const double* y, double* out) {
  clear_error();
  if (n == 0) return hipSuccess;
  k_add<<<grid_for(n, 256), 256, 0, s>>>(n, x, y, out);
  return hipGetLastError();
}

}  // namespace akx
