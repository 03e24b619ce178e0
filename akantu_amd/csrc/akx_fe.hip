// akx_fe.hip -- FEEngine and Material operators on caller fields, in FP64 and
// in the reference's operation order (the building blocks a driver composes
// itself, next to the fused explicit step of akx_kernels.hip):
//
//   k_fe_gradient      gradientOnIntegrationPoints  (shape_functions_inline_impl.hh:352-395)
//   k_fe_interpolate   interpolateOnIntegrationPoints (shape_functions_inline_impl.hh:300-335)
//   k_fe_btd           computeBtD                   (shape_lagrange_inline_impl.hh:382-423)
//   k_fe_ntb           computeNtb                   (shape_lagrange_inline_impl.hh:495-533)
//   k_fe_integrate     IntegratorGauss::integrate   (integrator_gauss_inline_impl.hh:426-483)
//   k_fe_integrate_pts integrateOnIntegrationPoints (integrator_gauss_inline_impl.hh:545-597)
//   k_reduce_pairs     Math::reduce's pairwise passes (aka_math.cc:225-250), after a sort
//   k_fe_lumped_*      assembleFieldLumped row sum  (fe_engine_template_tmpl_field.hh:126-199)
//   k_mat_*            Material::computeStress / computeAllStresses / assembleInternalForces
//                      (material.hh:110-195, material.cc:178-279,637-722)
//
// Element-indexed inputs and outputs are rows in the order of the element list
// the caller gave (t), while the element tables are read at the element's
// internal id (filt[t]).  Nodal arrays are in the internal node numbering.
// These run once per operator call; they are written for clarity of the
// reference order, one thread per element (or per element and component).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "akx_internal.hpp"

namespace akx {

namespace {

__device__ __forceinline__ int64_t elem_of(const FeArgs& A, int64_t t) {
  return A.filt ? int64_t(A.filt[t]) : t;
}

template <int ET>
__global__ void __launch_bounds__(256) k_fe_gradient(FeArgs A) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, Q = T::Q, dn = d * n;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= A.count) return;
  const int64_t e = elem_of(A, t), E = A.E;
  const int nb = A.nb;
  double ue[FE_MAX_DOF * 8];
  for (int a = 0; a < n; ++a) {
    const int64_t node = A.conn[a * E + e];
    for (int i = 0; i < nb; ++i) ue[i + a * nb] = A.in[node * nb + i];
  }
  for (int q = 0; q < Q; ++q) {
    double Bq[dn];
#pragma unroll
    for (int k = 0; k < dn; ++k) Bq[k] = A.B[(int64_t(q) * dn + k) * E + e];
    double* o = A.out + (t * Q + q) * nb * d;
    // nabla_u (nb x d) = u_e (nb x n) * B^T, B (d x n)
    for (int j = 0; j < d; ++j)
      for (int i = 0; i < nb; ++i) {
        double s = 0.;
#pragma unroll
        for (int l = 0; l < n; ++l) s += ue[i + l * nb] * Bq[j + l * d];
        o[i + j * nb] = s;
      }
  }
}

template <int ET>
__global__ void __launch_bounds__(256) k_fe_interpolate(FeArgs A) {
  using T = Elem<ET>;
  constexpr int n = T::n, Q = T::Q;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= A.count) return;
  const int64_t e = elem_of(A, t), E = A.E;
  const int nb = A.nb;
  double ue[FE_MAX_DOF * 8];
  for (int a = 0; a < n; ++a) {
    const int64_t node = A.conn[a * E + e];
    for (int i = 0; i < nb; ++i) ue[i + a * nb] = A.in[node * nb + i];
  }
  // uq (nb x Q) = u_e (nb x n) * N (n x Q)
  double* o = A.out + t * Q * nb;
  for (int q = 0; q < Q; ++q)
    for (int i = 0; i < nb; ++i) {
      double s = 0.;
#pragma unroll
      for (int l = 0; l < n; ++l) s += ue[i + l * nb] * A.tab.N[q][l];
      o[i + q * nb] = s;
    }
}

template <int ET>
__global__ void __launch_bounds__(256) k_fe_btd(FeArgs A) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, Q = T::Q, dn = d * n;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= A.count) return;
  const int64_t e = elem_of(A, t), E = A.E;
  const int rows = A.nb;
  for (int q = 0; q < Q; ++q) {
    double Bq[dn];
#pragma unroll
    for (int k = 0; k < dn; ++k) Bq[k] = A.B[(int64_t(q) * dn + k) * E + e];
    const int64_t r = t * Q + q;
    const double* D = A.in + r * rows * d;
    double* o = A.out + r * rows * n;
    // Bt_D (rows x n) = D (rows x d) * B (d x n)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < rows; ++i) {
        double s = 0.;
#pragma unroll
        for (int l = 0; l < d; ++l) s += D[i + l * rows] * Bq[l + j * d];
        o[i + j * rows] = s;
      }
  }
}

template <int ET>
__global__ void __launch_bounds__(256) k_fe_ntb(FeArgs A) {
  using T = Elem<ET>;
  constexpr int n = T::n, Q = T::Q;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= A.count) return;
  const int nb = A.nb;
  for (int q = 0; q < Q; ++q) {
    const int64_t r = t * Q + q;
    const double* b = A.in + r * nb;
    double* o = A.out + r * nb * n;
    // Ntb (nb x n) = b (nb x 1) * N (1 x n)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < nb; ++i) {
        double s = 0.;
        s += b[i] * A.tab.N[q][j];
        o[i + j * nb] = s;
      }
  }
}

// one thread per (element, component): intf[t][k] = sum_q f[t][q][k] J[q]
template <int ET>
__global__ void __launch_bounds__(256) k_fe_integrate(FeArgs A) {
  constexpr int Q = Elem<ET>::Q;
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int nb = A.nb;
  if (k >= A.count * nb) return;
  const int64_t t = k / nb;
  const int i = int(k - t * nb);
  const int64_t e = elem_of(A, t);
  const double* f = A.in + t * Q * nb;
  double s = 0.;
#pragma unroll
  for (int q = 0; q < Q; ++q) s += f[i + q * nb] * A.J[int64_t(q) * A.E + e];
  A.out[t * nb + i] = s;
}

template <int ET>
__global__ void __launch_bounds__(256) k_fe_integrate_pts(FeArgs A) {
  constexpr int Q = Elem<ET>::Q;
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int nb = A.nb;
  if (k >= A.count * Q * nb) return;
  const int64_t r = k / nb;
  const int64_t t = r / Q;
  const int q = int(r - t * Q);
  double v = A.in[k];
  v *= A.J[int64_t(q) * A.E + elem_of(A, t)];
  A.out[k] = v;
}

// One pass of Math::reduce on a sorted array of m values: the last value is
// folded into the one before it when m is odd, then out[i] = a[2i] + a[2i+1].
__global__ void __launch_bounds__(256) k_reduce_pairs(int64_t m, const double* __restrict__ a,
                                                      double* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t half = m >> 1;
  if (i >= half) return;
  double lo = a[2 * i], hi = a[2 * i + 1];
  if ((m & 1) && 2 * i + 1 == m - 2) hi += a[m - 1];
  out[i] = lo + hi;
}

// Row-sum lumping, element part: integ[e][a * nb + k] = sum_q (0 + field[e][q][k]
// N_a(q)) J[e][q], field in internal element order.
template <int ET>
__global__ void __launch_bounds__(256) k_fe_lumped_elements(FeArgs A) {
  using T = Elem<ET>;
  constexpr int n = T::n, Q = T::Q;
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int nb = A.nb;
  if (k >= A.E * n * nb) return;
  const int64_t e = k / (n * nb);
  const int c = int(k - e * n * nb);
  const int a = c / nb, i = c - a * nb;
  double s = 0.;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    double ntb = 0.;
    ntb += A.in[(e * Q + q) * nb + i] * A.tab.N[q][a];
    s += ntb * A.J[int64_t(q) * A.E + e];
  }
  A.out[k] = s;
}

// Row-sum lumping, node part: each node adds its element contributions in
// ascending element order (the mass CSR), then the sum once to lumped.
__global__ void __launch_bounds__(256) k_fe_lumped_nodes(int64_t N, int nb, int n,
                                                         const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ slot,
                                                         const double* __restrict__ integ,
                                                         double* __restrict__ lumped) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= N * nb) return;
  const int64_t node = k / nb;
  const int i = int(k - node * nb);
  double local = 0.;
  for (uint32_t p = off[node]; p < off[node + 1]; ++p) {
    const uint32_t s = slot[p];  // element * n + local node
    local += 1. * integ[int64_t(s) * nb + i];
  }
  double r = lumped[k];
  r += 1. * local;
  lumped[k] = r;
}

// ---------------------------------------------------------------------------
// Material operators on one material's elements (FeArgs.filt: its internal
// element ids, in the material's element order).
// ---------------------------------------------------------------------------
template <int d, int KIND>
__device__ __forceinline__ void point_law(const MatArgs& M, int64_t eq, const double* G, double* S) {
  if constexpr (KIND == ELASTIC) {
    elastic_stress_ref<d>(M.mat, G, S);
  } else if constexpr (KIND == NEOHOOKEAN) {
    neohookean_stress_ref<d>(M.mat, G, S);
  } else {
    double dam = M.damage[eq];
    nh_damage_stress_ref<d>(M.mat, G, S, dam, M.Yd[eq]);
    M.damage[eq] = dam;
  }
}

// computeStress from given gradients (rows t*Q + q of M.in_grad); the
// material's internals gradu / stress (internal rows e*Q + q) are kept.
template <int ET, int KIND>
__global__ void __launch_bounds__(256) k_mat_stress(FeArgs A, MatArgs M) {
  using T = Elem<ET>;
  constexpr int d = T::d, Q = T::Q, dd = d * d;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= A.count) return;
  const int64_t e = elem_of(A, t);
  for (int q = 0; q < Q; ++q) {
    const int64_t eq = e * Q + q, r = t * Q + q;
    double G[dd], S[dd];
#pragma unroll
    for (int k = 0; k < dd; ++k) G[k] = A.in[r * dd + k];
    point_law<d, KIND>(M, eq, G, S);
#pragma unroll
    for (int k = 0; k < dd; ++k) {
      M.gradu[eq * dd + k] = G[k];
      M.stress[eq * dd + k] = S[k];
      if (A.out) A.out[r * dd + k] = S[k];
    }
  }
}

// computeAllStresses: gradient of the current displacement (minus a zero
// eigengradient) and computeStress.
template <int ET, int KIND>
__global__ void __launch_bounds__(256) k_mat_all_stresses(FeArgs A, MatArgs M) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, Q = T::Q, dn = d * n, dd = d * d;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= A.count) return;
  const int64_t e = elem_of(A, t), E = A.E;
  double ue[dn];
#pragma unroll
  for (int a = 0; a < n; ++a) {
    const int64_t node = A.conn[a * E + e];
#pragma unroll
    for (int i = 0; i < d; ++i) ue[i + a * d] = A.in[node * d + i];
  }
  for (int q = 0; q < Q; ++q) {
    double Bq[dn], G[dd], S[dd];
#pragma unroll
    for (int k = 0; k < dn; ++k) Bq[k] = A.B[(int64_t(q) * dn + k) * E + e];
    matrix_matrixt<d, d, n>(ue, Bq, G);
#pragma unroll
    for (int k = 0; k < dd; ++k) G[k] -= 0.;
    const int64_t eq = e * Q + q;
    point_law<d, KIND>(M, eq, G, S);
#pragma unroll
    for (int k = 0; k < dd; ++k) {
      M.gradu[eq * dd + k] = G[k];
      M.stress[eq * dd + k] = S[k];
    }
  }
}

// assembleInternalForces, element part: B^T sigma (small strain) or the
// finite-strain residual from the stored internals, integrated over the
// points, into element-ordered slots [e][a*d + i].
template <int ET, int KIND>
__global__ void __launch_bounds__(256) k_mat_forces(FeArgs A, MatArgs M) {
  using T = Elem<ET>;
  constexpr int d = T::d, n = T::n, Q = T::Q, dn = d * n, dd = d * d;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= A.count) return;
  const int64_t e = elem_of(A, t), E = A.E;
  double fe[dn];
#pragma unroll
  for (int k = 0; k < dn; ++k) fe[k] = 0.;
  for (int q = 0; q < Q; ++q) {
    double Bq[dn], R[dn], G[dd], S[dd];
#pragma unroll
    for (int k = 0; k < dn; ++k) Bq[k] = A.B[(int64_t(q) * dn + k) * E + e];
    const int64_t eq = e * Q + q;
#pragma unroll
    for (int k = 0; k < dd; ++k) {
      G[k] = M.gradu[eq * dd + k];
      S[k] = M.stress[eq * dd + k];
    }
    if constexpr (KIND == ELASTIC) matrix_matrix<d, n, d>(S, Bq, R);
    else finite_residual_ref<d, n>(Bq, G, S, R);
    const double jq = A.J[int64_t(q) * E + e];
#pragma unroll
    for (int k = 0; k < dn; ++k) fe[k] += R[k] * jq;
  }
#pragma unroll
  for (int k = 0; k < dn; ++k) A.out[e * dn + k] = fe[k];
}

// assembleInternalForces, node part: f_int += -1 * f_e of the material's
// elements, in the assembly order of the force CSR (material by material,
// ascending elements), on top of the value already there.
__global__ void __launch_bounds__(256) k_mat_gather(int64_t N, int d, int n, int mat,
                                                    const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ slot,
                                                    const uint16_t* __restrict__ elem_mat,
                                                    const double* __restrict__ fe,
                                                    double* __restrict__ fint) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= N * d) return;
  const int64_t node = k / d;
  const int i = int(k - node * d);
  double r = fint[k];
  for (uint32_t p = off[node]; p < off[node + 1]; ++p) {
    const uint32_t s = slot[p];
    const uint32_t e = s / uint32_t(n);
    if ((elem_mat ? int(elem_mat[e]) : 0) != mat) continue;
    r += -1. * fe[int64_t(s) * d + i];
  }
  fint[k] = r;
}

inline void clear_error() { (void)hipGetLastError(); }
inline unsigned grid_for(int64_t n) { return unsigned((n + 255) / 256); }

#define AKX_FE_SWITCH(et, F)              \
  switch (et) {                           \
  case TRI3: F(TRI3); break;              \
  case QUAD4: F(QUAD4); break;            \
  case TET4: F(TET4); break;              \
  case HEX8: F(HEX8); break;              \
  default: return hipErrorInvalidValue;   \
  }

}  // namespace

hipError_t launch_fe(int op, int et, hipStream_t s, const FeArgs& A) {
  clear_error();
  if (A.count <= 0 && op != FE_LUMPED_ELEMENTS) return hipSuccess;
  const int n = et == TRI3 ? 3 : et == HEX8 ? 8 : 4;
  const int Q = et == QUAD4 ? 4 : et == HEX8 ? 8 : 1;
  switch (op) {
  case FE_GRADIENT:
#define F(ET) k_fe_gradient<ET><<<grid_for(A.count), 256, 0, s>>>(A)
    AKX_FE_SWITCH(et, F)
#undef F
    break;
  case FE_INTERPOLATE:
#define F(ET) k_fe_interpolate<ET><<<grid_for(A.count), 256, 0, s>>>(A)
    AKX_FE_SWITCH(et, F)
#undef F
    break;
  case FE_BTD:
#define F(ET) k_fe_btd<ET><<<grid_for(A.count), 256, 0, s>>>(A)
    AKX_FE_SWITCH(et, F)
#undef F
    break;
  case FE_NTB:
#define F(ET) k_fe_ntb<ET><<<grid_for(A.count), 256, 0, s>>>(A)
    AKX_FE_SWITCH(et, F)
#undef F
    break;
  case FE_INTEGRATE:
#define F(ET) k_fe_integrate<ET><<<grid_for(A.count * A.nb), 256, 0, s>>>(A)
    AKX_FE_SWITCH(et, F)
#undef F
    break;
  case FE_INTEGRATE_POINTS:
#define F(ET) k_fe_integrate_pts<ET><<<grid_for(A.count * Q * A.nb), 256, 0, s>>>(A)
    AKX_FE_SWITCH(et, F)
#undef F
    break;
  case FE_LUMPED_ELEMENTS:
    if (A.E <= 0) return hipSuccess;
#define F(ET) k_fe_lumped_elements<ET><<<grid_for(A.E * n * A.nb), 256, 0, s>>>(A)
    AKX_FE_SWITCH(et, F)
#undef F
    break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_fe_lumped_nodes(hipStream_t s, int64_t N, int nb, int n, const uint32_t* off,
                                  const uint32_t* slot, const double* integ, double* lumped) {
  clear_error();
  if (N * nb > 0)
    k_fe_lumped_nodes<<<grid_for(N * nb), 256, 0, s>>>(N, nb, n, off, slot, integ, lumped);
  return hipGetLastError();
}

// Math::reduce of m device values (in place: `a` is sorted and then used as
// scratch, `tmp` holds m values).  The result lands in *out (device).
hipError_t launch_reduce_sorted(hipStream_t s, int64_t m, double* a, double* tmp, double* out,
                                void* sort_tmp, size_t* sort_bytes) {
  clear_error();
  if (sort_tmp == nullptr) {
    // size query for the sort's temporary storage
    return hipcub::DeviceRadixSort::SortKeys(nullptr, *sort_bytes, a, tmp, int(m), 0, 64, s);
  }
  hipError_t e = hipcub::DeviceRadixSort::SortKeys(sort_tmp, *sort_bytes, a, tmp, int(m), 0, 64, s);
  if (e != hipSuccess) return e;
  // the sorted values are in tmp; ping-pong between tmp and a
  double* src = tmp;
  double* dst = a;
  while (m > 1) {
    k_reduce_pairs<<<grid_for(m >> 1), 256, 0, s>>>(m, src, dst);
    m >>= 1;
    double* x = src;
    src = dst;
    dst = x;
  }
  return hipMemcpyAsync(out, src, sizeof(double), hipMemcpyDeviceToDevice, s);
}

hipError_t launch_material(int op, int et, int kind, hipStream_t s, const FeArgs& A,
                           const MatArgs& M) {
  clear_error();
  if (A.count <= 0) return hipSuccess;
  const unsigned g = grid_for(A.count);
#define K(KERNEL, ET)                                                          \
  switch (kind) {                                                              \
  case ELASTIC: KERNEL<ET, ELASTIC><<<g, 256, 0, s>>>(A, M); break;            \
  case NEOHOOKEAN: KERNEL<ET, NEOHOOKEAN><<<g, 256, 0, s>>>(A, M); break;      \
  case NH_DAMAGE: KERNEL<ET, NH_DAMAGE><<<g, 256, 0, s>>>(A, M); break;        \
  default: return hipErrorInvalidValue;                                        \
  }
  switch (op) {
  case MAT_STRESS:
#define F(ET) K(k_mat_stress, ET)
    AKX_FE_SWITCH(et, F)
#undef F
    break;
  case MAT_ALL_STRESSES:
#define F(ET) K(k_mat_all_stresses, ET)
    AKX_FE_SWITCH(et, F)
#undef F
    break;
  case MAT_FORCES:
#define F(ET) K(k_mat_forces, ET)
    AKX_FE_SWITCH(et, F)
#undef F
    break;
  default: return hipErrorInvalidValue;
  }
#undef K
  return hipGetLastError();
}

hipError_t launch_material_gather(hipStream_t s, int64_t N, int d, int n, int mat,
                                  const uint32_t* off, const uint32_t* slot,
                                  const uint16_t* elem_mat, const double* fe, double* fint) {
  clear_error();
  if (N * d > 0)
    k_mat_gather<<<grid_for(N * d), 256, 0, s>>>(N, d, n, mat, off, slot, elem_mat, fe, fint);
  return hipGetLastError();
}

}  // namespace akx
