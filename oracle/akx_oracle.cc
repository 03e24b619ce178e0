// akx_oracle.cc -- CPU FP64 restatement of Akantu's explicit central-difference
// path (SolidMechanicsModel + lumped mass + CentralDifference).
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: only tests/, the smoke
// check and the CPU-baseline leg of bench.py may load it, and only as the
// checker (or as the timed CPU reference loop).  The product path never links
// or calls it.
//
// Every routine below restates one reference routine in the reference's own
// operation order (row-major nodal arrays, column-major per-point matrices,
// element-ascending scatter-add, the non-BLAS small-matrix loops), so that a
// build with -ffp-contract=off reproduces the reference arithmetic bit for bit.
// Reference paths are relative to the Akantu 3.1.0 source tree.
//
// Pinning (see tests/test_oracle_*.py):
//   * integrator (predictor / residual / lumped solve / corrector):
//     test/test_model/test_common/test_model_solver/
//       test_model_solver_dynamic_explicit.verified  (402-line golden, KAT)
//   * precompute, gradient, elastic stress, B^T sigma, integrate, assemble:
//     patch tests on the reference's own meshes (linear field exact,
//     f_int + f_ext = 0, E_kin = 0 after 100 steps), gradient/integrate KATs,
//     lumped-mass sum, elastic-stress rotated-strain KAT.
//   * neo-Hookean and the neo-Hookean + Marigo damage composition are
//     not covered by any reference test (they are disabled there); they are
//     pinned by invariants: zero strain -> zero stress, small-strain limit ->
//     linear elastic, objectivity, S = 2 dW/dC by finite differences, and
//     the Marigo branch special cases.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>
#include <algorithm>

namespace {

enum ElemType { TRIANGLE_3 = 0, QUADRANGLE_4 = 1, TETRAHEDRON_4 = 2, HEXAHEDRON_8 = 3 };
enum MatKind { ELASTIC = 0, NEOHOOKEAN = 1, NEOHOOKEAN_DAMAGE = 2 };

struct ElemInfo {
  int dim, nnodes, nqp;
  double qp[8][3];
  double w[8];
};

// Gauss tables: src/fe_engine/gauss_integration.cc:72-73 (triangle, 1 pt),
// :122-123 (tetrahedron, 1 pt), :55-56 (segment, 2 pts -> tensor product for
// quadrangle/hexahedron, point ordering xi fastest, gauss_integration_tmpl.hh:131-160).
ElemInfo elem_info(int type) {
  ElemInfo e{};
  const double g = 1. / std::sqrt(3.);
  switch (type) {
  case TRIANGLE_3:
    e.dim = 2; e.nnodes = 3; e.nqp = 1;
    e.qp[0][0] = 1. / 3.; e.qp[0][1] = 1. / 3.; e.w[0] = 1. / 2.;
    break;
  case QUADRANGLE_4:
    e.dim = 2; e.nnodes = 4; e.nqp = 4;
    for (int q = 0; q < 4; ++q) {
      e.qp[q][0] = (q & 1) ? g : -g;
      e.qp[q][1] = (q & 2) ? g : -g;
      e.w[q] = 1.;
    }
    break;
  case TETRAHEDRON_4:
    e.dim = 3; e.nnodes = 4; e.nqp = 1;
    e.qp[0][0] = e.qp[0][1] = e.qp[0][2] = 1. / 4.; e.w[0] = 1. / 6.;
    break;
  case HEXAHEDRON_8:
    e.dim = 3; e.nnodes = 8; e.nqp = 8;
    for (int q = 0; q < 8; ++q) {
      e.qp[q][0] = (q & 1) ? g : -g;
      e.qp[q][1] = (q & 2) ? g : -g;
      e.qp[q][2] = (q & 4) ? g : -g;
      e.w[q] = 1.;
    }
    break;
  default:
    e.dim = 0;
  }
  return e;
}

// Shape functions N(xi).
//   triangle_3:     element_class_triangle_3_inline_impl.hh:62-76
//   quadrangle_4:   element_class_quadrangle_4_inline_impl.hh (bilinear, .25 prefactor)
//   tetrahedron_4:  element_class_tetrahedron_4_inline_impl.hh:68-83 (permuted coords)
//   hexahedron_8:   element_class_hexahedron_8_inline_impl.hh:117-130
void shapes(int type, const double* c, double* N) {
  switch (type) {
  case TRIANGLE_3: {
    double c0 = 1 - c[0] - c[1];
    N[0] = c0; N[1] = c[0]; N[2] = c[1];
  } break;
  case QUADRANGLE_4:
    N[0] = .25 * (1 - c[0]) * (1 - c[1]);
    N[1] = .25 * (1 + c[0]) * (1 - c[1]);
    N[2] = .25 * (1 + c[0]) * (1 + c[1]);
    N[3] = .25 * (1 - c[0]) * (1 + c[1]);
    break;
  case TETRAHEDRON_4: {
    double c0 = 1 - c[0] - c[1] - c[2];
    N[0] = c0; N[1] = c[1]; N[2] = c[2]; N[3] = c[0];
  } break;
  case HEXAHEDRON_8:
    N[0] = .125 * (1 - c[0]) * (1 - c[1]) * (1 - c[2]);
    N[1] = .125 * (1 + c[0]) * (1 - c[1]) * (1 - c[2]);
    N[2] = .125 * (1 + c[0]) * (1 + c[1]) * (1 - c[2]);
    N[3] = .125 * (1 - c[0]) * (1 + c[1]) * (1 - c[2]);
    N[4] = .125 * (1 - c[0]) * (1 - c[1]) * (1 + c[2]);
    N[5] = .125 * (1 + c[0]) * (1 - c[1]) * (1 + c[2]);
    N[6] = .125 * (1 + c[0]) * (1 + c[1]) * (1 + c[2]);
    N[7] = .125 * (1 - c[0]) * (1 + c[1]) * (1 + c[2]);
    break;
  }
}

// dN/dxi, stored column-major natdim x n: out[i + a*natdim] = dN_a/dxi_i.
//   triangle_3:    element_class_triangle_3_inline_impl.hh:79-102
//   quadrangle_4:  element_class_quadrangle_4_inline_impl.hh
//   tetrahedron_4: element_class_tetrahedron_4_inline_impl.hh:86-125
//   hexahedron_8:  element_class_hexahedron_8_inline_impl.hh:133-214
void dnds(int type, const double* c, double* D) {
  auto set = [&](int nd, int i, int a, double v) { D[i + a * nd] = v; };
  switch (type) {
  case TRIANGLE_3:
    set(2, 0, 0, -1.); set(2, 0, 1, 1.); set(2, 0, 2, 0.);
    set(2, 1, 0, -1.); set(2, 1, 1, 0.); set(2, 1, 2, 1.);
    break;
  case QUADRANGLE_4:
    set(2, 0, 0, -.25 * (1 - c[1])); set(2, 0, 1, .25 * (1 - c[1]));
    set(2, 0, 2, .25 * (1 + c[1]));  set(2, 0, 3, -.25 * (1 + c[1]));
    set(2, 1, 0, -.25 * (1 - c[0])); set(2, 1, 1, -.25 * (1 + c[0]));
    set(2, 1, 2, .25 * (1 + c[0]));  set(2, 1, 3, .25 * (1 - c[0]));
    break;
  case TETRAHEDRON_4:
    set(3, 0, 0, -1.); set(3, 0, 1, 1.); set(3, 0, 2, 0.); set(3, 0, 3, 0.);
    set(3, 1, 0, -1.); set(3, 1, 1, 0.); set(3, 1, 2, 1.); set(3, 1, 3, 0.);
    set(3, 2, 0, -1.); set(3, 2, 1, 0.); set(3, 2, 2, 0.); set(3, 2, 3, 1.);
    break;
  case HEXAHEDRON_8:
    set(3, 0, 0, -.125 * (1 - c[1]) * (1 - c[2]));
    set(3, 0, 1, .125 * (1 - c[1]) * (1 - c[2]));
    set(3, 0, 2, .125 * (1 + c[1]) * (1 - c[2]));
    set(3, 0, 3, -.125 * (1 + c[1]) * (1 - c[2]));
    set(3, 0, 4, -.125 * (1 - c[1]) * (1 + c[2]));
    set(3, 0, 5, .125 * (1 - c[1]) * (1 + c[2]));
    set(3, 0, 6, .125 * (1 + c[1]) * (1 + c[2]));
    set(3, 0, 7, -.125 * (1 + c[1]) * (1 + c[2]));
    set(3, 1, 0, -.125 * (1 - c[0]) * (1 - c[2]));
    set(3, 1, 1, -.125 * (1 + c[0]) * (1 - c[2]));
    set(3, 1, 2, .125 * (1 + c[0]) * (1 - c[2]));
    set(3, 1, 3, .125 * (1 - c[0]) * (1 - c[2]));
    set(3, 1, 4, -.125 * (1 - c[0]) * (1 + c[2]));
    set(3, 1, 5, -.125 * (1 + c[0]) * (1 + c[2]));
    set(3, 1, 6, .125 * (1 + c[0]) * (1 + c[2]));
    set(3, 1, 7, .125 * (1 - c[0]) * (1 + c[2]));
    set(3, 2, 0, -.125 * (1 - c[0]) * (1 - c[1]));
    set(3, 2, 1, -.125 * (1 + c[0]) * (1 - c[1]));
    set(3, 2, 2, -.125 * (1 + c[0]) * (1 + c[1]));
    set(3, 2, 3, -.125 * (1 - c[0]) * (1 + c[1]));
    set(3, 2, 4, .125 * (1 - c[0]) * (1 - c[1]));
    set(3, 2, 5, .125 * (1 + c[0]) * (1 - c[1]));
    set(3, 2, 6, .125 * (1 + c[0]) * (1 + c[1]));
    set(3, 2, 7, .125 * (1 - c[0]) * (1 + c[1]));
    break;
  }
}

// Small-matrix helpers, the non-BLAS loops of src/common/aka_math_tmpl.hh.
// All matrices column-major.  C(im x in) = A(im x ik) * B(ik x in)   (:98-122)
inline void matrix_matrix(int im, int in, int ik, const double* A, const double* B, double* C) {
  for (int j = 0; j < in; ++j)
    for (int i = 0; i < im; ++i) {
      double s = 0.;
      for (int l = 0; l < ik; ++l) s += A[i + l * im] * B[l + j * ik];
      C[i + j * im] = s * 1.;
    }
}
// C(im x in) = A(im x ik) * B(in x ik)^T   (:152-176)
inline void matrix_matrixt(int im, int in, int ik, const double* A, const double* B, double* C) {
  for (int j = 0; j < in; ++j)
    for (int i = 0; i < im; ++i) {
      double s = 0.;
      for (int l = 0; l < ik; ++l) s += A[i + l * im] * B[j + l * in];
      C[i + j * im] = s * 1.;
    }
}
// C(im x in) = A(ik x im)^T * B(ik x in)   (:125-149)
inline void matrixt_matrix(int im, int in, int ik, const double* A, const double* B, double* C) {
  for (int j = 0; j < in; ++j)
    for (int i = 0; i < im; ++i) {
      double s = 0.;
      for (int l = 0; l < ik; ++l) s += A[l + i * ik] * B[l + j * ik];
      C[i + j * im] = s * 1.;
    }
}
// det2 / det3 / inv2 / inv3: aka_math_tmpl.hh:330-339, 427-449
inline double det2(const double* m) { return m[0] * m[3] - m[1] * m[2]; }
inline double det3(const double* m) {
  return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
         m[6] * (m[1] * m[5] - m[4] * m[2]);
}
inline void inv2(const double* m, double* r) {
  double d = det2(m);
  r[0] = m[3] / d; r[1] = -m[1] / d; r[2] = -m[2] / d; r[3] = m[0] / d;
}
inline void inv3(const double* m, double* r) {
  double d = det3(m);
  r[0] = (m[4] * m[8] - m[7] * m[5]) / d;
  r[1] = (m[2] * m[7] - m[8] * m[1]) / d;
  r[2] = (m[1] * m[5] - m[4] * m[2]) / d;
  r[3] = (m[5] * m[6] - m[8] * m[3]) / d;
  r[4] = (m[0] * m[8] - m[6] * m[2]) / d;
  r[5] = (m[2] * m[3] - m[5] * m[0]) / d;
  r[6] = (m[3] * m[7] - m[6] * m[4]) / d;
  r[7] = (m[1] * m[6] - m[7] * m[0]) / d;
  r[8] = (m[0] * m[4] - m[3] * m[1]) / d;
}
inline double detn(int n, const double* m) { return n == 2 ? det2(m) : det3(m); }
inline void invn(int n, const double* m, double* r) { if (n == 2) inv2(m, r); else inv3(m, r); }

// Inradius used by the stable time step (aka_math_tmpl.hh:563-679,
// element_class_*::getInradius).
inline double dist2d(const double* x, const double* y) {
  return std::sqrt((y[0] - x[0]) * (y[0] - x[0]) + (y[1] - x[1]) * (y[1] - x[1]));
}
inline double dist3d(const double* x, const double* y) {
  return std::sqrt((y[0] - x[0]) * (y[0] - x[0]) + (y[1] - x[1]) * (y[1] - x[1]) +
                   (y[2] - x[2]) * (y[2] - x[2]));
}
double triangle_inradius(const double* c1, const double* c2, const double* c3) {
  double a = dist2d(c1, c2), b = dist2d(c2, c3), c = dist2d(c1, c3);
  double s = (a + b + c) * 0.5;
  return std::sqrt((s - a) * (s - b) * (s - c) / s);
}
double tet_volume(const double* c1, const double* c2, const double* c3, const double* c4) {
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
  return vol;
}
double tet_inradius(const double* c1, const double* c2, const double* c3, const double* c4) {
  double l12 = dist3d(c1, c2), l13 = dist3d(c1, c3), l14 = dist3d(c1, c4);
  double l23 = dist3d(c2, c3), l24 = dist3d(c2, c4), l34 = dist3d(c3, c4);
  double s1 = (l12 + l23 + l13) * 0.5; s1 = std::sqrt(s1 * (s1 - l12) * (s1 - l23) * (s1 - l13));
  double s2 = (l12 + l24 + l14) * 0.5; s2 = std::sqrt(s2 * (s2 - l12) * (s2 - l24) * (s2 - l14));
  double s3 = (l23 + l34 + l24) * 0.5; s3 = std::sqrt(s3 * (s3 - l23) * (s3 - l34) * (s3 - l24));
  double s4 = (l13 + l34 + l14) * 0.5; s4 = std::sqrt(s4 * (s4 - l13) * (s4 - l34) * (s4 - l14));
  double volume = tet_volume(c1, c2, c3, c4);
  return 3 * volume / (s1 + s2 + s3 + s4);
}
// element_class_quadrangle_4_inline_impl.hh:131-154: smallest of the 4 edges.
double quad_inradius(const double* x /* 4 points, d=2 */) {
  double a = dist2d(x + 0, x + 2), b = dist2d(x + 2, x + 4);
  double c = dist2d(x + 4, x + 6), d = dist2d(x + 6, x + 0);
  return std::min({a, b, c, d});
}
double hex_inradius(const double* x /* 8 points, d=3 */) {
  auto P = [&](int i) { return x + 3 * i; };
  double e[12] = {dist3d(P(0), P(1)), dist3d(P(1), P(2)), dist3d(P(2), P(3)),
                  dist3d(P(3), P(0)), dist3d(P(0), P(4)), dist3d(P(1), P(5)),
                  dist3d(P(2), P(6)), dist3d(P(3), P(7)), dist3d(P(4), P(5)),
                  dist3d(P(5), P(6)), dist3d(P(6), P(7)), dist3d(P(7), P(4))};
  double p = e[0];
  for (int i = 1; i < 12; ++i) p = std::min(p, e[i]);
  return p;
}

/* ------------------------------------------------------------------------ */
/* Central difference = NewmarkBeta(alpha=0, beta=1/2), k = h = 0             */
/* src/model/common/integration_scheme/newmark-beta.cc:60-89, 176-210         */
/* ------------------------------------------------------------------------ */
const double NB_ALPHA = 0., NB_BETA = .5, NB_K = 0., NB_H = 0.;

void cd_predictor(size_t ndof, double dt, double* u, double* v, double* a, const uint8_t* blocked) {
  for (size_t d = 0; d < ndof; ++d) {
    if (!blocked[d]) {
      double dt_a_n = dt * a[d];
      u[d] += (1 - NB_K * NB_ALPHA) * dt * v[d] + (.5 - NB_H * NB_ALPHA * NB_BETA) * dt * dt_a_n;
      v[d] = (1 - NB_K) * v[d] + (1 - NB_H * NB_BETA) * dt_a_n;
      a[d] = (1 - NB_H) * a[d];
    }
  }
}

// Residual (TimeStepSolver::assembleResidual -> clearResidual;
// SMM::assembleResidual solid_mechanics_model.cc:279-293: r += 1 f_ext, r += 1 f_int;
// IntegrationScheme2ndOrder::assembleResidual + DOFManagerDefault::
// assembleLumpedMatMulVectToResidual dof_manager_default.cc:271-287: r += M (-1 a)),
// lumped solve (non_linear_solver_lumped.cc:85-98, alpha = 1) and
// allCorrector<_acceleration> (newmark-beta.cc:176-210; c = 1, d = beta dt,
// e = alpha beta dt^2).
void cd_residual_solve_correct(size_t ndof, double dt, double* u, double* v, double* a,
                               const uint8_t* blocked, const double* M, const double* fext,
                               const double* fint, double fint_scale, double* delta) {
  const double c = 1., dcoef = NB_BETA * dt, e = NB_ALPHA * NB_BETA * dt * dt;
  const double b_a2x = 1.;
  for (size_t d = 0; d < ndof; ++d) {
    double r = 0.;
    r += 1. * fext[d];
    r += fint_scale * fint[d];
    double cache = 0.;
    cache += -1. * a[d];
    r += M[d] * cache;
    if (!blocked[d]) delta[d] = b_a2x * (r / M[d]);
  }
  for (size_t d = 0; d < ndof; ++d) {
    if (!blocked[d]) {
      u[d] += e * delta[d];
      v[d] += dcoef * delta[d];
      a[d] += c * delta[d];
    }
  }
}

struct Material {
  int kind = ELASTIC;
  double E = 0, nu = 0, rho = 0;
  int plane_stress = 0;
  double lambda = 0, mu = 0;
  double Sd = 0;
  std::vector<uint32_t> filter; // element ids in ascending order
};

struct Model {
  int type, d, n, nq;
  ElemInfo info;
  size_t N, E;
  std::vector<double> X;
  std::vector<uint32_t> conn;
  std::vector<Material> mats;
  std::vector<int> elem_mat;
  std::vector<double> B, JxW, Nq; // (E*Q)*(d*n), E*Q, Q*n
  std::vector<double> u, v, a, fext, fint, mass, prev, inc, delta;
  std::vector<uint8_t> blocked;
  std::vector<double> gradu, stress; // (E*Q)*d*d
  std::vector<double> damage, Yd;    // E*Q
  double dt = 0.;
  int nthreads = 1;
  std::vector<std::vector<double>> part;  // per-thread force arrays (CPU baseline)
  int initialized = 0;
  std::string err;
};

void material_update_params(Material& m, int dim) {
  // MaterialElastic<dim>::updateInternalParameters (material_elastic.cc:83-107),
  // MaterialNeohookean::updateInternalParameters (material_neohookean.cc:83-89)
  m.lambda = m.nu * m.E / ((1 + m.nu) * (1 - 2 * m.nu));
  m.mu = m.E / (2 * (1 + m.nu));
  if (dim == 2 && m.plane_stress && m.kind == ELASTIC)
    m.lambda = m.nu * m.E / ((1 + m.nu) * (1 - m.nu));
}

// One quadrature point: elastic stress, material_elastic_inline_impl.hh:41-55
inline void elastic_stress(int d, const Material& m, const double* G, double* S) {
  double trace = 0.;
  for (int i = 0; i < d; ++i) trace += G[i + i * d];
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j)
      S[i + j * d] = double(i == j) * m.lambda * trace + m.mu * (G[i + j * d] + G[j + i * d]) +
                     double(i == j) * 0.;
}

// neo-Hookean second Piola-Kirchhoff stress, material_neohookean_inline_impl.hh:46-65
// (C33 = 1: plane strain / 3D).  Returns J.
inline double neohookean_stress(int d, const Material& m, const double* G, double* S) {
  double F[9] = {0}, C[9] = {0}, Ci[9] = {0};
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) F[i + j * d] = double(i == j);
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) F[i + j * d] += G[i + j * d];
  matrixt_matrix(d, d, d, F, F, C);
  double J = detn(d, F) * std::sqrt(1.);
  invn(d, C, Ci);
  double lj = std::log(J);
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j)
      S[i + j * d] = double(i == j) * m.mu + (m.lambda * lj - m.mu) * Ci[i + j * d];
  return J;
}

// neo-Hookean strain energy density, material_neohookean_inline_impl.hh:145-158
inline double neohookean_energy(int d, const Material& m, const double* G) {
  double F[9] = {0}, C[9] = {0};
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) F[i + j * d] = double(i == j);
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) F[i + j * d] += G[i + j * d];
  matrixt_matrix(d, d, d, F, F, C);
  double J = detn(d, F);
  double trC = 0.;
  for (int i = 0; i < d; ++i) trC += C[i + i * d];
  return 0.5 * m.lambda * std::pow(std::log(J), 2.) + m.mu * (-std::log(J) + 0.5 * (trC - d));
}

// Neo-Hookean + isotropic damage composition (no counterpart in the reference;
// the damage update restates MaterialMarigo::computeDamageAndStressOnQuad,
// material_marigo_inline_impl.hh:67-76, with the driving force Y taken as the
// undamaged neo-Hookean energy density W(F)):
//   Fd = Y - Yd - Sd d;  if Fd > 0: d = (Y - Yd)/Sd;  d = min(d, 1);  S *= (1 - d)
inline void neohookean_damage_stress(int d, const Material& m, const double* G, double* S,
                                     double& dam, double Ydq) {
  neohookean_stress(d, m, G, S);
  double Y = neohookean_energy(d, m, G);
  double Fd = Y - Ydq - m.Sd * dam;
  if (Fd > 0) dam = (Y - Ydq) / m.Sd;
  dam = std::min(dam, 1.);
  for (int k = 0; k < d * d; ++k) S[k] *= 1 - dam;
}

// Finite-strain residual at one point: r = (B_L0 + B_L2(gradu))^T S_voigt
// material.cc:689-706, aka_voigthelper_tmpl.hh:104-143 (sym Voigt B),
// :177-203 (3D B_L2), :207-226 (2D B_L2), Voigt order aka_voigthelper.cc:52-61.
void finite_strain_residual(int d, int n, const double* Bq, const double* G, const double* S,
                            double* r) {
  const int vs = d == 2 ? 3 : 6;
  const int cols = d * n;
  double Bv[6 * 24], B2[6 * 24];
  std::fill(Bv, Bv + vs * cols, 0.);
  std::fill(B2, B2 + vs * cols, 0.);
  auto BV = [&](int I, int c) -> double& { return Bv[I + c * vs]; };
  auto BL = [&](int I, int c) -> double& { return B2[I + c * vs]; };
  auto Bm = [&](int i, int a) { return Bq[i + a * d]; };
  auto Gm = [&](int i, int j) { return G[i + j * d]; };
  for (int i = 0; i < d; ++i)
    for (int a = 0; a < n; ++a) BV(i, i + a * d) = Bm(i, a);
  if (d == 2) {
    for (int a = 0; a < n; ++a) {
      BV(2, 1 + a * 2) = Bm(0, a);
      BV(2, 0 + a * 2) = Bm(1, a);
    }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < n; ++j)
        for (int k = 0; k < 2; ++k) BL(i, j * 2 + k) = Gm(k, i) * Bm(i, j);
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < 2; ++k)
        for (int mm = 0; mm < 2; ++mm) BL(2, j * 2 + k) += Gm(k, mm) * Bm(1 - mm, j);
  } else {
    for (int a = 0; a < n; ++a) {
      double dndx = Bm(0, a), dndy = Bm(1, a), dndz = Bm(2, a);
      BV(3, 1 + a * 3) = dndz; BV(3, 2 + a * 3) = dndy;
      BV(4, 0 + a * 3) = dndz; BV(4, 2 + a * 3) = dndx;
      BV(5, 0 + a * 3) = dndy; BV(5, 1 + a * 3) = dndx;
    }
    for (int I = 0; I < 3; ++I)
      for (int a = 0; a < n; ++a)
        for (int i = 0; i < 3; ++i) BL(I, a * 3 + i) = Gm(i, I) * Bm(I, a);
    for (int Iv = 3; Iv < 6; ++Iv)
      for (int a = 0; a < n; ++a)
        for (int k = 0; k < 3; ++k) {
          int aux = Iv - 3;
          for (int mm = 0; mm < 3; ++mm)
            if (mm != aux) BL(Iv, a * 3 + k) += Gm(k, mm) * Bm(3 - mm - aux, a);
        }
  }
  for (int k = 0; k < vs * cols; ++k) Bv[k] += B2[k];
  double Sv[6];
  if (d == 2) { Sv[0] = S[0]; Sv[1] = S[3]; Sv[2] = S[0 + 1 * 2]; }
  else {
    Sv[0] = S[0]; Sv[1] = S[4]; Sv[2] = S[8];
    Sv[3] = S[1 + 2 * 3]; Sv[4] = S[0 + 2 * 3]; Sv[5] = S[0 + 1 * 3];
  }
  matrixt_matrix(cols, 1, vs, Bv, Sv, r);
}

// Element force vector fe (d*n, layout [a*d+i]) of element e, with gradients and
// stresses stored per point.  Restates Material::computeAllStresses
// (material.cc:254-279) + ShapeFunctions::gradientElementalFieldOnIntegrationPoints
// (shape_functions_inline_impl.hh:383-390) + computeStress + computeBtD
// (shape_lagrange_inline_impl.hh:411-422) / finite-strain residual
// (material.cc:637-718) + IntegratorGauss::integrate (integrator_gauss_inline_impl.hh:426-455).
void element_force(Model& M, const Material& mat, size_t e, double* fe) {
  const int d = M.d, n = M.n, Q = M.nq, dn = d * n;
  double ue[24];
  for (int a = 0; a < n; ++a)
    for (int i = 0; i < d; ++i) ue[i + a * d] = M.u[size_t(M.conn[e * n + a]) * d + i];
  double Rq[8][24];
  for (int q = 0; q < Q; ++q) {
    size_t eq = e * Q + q;
    const double* Bq = &M.B[eq * dn];
    double* G = &M.gradu[eq * d * d];
    double* S = &M.stress[eq * d * d];
    matrix_matrixt(d, d, n, ue, Bq, G);
    for (int k = 0; k < d * d; ++k) G[k] -= 0.; // gradu -= eigengradu (zero)
    if (mat.kind == ELASTIC) {
      elastic_stress(d, mat, G, S);
      matrix_matrix(d, n, d, S, Bq, Rq[q]);
    } else {
      if (mat.kind == NEOHOOKEAN) neohookean_stress(d, mat, G, S);
      else neohookean_damage_stress(d, mat, G, S, M.damage[eq], M.Yd[eq]);
      finite_strain_residual(d, n, Bq, G, S, Rq[q]);
    }
  }
  for (int k = 0; k < dn; ++k) {
    double s = 0.;
    for (int q = 0; q < Q; ++q) s += Rq[q][k] * M.JxW[e * Q + q];
    fe[k] = s * 1.;
  }
}

// SMM::assembleInternalForces (solid_mechanics_model.cc:375-414) with the
// element-order scatter of DOFManager::assembleElementalArrayLocalArray
// (dof_manager.cc:83-137, scale -1, Vector::aXplusY).
void assemble_internal_forces(Model& M) {
  const int d = M.d, n = M.n, dn = d * n;
  std::fill(M.fint.begin(), M.fint.end(), 0.);
  if (M.nthreads <= 1) {
    double fe[24];
    for (auto& mat : M.mats)
      for (uint32_t e : mat.filter) {
        element_force(M, mat, e, fe);
        for (int a = 0; a < n; ++a) {
          double* dst = &M.fint[size_t(M.conn[size_t(e) * n + a]) * d];
          for (int i = 0; i < d; ++i) dst[i] += -1. * fe[i + a * d];
        }
      }
    return;
  }
  // Multi-threaded variant (CPU baseline only): contiguous element chunks per
  // thread into private force arrays (allocated once per model), then every
  // thread adds the T partial arrays over its own range of DOFs, in thread
  // order.  This mirrors an MPI domain decomposition; it is not
  // bit-identical to the single-thread order.
  const int T = M.nthreads;
  const size_t nd = M.fint.size();
  if (M.part.size() != size_t(T) || (T && M.part[0].size() != nd))
    M.part.assign(T, std::vector<double>(nd, 0.));
  auto run = [&](auto&& body) {
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) th.emplace_back(body, t);
    for (auto& x : th) x.join();
  };
  run([&](int t) {
    std::fill(M.part[t].begin(), M.part[t].end(), 0.);
    double fe[24];
    for (auto& mat : M.mats) {
      const size_t ne = mat.filter.size(), b = ne * t / T, en = ne * (t + 1) / T;
      for (size_t k = b; k < en; ++k) {
        uint32_t e = mat.filter[k];
        element_force(M, mat, e, fe);
        for (int a = 0; a < n; ++a) {
          double* dst = &M.part[t][size_t(M.conn[size_t(e) * n + a]) * d];
          for (int i = 0; i < d; ++i) dst[i] += -1. * fe[i + a * d];
        }
      }
    }
  });
  run([&](int t) {
    const size_t b = nd * t / T, en = nd * (t + 1) / T;
    for (int s = 0; s < T; ++s) {
      const double* src = M.part[s].data();
      for (size_t k = b; k < en; ++k) M.fint[k] += src[k];
    }
  });
  (void)dn;
}

int precompute(Model& M) {
  const int d = M.d, n = M.n, Q = M.nq, dn = d * n;
  const ElemInfo& I = M.info;
  M.B.assign(M.E * Q * dn, 0.);
  M.JxW.assign(M.E * Q, 0.);
  M.Nq.assign(size_t(Q) * n, 0.);
  double D[8][3 * 8];
  for (int q = 0; q < Q; ++q) {
    dnds(M.type, I.qp[q], D[q]);
    shapes(M.type, I.qp[q], &M.Nq[size_t(q) * n]);
  }
  // ShapeLagrange::computeShapeDerivativesOnIntegrationPoints
  // (shape_lagrange_inline_impl.hh:204-252 -> :62-82; element_class_tmpl.hh:282-355)
  // and IntegratorGauss::computeJacobiansOnIntegrationPoints (+ weights)
  // (integrator_gauss_inline_impl.hh:208-252, 396-423, check :182-205).
  for (size_t e = 0; e < M.E; ++e) {
    double xe[24];
    for (int a = 0; a < n; ++a)
      for (int i = 0; i < d; ++i) xe[i + a * d] = M.X[size_t(M.conn[e * n + a]) * d + i];
    for (int q = 0; q < Q; ++q) {
      double J[9], Ji[9];
      matrix_matrixt(d, d, n, D[q], xe, J);
      invn(d, J, Ji);
      matrix_matrix(d, n, d, Ji, D[q], &M.B[(e * Q + q) * dn]);
      double detJ = detn(d, J);
      M.JxW[e * Q + q] = detJ * I.w[q];
      if (M.JxW[e * Q + q] < 0) {
        M.err = "Negative jacobian computed, possible problem in the element node ordering";
        return -2;
      }
    }
  }
  // Lumped mass by row sum: SMM::assembleMassLumped (solid_mechanics_model_mass.cc:58-102)
  // -> assembleLumpedRowSum (fe_engine_template_tmpl_field.hh:165-199), computeNtb
  // (shape_lagrange_inline_impl.hh:495-533), integrate, element-order assembly.
  std::vector<double> local(M.N * d, 0.);
  for (size_t e = 0; e < M.E; ++e) {
    double rho = M.mats[M.elem_mat[e]].rho;
    double integ[24];
    for (int k = 0; k < dn; ++k) {
      double s = 0.;
      for (int q = 0; q < Q; ++q) {
        int i = k % d, a = k / d;
        double ntb = 0.;
        ntb += rho * M.Nq[size_t(q) * n + a];
        (void)i;
        ntb *= 1.;
        s += ntb * M.JxW[e * Q + q];
      }
      integ[k] = s * 1.;
    }
    for (int a = 0; a < n; ++a)
      for (int i = 0; i < d; ++i) local[size_t(M.conn[e * n + a]) * d + i] += 1. * integ[i + a * d];
  }
  for (size_t k = 0; k < M.mass.size(); ++k) {
    double m = 0.;
    m += 1. * local[k];
    M.mass[k] = m;
  }
  M.initialized = 1;
  return 0;
}

} // namespace

/* ======================================================================== */
/* C ABI (ctypes)                                                            */
/* ======================================================================== */
extern "C" {

void* orc_create(int type, int dim, int64_t nb_nodes, const double* X, int64_t nb_elem,
                 const uint32_t* conn) {
  ElemInfo I = elem_info(type);
  if (I.dim == 0 || I.dim != dim || nb_nodes <= 0 || nb_elem < 0) return nullptr;
  auto* M = new Model();
  M->type = type; M->d = dim; M->n = I.nnodes; M->nq = I.nqp; M->info = I;
  M->N = size_t(nb_nodes); M->E = size_t(nb_elem);
  M->X.assign(X, X + M->N * dim);
  M->conn.assign(conn, conn + M->E * M->n);
  size_t nd = M->N * dim;
  M->u.assign(nd, 0.); M->v.assign(nd, 0.); M->a.assign(nd, 0.);
  M->fext.assign(nd, 0.); M->fint.assign(nd, 0.); M->mass.assign(nd, 0.);
  M->prev.assign(nd, 0.); M->inc.assign(nd, 0.); M->delta.assign(nd, 0.);
  M->blocked.assign(nd, 0);
  size_t eq = M->E * M->nq;
  M->gradu.assign(eq * dim * dim, 0.); M->stress.assign(eq * dim * dim, 0.);
  M->damage.assign(eq, 0.); M->Yd.assign(eq, 0.);
  M->elem_mat.assign(M->E, 0);
  return M;
}

void orc_destroy(void* h) { delete static_cast<Model*>(h); }

const char* orc_last_error(void* h) { return static_cast<Model*>(h)->err.c_str(); }

int orc_add_material(void* h, int kind, double E, double nu, double rho, int plane_stress) {
  auto* M = static_cast<Model*>(h);
  Material m;
  m.kind = kind; m.E = E; m.nu = nu; m.rho = rho; m.plane_stress = plane_stress;
  material_update_params(m, M->d);
  M->mats.push_back(m);
  return int(M->mats.size()) - 1;
}

int orc_set_damage(void* h, int mat, double Sd, const double* Yd_per_point) {
  auto* M = static_cast<Model*>(h);
  if (mat < 0 || mat >= int(M->mats.size())) return -1;
  M->mats[mat].Sd = Sd;
  if (Yd_per_point) std::copy(Yd_per_point, Yd_per_point + M->E * M->nq, M->Yd.begin());
  return 0;
}

int orc_assign_materials(void* h, const int32_t* elem_mat) {
  auto* M = static_cast<Model*>(h);
  for (auto& m : M->mats) m.filter.clear();
  for (size_t e = 0; e < M->E; ++e) {
    int k = elem_mat ? elem_mat[e] : 0;
    if (k < 0 || k >= int(M->mats.size())) return -1;
    M->elem_mat[e] = k;
    M->mats[k].filter.push_back(uint32_t(e));
  }
  return 0;
}

int orc_init(void* h) {
  auto* M = static_cast<Model*>(h);
  if (M->mats.empty()) { M->err = "no material"; return -1; }
  bool any = false;
  for (auto& m : M->mats) any |= !m.filter.empty();
  if (!any) orc_assign_materials(h, nullptr);
  return precompute(*M);
}

void orc_set_threads(void* h, int n) { static_cast<Model*>(h)->nthreads = std::max(1, n); }

// field ids: 0 u, 1 v, 2 a, 3 f_ext, 4 f_int, 5 mass, 6 B, 7 JxW, 8 gradu, 9 stress,
// 10 damage, 11 Yd, 12 previous u, 13 increment, 14 X
double* orc_field(void* h, int which, int64_t* size) {
  auto* M = static_cast<Model*>(h);
  std::vector<double>* f = nullptr;
  switch (which) {
  case 0: f = &M->u; break;   case 1: f = &M->v; break;    case 2: f = &M->a; break;
  case 3: f = &M->fext; break; case 4: f = &M->fint; break; case 5: f = &M->mass; break;
  case 6: f = &M->B; break;   case 7: f = &M->JxW; break;  case 8: f = &M->gradu; break;
  case 9: f = &M->stress; break; case 10: f = &M->damage; break; case 11: f = &M->Yd; break;
  case 12: f = &M->prev; break; case 13: f = &M->inc; break; case 14: f = &M->X; break;
  default: if (size) *size = 0; return nullptr;
  }
  if (size) *size = int64_t(f->size());
  return f->data();
}

uint8_t* orc_blocked(void* h) { return static_cast<Model*>(h)->blocked.data(); }

void orc_set_time_step(void* h, double dt) { static_cast<Model*>(h)->dt = dt; }

int orc_assemble_internal_forces(void* h) {
  auto* M = static_cast<Model*>(h);
  if (!M->initialized) return -1;
  assemble_internal_forces(*M);
  return 0;
}

// One SolidMechanicsModel::solveStep() with the explicit lumped-mass solver,
// repeated nsteps times: NonLinearSolverLumped::solve (non_linear_solver_lumped.cc:58-82)
// -> TimeStepSolverDefault::predictor (savePreviousDOFs + NewmarkBeta::predictor,
// time_step_solver_default.cc:172-186) -> assembleResidual -> solveLumped ->
// corrector (+ increment = u - previous, :189-224).
int orc_solve_step(void* h, int nsteps) {
  auto* M = static_cast<Model*>(h);
  if (!M->initialized) return -1;
  const size_t nd = M->N * M->d;
  for (int s = 0; s < nsteps; ++s) {
    std::copy(M->u.begin(), M->u.end(), M->prev.begin());
    cd_predictor(nd, M->dt, M->u.data(), M->v.data(), M->a.data(), M->blocked.data());
    assemble_internal_forces(*M);
    cd_residual_solve_correct(nd, M->dt, M->u.data(), M->v.data(), M->a.data(), M->blocked.data(),
                              M->mass.data(), M->fext.data(), M->fint.data(), 1.,
                              M->delta.data());
    for (size_t k = 0; k < nd; ++k) { M->inc[k] = M->u[k]; M->inc[k] -= M->prev[k]; }
  }
  return 0;
}

// SMM::getStableTimeStep (solid_mechanics_model.cc:482-533) on X + u.
double orc_stable_dt(void* h) {
  auto* M = static_cast<Model*>(h);
  const int d = M->d, n = M->n;
  std::vector<double> cur(M->X);
  for (size_t k = 0; k < cur.size(); ++k) cur[k] += M->u[k];
  double min_dt = std::numeric_limits<double>::max();
  for (size_t e = 0; e < M->E; ++e) {
    double x[24];
    for (int a = 0; a < n; ++a)
      for (int i = 0; i < d; ++i) x[i + a * d] = cur[size_t(M->conn[e * n + a]) * d + i];
    double hh = 0.;
    switch (M->type) {
    case TRIANGLE_3: hh = 2. * triangle_inradius(x, x + 2, x + 4); break;
    case QUADRANGLE_4: hh = quad_inradius(x); break;
    case TETRAHEDRON_4: hh = 2. * tet_inradius(x, x + 3, x + 6, x + 9); break;
    case HEXAHEDRON_8: hh = hex_inradius(x); break;
    }
    const Material& m = M->mats[M->elem_mat[e]];
    double c = std::sqrt((m.lambda + 2 * m.mu) / m.rho);
    min_dt = std::min(min_dt, hh / c);
  }
  return min_dt;
}

// kind 0: kinetic (solid_mechanics_model.cc:536-582)
// kind 1: potential (material.cc:743-757; elastic 0.5 sigma:gradu,
//         material_elastic_inline_impl.hh:104-108; neo-Hookean W(F))
// kind 2: external work (solid_mechanics_model.cc:615-662, dynamic: velocity * dt)
double orc_energy(void* h, int kind) {
  auto* M = static_cast<Model*>(h);
  const int d = M->d;
  if (kind == 0) {
    double ekin = 0.;
    for (size_t nn = 0; nn < M->N; ++nn) {
      double mv2 = 0.;
      for (int i = 0; i < d; ++i) {
        double m = M->mass[nn * d + i], v = M->v[nn * d + i];
        if (m > std::numeric_limits<double>::epsilon()) mv2 += v * v * m;
      }
      ekin += mv2;
    }
    return ekin * .5;
  }
  if (kind == 1) {
    // stresses/gradients on the current displacement
    double fe[24];
    const int Q = M->nq;
    double epot = 0.;
    for (auto& mat : M->mats) {
      std::vector<double> dam_save;
      if (mat.kind == NEOHOOKEAN_DAMAGE) dam_save = M->damage;
      double part = 0.;
      for (uint32_t e : mat.filter) {
        element_force(*M, mat, e, fe);
        double s_e = 0.;
        for (int q = 0; q < Q; ++q) {
          size_t eq = size_t(e) * Q + q;
          const double* G = &M->gradu[eq * d * d];
          const double* S = &M->stress[eq * d * d];
          double w;
          if (mat.kind == ELASTIC) {
            double dd = 0.;
            for (int k = 0; k < d * d; ++k) dd += S[k] * G[k];
            w = .5 * dd;
          } else {
            w = neohookean_energy(d, mat, G);
            if (mat.kind == NEOHOOKEAN_DAMAGE) w *= (1 - M->damage[eq]);
          }
          s_e += w * M->JxW[eq];
        }
        part += s_e;
      }
      if (mat.kind == NEOHOOKEAN_DAMAGE) M->damage = dam_save;
      epot += part;
    }
    return epot;
  }
  if (kind == 2) {
    double work = 0.;
    for (size_t k = 0; k < M->N * d; ++k) {
      if (M->blocked[k]) work -= M->fint[k] * M->v[k];
      else work += M->fext[k] * M->v[k];
    }
    return work * M->dt;
  }
  return 0.;
}

// Material-point stress evaluation for unit tests (no mesh): kind as above,
// G column-major d x d.  Returns J for finite strain, 0 otherwise.
double orc_point_stress(int kind, int dim, double E, double nu, int plane_stress, double Sd,
                        double Ydq, double* damage_inout, const double* G, double* S) {
  Material m;
  m.kind = kind; m.E = E; m.nu = nu; m.rho = 1.; m.plane_stress = plane_stress; m.Sd = Sd;
  material_update_params(m, dim);
  if (kind == ELASTIC) { elastic_stress(dim, m, G, S); return 0.; }
  if (kind == NEOHOOKEAN) return neohookean_stress(dim, m, G, S);
  double dam = damage_inout ? *damage_inout : 0.;
  neohookean_damage_stress(dim, m, G, S, dam, Ydq);
  if (damage_inout) *damage_inout = dam;
  return 0.;
}

double orc_point_energy(int kind, int dim, double E, double nu, int plane_stress, const double* G) {
  Material m;
  m.kind = kind; m.E = E; m.nu = nu; m.rho = 1.; m.plane_stress = plane_stress;
  material_update_params(m, dim);
  if (kind == ELASTIC) {
    double S[9];
    elastic_stress(dim, m, G, S);
    double dd = 0.;
    for (int k = 0; k < dim * dim; ++k) dd += S[k] * G[k];
    return .5 * dd;
  }
  return neohookean_energy(dim, m, G);
}

// Finite-strain point residual (for invariant tests): B (d x n col-major)
void orc_point_finite_residual(int dim, int n, const double* B, const double* G, const double* S,
                               double* r) {
  finite_strain_residual(dim, n, B, G, S, r);
}

/* ------------------------------------------------------------------------ */
/* Known-answer driver: test_model_solver_dynamic.cc (EXPLICIT=true) with the */
/* 1D truss MyModel (test_model_solver_my_model.hh).  Writes the same table  */
/* the reference writes to stdout; the golden is                              */
/* test_model_solver_dynamic_explicit.verified.                               */
/* ------------------------------------------------------------------------ */
int orc_truss_dynamic_explicit(char* out, int64_t cap) {
  const size_t nb_nodes = 201, nel = nb_nodes - 1;
  const int max_steps = 400;
  const double time_step = 0.001, pulse_width = 0.2, A0 = 0.01;
  const double E = 1., Asec = 1., rho = 1.;
  std::vector<double> x(nb_nodes), u(nb_nodes, 0.), v(nb_nodes, 0.), a(nb_nodes, 0.);
  std::vector<double> fext(nb_nodes, 0.), fint(nb_nodes, 0.), M(nb_nodes, 0.), delta(nb_nodes, 0.);
  std::vector<uint8_t> blocked(nb_nodes, 0);
  std::vector<double> L(nel), strain(nel, 0.), stress(nel, 0.);
  for (size_t i = 0; i < nb_nodes; ++i) x[i] = i * (1. / (nb_nodes - 1));
  for (size_t e = 0; e < nel; ++e) L[e] = std::abs(x[e + 1] - x[e]);

  // Sinusoidal Dirichlet functor (test_model_solver_dynamic.cc:52-72)
  const double k = 2 * M_PI / pulse_width, cv = std::sqrt(E / rho);
  auto sinusoidal = [&](size_t n, double t) {
    double xx = x[n];
    v[n] = k * cv * A0 * std::sin(k * (xx - cv * t));
    u[n] = A0 * std::cos(k * (xx - cv * t));
  };
  // MyModel::assembleResidual (my_model.hh:258-300): f_int scatter with +-f_n,
  // r += f_ext, r += -1 f_int.
  auto internal = [&]() {
    std::fill(fint.begin(), fint.end(), 0.);
    for (size_t e = 0; e < nel; ++e) {
      double u1 = u[e], u2 = u[e + 1];
      strain[e] = (u2 - u1) / L[e];
      stress[e] = E * strain[e];
      double f_n = Asec * stress[e];
      double f0 = -f_n, f1 = f_n;
      fint[e] += 1. * f0;
      fint[e + 1] += 1. * f1;
    }
  };
  // MyModel::assembleLumpedMass (my_model.hh:110-138)
  {
    std::vector<double> Ml(nb_nodes, 0.);
    for (size_t e = 0; e < nel; ++e) {
      double Ll = std::abs(x[e + 1] - x[e]);
      double M_n = rho * Asec * Ll / 2;
      Ml[e] += 1. * M_n;
      Ml[e + 1] += 1. * M_n;
    }
    for (size_t i = 0; i < nb_nodes; ++i) { M[i] = 0.; M[i] += 1. * Ml[i]; }
  }
  for (size_t n = 0; n < nb_nodes; ++n) sinusoidal(n, 0.);
  blocked[0] = blocked[nb_nodes - 1] = 1;

  std::string s;
  char line[256];
  snprintf(line, sizeof line, "%14s,%14s,%14s,%14s,%14s,%14s,%14s\n", "time", "wext", "epot",
           "ekin", "total", "max_disp", "min_disp");
  s += line;
  double wext = 0.;
  internal();
  double epot = 0, ekin = 0, einit = ekin + epot, etot = ekin + epot - wext - einit;
  auto minmax = [&](double& mx, double& mn) {
    mx = 0.; mn = 0.;
    for (double d : u) { mx = std::max(mx, d); mn = std::min(mn, d); }
  };
  double mx, mn;
  minmax(mx, mn);
  auto row = [&](double t) {
    snprintf(line, sizeof line, "%14.6e,%14.6e,%14.6e,%14.6e,%14.6e,%14.6e,%14.6e\n", t, wext,
             epot, ekin, etot, mx, mn);
    s += line;
  };
  row(0.);
  for (int i = 1; i < max_steps + 1; ++i) {
    sinusoidal(0, time_step * (i - 1));
    sinusoidal(nb_nodes - 1, time_step * (i - 1));
    cd_predictor(nb_nodes, time_step, u.data(), v.data(), a.data(), blocked.data());
    internal();
    cd_residual_solve_correct(nb_nodes, time_step, u.data(), v.data(), a.data(), blocked.data(),
                              M.data(), fext.data(), fint.data(), -1., delta.data());
    // energies (my_model.hh:302-360)
    double res = 0;
    for (size_t e = 0; e < nel; ++e) res += strain[e] * stress[e] * Asec * L[e];
    epot = res / 2.;
    res = 0;
    for (size_t n = 0; n < nb_nodes; ++n) res += M[n] * v[n] * v[n];
    ekin = res / 2.;
    res = 0;
    for (size_t n = 0; n < nb_nodes; ++n) res += (blocked[n] ? -fint[n] : fext[n]) * v[n];
    wext += res * time_step;
    etot = ekin + epot - wext - einit;
    minmax(mx, mn);
    row(time_step * i);
  }
  if (int64_t(s.size()) + 1 > cap) return -int(s.size() + 1);
  std::memcpy(out, s.c_str(), s.size() + 1);
  return int(s.size());
}

/* ------------------------------------------------------------------------ */
/* FEEngine operators on a caller's field (fe_engine.hh:95-189,240-244).     */
/* filter: element ids (nullptr = every element, in order); per-point        */
/* inputs and outputs are in filter order.                                   */
/* ------------------------------------------------------------------------ */
namespace {
size_t filtered(const Model& M, const uint32_t* filter, int64_t nf, int64_t t) {
  return filter ? size_t(filter[t]) : size_t(t);
}
int64_t filter_count(const Model& M, const uint32_t* filter, int64_t nf) {
  return filter ? nf : int64_t(M.E);
}

// Math::reduce (aka_math.cc:225-250): sort ascending, then pairwise sums.
double math_reduce(std::vector<double> a) {
  size_t nb_values = a.size();
  if (nb_values == 0) return 0.;
  size_t half = nb_values >> 1;
  std::sort(a.begin(), a.end());
  while (half) {
    size_t remaining = nb_values - 2 * half;
    if (remaining) a[nb_values - 2] += a[nb_values - 1];
    for (size_t i = 0; i < half; ++i) a[i] = a[2 * i] + a[2 * i + 1];
    nb_values = half;
    half >>= 1;
  }
  return a[0];
}
}  // namespace

// ShapeLagrange::gradientOnIntegrationPoints (shape_lagrange_inline_impl.hh:353-379):
// extractNodalToElementField (fe_engine_inline_impl.hh:104-143) then, per point,
// nabla_u = u_e * B^T (shape_functions_inline_impl.hh:352-395): out (nf*Q, nb_dof*d).
int orc_fe_gradient(void* h, const double* u, int nb_dof, const uint32_t* filter, int64_t nf,
                    double* out) {
  auto* M = static_cast<Model*>(h);
  const int d = M->d, n = M->n, Q = M->nq, dn = d * n;
  if (nb_dof < 1 || nb_dof > 9) return -1;
  const int64_t cnt = filter_count(*M, filter, nf);
  double ue[9 * 8];
  for (int64_t t = 0; t < cnt; ++t) {
    const size_t e = filtered(*M, filter, nf, t);
    for (int a = 0; a < n; ++a)
      for (int i = 0; i < nb_dof; ++i)
        ue[i + a * nb_dof] = u[size_t(M->conn[e * n + a]) * nb_dof + i];
    for (int q = 0; q < Q; ++q)
      matrix_matrixt(nb_dof, d, n, ue, &M->B[(e * Q + q) * dn],
                     out + (size_t(t) * Q + q) * nb_dof * d);
  }
  return 0;
}

// ShapeFunctions::interpolateElementalFieldOnIntegrationPoints
// (shape_functions_inline_impl.hh:300-335): uq = u_e * N, out (nf*Q, nb_dof).
int orc_fe_interpolate(void* h, const double* u, int nb_dof, const uint32_t* filter, int64_t nf,
                       double* out) {
  auto* M = static_cast<Model*>(h);
  const int n = M->n, Q = M->nq;
  if (nb_dof < 1 || nb_dof > 9) return -1;
  const int64_t cnt = filter_count(*M, filter, nf);
  double ue[9 * 8], Nm[8 * 8], uq[9 * 8];
  for (int q = 0; q < Q; ++q)
    for (int a = 0; a < n; ++a) Nm[a + q * n] = M->Nq[size_t(q) * n + a];
  for (int64_t t = 0; t < cnt; ++t) {
    const size_t e = filtered(*M, filter, nf, t);
    for (int a = 0; a < n; ++a)
      for (int i = 0; i < nb_dof; ++i)
        ue[i + a * nb_dof] = u[size_t(M->conn[e * n + a]) * nb_dof + i];
    matrix_matrix(nb_dof, Q, n, ue, Nm, uq);
    std::memcpy(out + size_t(t) * Q * nb_dof, uq, sizeof(double) * Q * nb_dof);
  }
  return 0;
}

// ShapeLagrange::computeBtD (shape_lagrange_inline_impl.hh:382-423):
// Bt_D = D (rows x d) * B (d x n), out (nf*Q, rows*n).
int orc_fe_btd(void* h, const double* D, int rows, const uint32_t* filter, int64_t nf,
               double* out) {
  auto* M = static_cast<Model*>(h);
  const int d = M->d, n = M->n, Q = M->nq, dn = d * n;
  if (rows < 1 || rows > 9) return -1;
  const int64_t cnt = filter_count(*M, filter, nf);
  for (int64_t t = 0; t < cnt; ++t) {
    const size_t e = filtered(*M, filter, nf, t);
    for (int q = 0; q < Q; ++q) {
      const size_t r = size_t(t) * Q + q;
      matrix_matrix(rows, n, d, D + r * rows * d, &M->B[(e * Q + q) * dn], out + r * rows * n);
    }
  }
  return 0;
}

// ShapeLagrange::computeNtb (shape_lagrange_inline_impl.hh:495-533):
// Ntb = b (nb_dof x 1) * N (1 x n), out (nf*Q, nb_dof*n).
int orc_fe_ntb(void* h, const double* b, int nb_dof, const uint32_t* filter, int64_t nf,
               double* out) {
  auto* M = static_cast<Model*>(h);
  const int n = M->n, Q = M->nq;
  if (nb_dof < 1 || nb_dof > 24) return -1;
  const int64_t cnt = filter_count(*M, filter, nf);
  for (int64_t t = 0; t < cnt; ++t)
    for (int q = 0; q < Q; ++q) {
      const size_t r = size_t(t) * Q + q;
      matrix_matrix(nb_dof, n, 1, b + r * nb_dof, &M->Nq[size_t(q) * n], out + r * nb_dof * n);
    }
  return 0;
}

// IntegratorGauss::integrate (integrator_gauss_inline_impl.hh:426-483):
// intf = f (nb_dof x Q) * J (Q x 1), out (nf, nb_dof).
int orc_fe_integrate(void* h, const double* f, int nb_dof, const uint32_t* filter, int64_t nf,
                     double* out) {
  auto* M = static_cast<Model*>(h);
  const int Q = M->nq;
  if (nb_dof < 1 || nb_dof > 192) return -1;
  const int64_t cnt = filter_count(*M, filter, nf);
  double Jq[8];
  for (int64_t t = 0; t < cnt; ++t) {
    const size_t e = filtered(*M, filter, nf, t);
    for (int q = 0; q < Q; ++q) Jq[q] = M->JxW[e * Q + q];
    matrix_matrix(nb_dof, 1, Q, f + size_t(t) * Q * nb_dof, Jq, out + size_t(t) * nb_dof);
  }
  return 0;
}

// IntegratorGauss::integrateOnIntegrationPoints (integrator_gauss_inline_impl.hh:545-597):
// out = f * J per point.
int orc_fe_integrate_on_points(void* h, const double* f, int nb_dof, const uint32_t* filter,
                               int64_t nf, double* out) {
  auto* M = static_cast<Model*>(h);
  const int Q = M->nq;
  if (nb_dof < 1 || nb_dof > 192) return -1;
  const int64_t cnt = filter_count(*M, filter, nf);
  for (int64_t t = 0; t < cnt; ++t) {
    const size_t e = filtered(*M, filter, nf, t);
    for (int q = 0; q < Q; ++q) {
      const size_t r = size_t(t) * Q + q;
      for (int k = 0; k < nb_dof; ++k) {
        double v = f[r * nb_dof + k];
        v *= M->JxW[e * Q + q];
        out[r * nb_dof + k] = v;
      }
    }
  }
  return 0;
}

// IntegratorGauss::integrate of a scalar (integrator_gauss_inline_impl.hh:525-541):
// per-element integrals reduced by Math::reduce.
double orc_fe_integrate_scalar(void* h, const double* f, const uint32_t* filter, int64_t nf) {
  auto* M = static_cast<Model*>(h);
  const int64_t cnt = filter_count(*M, filter, nf);
  std::vector<double> per(size_t(cnt), 0.);
  orc_fe_integrate(h, f, 1, filter, nf, per.data());
  return math_reduce(per);
}

// FEEngineTemplate::assembleFieldLumped -> assembleLumpedRowSum
// (fe_engine_template_tmpl_field.hh:126-199): computeNtb, integrate, then
// DOFManager::assembleElementalArrayToLumpedMatrix: element-order scatter into a
// zeroed local array (dof_manager.cc:83-137), added once into the lumped vector
// (dof_manager_default.cc:227-251).  field (E*Q, nb_dof) -> lumped (N, nb_dof),
// which is added to (not overwritten).
int orc_fe_assemble_field_lumped(void* h, const double* field, int nb_dof, double* lumped) {
  auto* M = static_cast<Model*>(h);
  const int n = M->n, Q = M->nq;
  if (nb_dof < 1 || nb_dof > 24) return -1;
  const size_t E = M->E;
  std::vector<double> ntb(E * Q * nb_dof * n), integ(E * nb_dof * n);
  orc_fe_ntb(h, field, nb_dof, nullptr, 0, ntb.data());
  orc_fe_integrate(h, ntb.data(), nb_dof * n, nullptr, 0, integ.data());
  std::vector<double> local(M->N * nb_dof, 0.);
  for (size_t e = 0; e < E; ++e)
    for (int a = 0; a < n; ++a)
      for (int k = 0; k < nb_dof; ++k)
        local[size_t(M->conn[e * n + a]) * nb_dof + k] += 1. * integ[e * nb_dof * n + a * nb_dof + k];
  for (size_t k = 0; k < local.size(); ++k) lumped[k] += 1. * local[k];
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Material operators (material.hh:110-195) on one material's elements.     */
/* ------------------------------------------------------------------------ */
namespace {
void point_law(Model& M, const Material& mat, size_t eq, const double* G, double* S) {
  const int d = M.d;
  if (mat.kind == ELASTIC) elastic_stress(d, mat, G, S);
  else if (mat.kind == NEOHOOKEAN) neohookean_stress(d, mat, G, S);
  else neohookean_damage_stress(d, mat, G, S, M.damage[eq], M.Yd[eq]);
}
}  // namespace

// Material::computeStress (material.hh:110-113; material_elastic.cc:110-145,
// material_neohookean.cc:126-160, Marigo update): from a given gradient field of
// the material's points (in its element-filter order), the stress; the damage
// history of those points is updated, and both fields are kept as the
// material's internals.
int orc_material_compute_stress(void* h, int mat, const double* gradu, double* stress) {
  auto* M = static_cast<Model*>(h);
  if (mat < 0 || mat >= int(M->mats.size())) return -1;
  Material& m = M->mats[mat];
  const int d = M->d, Q = M->nq, dd = d * d;
  size_t t = 0;
  for (uint32_t e : m.filter)
    for (int q = 0; q < Q; ++q, ++t) {
      const size_t eq = size_t(e) * Q + q;
      std::memcpy(&M->gradu[eq * dd], gradu + t * dd, sizeof(double) * dd);
      point_law(*M, m, eq, &M->gradu[eq * dd], &M->stress[eq * dd]);
      std::memcpy(stress + t * dd, &M->stress[eq * dd], sizeof(double) * dd);
    }
  return 0;
}

// Material::computeAllStresses (material.cc:254-279): gradient of the current
// displacement on the material's points, minus the (zero) eigengradient, then
// computeStress.
int orc_material_compute_all_stresses(void* h, int mat) {
  auto* M = static_cast<Model*>(h);
  if (mat < 0 || mat >= int(M->mats.size())) return -1;
  Material& m = M->mats[mat];
  const int d = M->d, n = M->n, Q = M->nq, dn = d * n, dd = d * d;
  double ue[24];
  for (uint32_t e : m.filter) {
    for (int a = 0; a < n; ++a)
      for (int i = 0; i < d; ++i) ue[i + a * d] = M->u[size_t(M->conn[size_t(e) * n + a]) * d + i];
    for (int q = 0; q < Q; ++q) {
      const size_t eq = size_t(e) * Q + q;
      double* G = &M->gradu[eq * dd];
      matrix_matrixt(d, d, n, ue, &M->B[eq * dn], G);
      for (int k = 0; k < dd; ++k) G[k] -= 0.;
      point_law(*M, m, eq, G, &M->stress[eq * dd]);
    }
  }
  return 0;
}

// Material::assembleInternalForces (material.cc:178-246 small strain,
// :637-722 finite strain): B^T sigma (or the finite-strain residual) from the
// material's stored gradient and stress, integrated, and scattered with scale
// -1 into the model's internal force in element order (added to what is there).
int orc_material_assemble_internal_forces(void* h, int mat) {
  auto* M = static_cast<Model*>(h);
  if (mat < 0 || mat >= int(M->mats.size())) return -1;
  const Material& m = M->mats[mat];
  const int d = M->d, n = M->n, Q = M->nq, dn = d * n, dd = d * d;
  double Rq[8][24], fe[24];
  for (uint32_t e : m.filter) {
    for (int q = 0; q < Q; ++q) {
      const size_t eq = size_t(e) * Q + q;
      const double* Bq = &M->B[eq * dn];
      if (m.kind == ELASTIC) matrix_matrix(d, n, d, &M->stress[eq * dd], Bq, Rq[q]);
      else finite_strain_residual(d, n, Bq, &M->gradu[eq * dd], &M->stress[eq * dd], Rq[q]);
    }
    for (int k = 0; k < dn; ++k) {
      double s = 0.;
      for (int q = 0; q < Q; ++q) s += Rq[q][k] * M->JxW[size_t(e) * Q + q];
      fe[k] = s * 1.;
    }
    for (int a = 0; a < n; ++a) {
      double* dst = &M->fint[size_t(M->conn[size_t(e) * n + a]) * d];
      for (int i = 0; i < d; ++i) dst[i] += -1. * fe[i + a * d];
    }
  }
  return 0;
}

// Material::getPotentialEnergy (material.cc:743-757) of one material, from
// the current displacement (as orc_energy kind 1 computes it per material).
double orc_material_potential_energy(void* h, int mat) {
  auto* M = static_cast<Model*>(h);
  if (mat < 0 || mat >= int(M->mats.size())) return 0.;
  Material& m = M->mats[mat];
  const int d = M->d, Q = M->nq;
  double fe[24];
  std::vector<double> dam_save;
  if (m.kind == NEOHOOKEAN_DAMAGE) dam_save = M->damage;
  double part = 0.;
  for (uint32_t e : m.filter) {
    element_force(*M, m, e, fe);
    double s_e = 0.;
    for (int q = 0; q < Q; ++q) {
      size_t eq = size_t(e) * Q + q;
      const double* G = &M->gradu[eq * d * d];
      const double* S = &M->stress[eq * d * d];
      double w;
      if (m.kind == ELASTIC) {
        double dd = 0.;
        for (int k = 0; k < d * d; ++k) dd += S[k] * G[k];
        w = .5 * dd;
      } else {
        w = neohookean_energy(d, m, G);
        if (m.kind == NEOHOOKEAN_DAMAGE) w *= (1 - M->damage[eq]);
      }
      s_e += w * M->JxW[eq];
    }
    part += s_e;
  }
  if (m.kind == NEOHOOKEAN_DAMAGE) M->damage = dam_save;
  return part;
}

// element filter of a material (ascending element ids)
int64_t orc_material_elements(void* h, int mat, uint32_t* ids) {
  auto* M = static_cast<Model*>(h);
  if (mat < 0 || mat >= int(M->mats.size())) return -1;
  const auto& f = M->mats[mat].filter;
  if (ids) std::copy(f.begin(), f.end(), ids);
  return int64_t(f.size());
}

} // extern "C"
