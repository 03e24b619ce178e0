/**
 * patch_test_explicit.cc -- the linear patch test of the explicit path on the
 * akantu:: API of libakantu_amd.so.
 *
 * Follows test/test_model/patch_tests/patch_test_linear_elastic_explicit.cc:33-68
 * with the fixture of patch_test_linear_fixture.hh and
 * patch_test_linear_solid_mechanics_fixture.hh: every node starts on the
 * linear field u_i = alpha(i,0) + sum_j alpha(i,j+1) x_j, the boundary DOFs are
 * blocked, 100 steps at 0.8 of the stable time step must leave E_kin ~ 0,
 * the displacement on the linear field, the gradient and the stress constant
 * and equal to the prescribed ones, and the internal forces self-equilibrated.
 *
 *   patch_test_explicit <mesh.msh> <material.dat> <dim>
 *
 * Prints one line per check and returns 0 only if all pass, with the
 * reference's tolerances (patch_test_linear_fixture.hh:171-173).
 */
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "akantu/solid_mechanics_model.hh"

using namespace akantu;

namespace {

const Real alpha[3][4] = {{0.01, 0.02, 0.03, 0.04},
                          {0.05, 0.06, 0.07, 0.08},
                          {0.09, 0.10, 0.11, 0.12}};

int failures = 0;

void expect_below(const char* what, Real value, Real tol) {
  const bool ok = std::fabs(value) <= tol;
  std::cout << (ok ? "PASS " : "FAIL ") << what << " = " << value << " (tol " << tol << ")"
            << std::endl;
  if (!ok) ++failures;
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc < 4) {
    std::cerr << "usage: " << argv[0] << " mesh.msh material.dat dim\n";
    return 2;
  }
  const UInt dim = UInt(std::atoi(argv[3]));
  try {
    initialize(argv[2], argc, argv);
    Mesh mesh(dim);
    mesh.read(argv[1]);
    mesh.createBoundaryGroup();
    SolidMechanicsModel model(mesh);
    model.initFull(_analysis_method = _explicit_lumped_mass);
    model.setTimeStep(model.getStableTimeStep() * 0.8);

    const Array<Real>& X = mesh.getNodes();
    const UInt N = mesh.getNbNodes();
    Array<bool>& blocked = model.getBlockedDOFs();
    for (UInt n : mesh.getNodeGroup("boundary"))
      for (UInt i = 0; i < dim; ++i) blocked(n, i) = true;
    auto linear = [&](UInt n, UInt i) {
      Real v = alpha[i][0];
      for (UInt j = 0; j < dim; ++j) v += alpha[i][j + 1] * X(n, j);
      return v;
    };
    Array<Real>& u0 = model.getDisplacement();
    for (UInt n = 0; n < N; ++n)
      for (UInt i = 0; i < dim; ++i) u0(n, i) = linear(n, i);

    for (UInt s = 0; s < 100; ++s) model.solveStep();

    expect_below("kinetic energy", model.getEnergy("kinetic"), 1e-16);

    const Array<Real>& u = model.getDisplacement();
    Real dof_err = 0;
    for (UInt n = 0; n < N; ++n)
      for (UInt i = 0; i < dim; ++i) dof_err = std::max(dof_err, std::fabs(u(n, i) - linear(n, i)));
    expect_below("displacement error", dof_err, 1e-15);

    // prescribed gradient G_ij = alpha(i, j+1); stress of the plane-strain /
    // 3D elastic law (patch_test_linear_solid_mechanics_fixture.hh:98-124)
    Material& mat = model.getMaterial(0);
    const Real E = mat.get("E"), nu = mat.get("nu");
    const Real lambda = nu * E / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
    Real G[3][3] = {}, S[3][3] = {}, tr = 0, smax = 0;
    for (UInt i = 0; i < dim; ++i) {
      tr += alpha[i][i + 1];
      for (UInt j = 0; j < dim; ++j) G[i][j] = alpha[i][j + 1];
    }
    for (UInt i = 0; i < dim; ++i)
      for (UInt j = 0; j < dim; ++j) {
        S[i][j] = (i == j) * lambda * tr + mu * (G[i][j] + G[j][i]);
        smax = std::max(smax, std::fabs(S[i][j]));
      }
    const Array<Real>& grad = mat.getGradU(mesh.getElementType());
    const Array<Real>& stress = mat.getStress(mesh.getElementType());
    Real grad_err = 0, stress_err = 0;
    for (UInt q = 0; q < grad.size(); ++q) {
      Real gd = 0, gn = 0, sd = 0;
      for (UInt i = 0; i < dim; ++i)
        for (UInt j = 0; j < dim; ++j) {
          // column-major d x d per point: entry (i, j) at j * d + i
          const Real g = grad(q, j * dim + i), s = stress(q, j * dim + i);
          gd = std::max(gd, std::fabs(g - G[i][j]));
          gn = std::max(gn, std::fabs(g));
          sd = std::max(sd, std::fabs(s - S[i][j]));
        }
      grad_err = std::max(grad_err, gd / gn);
      stress_err = std::max(stress_err, sd / smax);
    }
    std::cout << "points checked: " << grad.size() << std::endl;
    expect_below("gradient error", grad_err, 1e-13);
    expect_below("stress error", stress_err, 1e-13);

    model.assembleInternalForces();
    const Array<Real>& fint = model.getInternalForce();
    Real total[3] = {0, 0, 0}, fmax = 0;
    for (UInt n = 0; n < N; ++n)
      for (UInt i = 0; i < dim; ++i) {
        total[i] += fint(n, i);
        fmax = std::max(fmax, std::fabs(fint(n, i)));
      }
    Real tmax = 0;
    for (UInt i = 0; i < dim; ++i) tmax = std::max(tmax, std::fabs(total[i]));
    expect_below("internal force balance", tmax / fmax, 1e-9);
    finalize();
  } catch (const debug::Exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  std::cout << (failures ? "FAILED" : "OK") << std::endl;
  return failures ? 1 : 0;
}
