/**
 * fe_engine_operators.cc -- the FEEngine and Material operators of the
 * akantu:: API, driven the way the reference's own FEEngine tests drive them:
 *
 *   - gradient of a linear field (test/test_fe_engine/test_gradient.cc:46-102):
 *     the constant gradient at every point, tolerance 5e-13;
 *   - integral of 1 and of a linear field (test/test_fe_engine/test_integrate.cc:
 *     41-72): the volume and the first moment of the mesh;
 *   - lumped assembly of rho (SolidMechanicsModel::assembleMassLumped,
 *     solid_mechanics_model_mass.cc:58-102) equal to the model's mass;
 *   - SolidMechanicsModel::assembleInternalForces (solid_mechanics_model.cc:
 *     375-414) restated on the Material API: clear the internal force,
 *     computeAllStresses, then assembleInternalForces, material by material;
 *     equal, bit for bit, to the model's own assembly.
 *
 *   fe_engine_operators <mesh.msh> <material.dat> <dim> <out_prefix>
 *
 * Writes <out_prefix>_u.bin (the displacement u_i = 1e-3 sin(x . k_i)),
 * <out_prefix>_gradient.bin (its gradient on the points),
 * <out_prefix>_fint.bin (the internal force of the Material-API assembly
 * under that displacement) and <out_prefix>_lumped.bin as raw float64, prints
 * one line per check and returns 0 only if all pass.
 */
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "akantu/solid_mechanics_model.hh"

using namespace akantu;

namespace {

int failures = 0;

void expect_below(const char* what, Real value, Real tol) {
  const bool ok = std::fabs(value) <= tol;
  std::cout << (ok ? "PASS " : "FAIL ") << what << " = " << value << " (tol " << tol << ")"
            << std::endl;
  if (!ok) ++failures;
}

void write(const std::string& path, const Array<Real>& a) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(a.storage()),
          std::streamsize(size_t(a.size()) * a.getNbComponent() * sizeof(Real)));
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc < 5) {
    std::cerr << "usage: " << argv[0] << " mesh.msh material.dat dim out_prefix\n";
    return 2;
  }
  const UInt dim = UInt(std::atoi(argv[3]));
  const std::string out = argv[4];
  try {
    initialize(argv[2], argc, argv);
    Mesh mesh(dim);
    mesh.read(argv[1]);
    const ElementType type = mesh.getElementType();
    SolidMechanicsModel model(mesh);
    // every third element to a second material, as a second section of the
    // input file would be assigned by a material selector
    if (model.getNbMaterials() == 0 && getStaticParser().getMaterialSections().size() > 1) {
      Array<UInt> em(mesh.getNbElement(), 1, 0u);
      for (UInt e = 0; e < mesh.getNbElement(); ++e) em(e) = (e % 3 == 1) ? 1 : 0;
      model.setMaterialIndex(em);
    }
    model.initFull(_analysis_method = _explicit_lumped_mass);
    FEEngine& fem = model.getFEEngine();
    const UInt N = mesh.getNbNodes(), E = mesh.getNbElement();
    const UInt Q = fem.getNbIntegrationPoints(type);
    const auto& X = mesh.getNodes();

    // gradient of a linear field, two components
    const Real A[2][3] = {{0.3, -1.2, 0.7}, {2.0, 0.5, -0.4}};
    Array<Real> lin(N, 2);
    for (UInt n = 0; n < N; ++n)
      for (UInt i = 0; i < 2; ++i) {
        lin(n, i) = 1.5;
        for (UInt j = 0; j < dim; ++j) lin(n, i) += A[i][j] * X(n, j);
      }
    Array<Real> grad;
    fem.gradientOnIntegrationPoints(lin, grad, 2, type);
    Real gerr = 0.;
    for (UInt p = 0; p < grad.size(); ++p)
      for (UInt j = 0; j < dim; ++j)
        for (UInt i = 0; i < 2; ++i) gerr = std::max(gerr, std::fabs(grad(p, i + 2 * j) - A[i][j]));
    expect_below("linear field gradient error", gerr, 5e-13);

    // volume and first moment
    Array<Real> ones(E * Q, 1, 1.), xq;
    const Real vol = fem.integrate(ones, type);
    Real lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (UInt n = 0; n < N; ++n)
      for (UInt j = 0; j < dim; ++j) {
        lo[j] = std::min(lo[j], X(n, j));
        hi[j] = std::max(hi[j], X(n, j));
      }
    Real box = 1.;
    for (UInt j = 0; j < dim; ++j) box *= hi[j] - lo[j];
    expect_below("volume - bounding box volume", vol - box, 1e-13 * box);
    fem.interpolateOnIntegrationPoints(X, xq, dim, type);
    Array<Real> moment;
    fem.integrate(xq, moment, dim, type);
    for (UInt j = 0; j < dim; ++j) {
      Real m = 0.;
      for (UInt e = 0; e < E; ++e) m += moment(e, j);
      expect_below("centroid - box centre", m / vol - 0.5 * (lo[j] + hi[j]),
                   1e-13 * (hi[j] - lo[j]));
    }

    // lumped assembly of rho is the model's lumped mass
    Array<Real> lumped(N, 1, 0.);
    fem.assembleFieldLumped(
        [&](Matrix<Real>& f, const Element& el) {
          f.set(model.getMaterial(model.getNbMaterials() > 1 && el.element % 3 == 1 ? 1 : 0)
                    .getParam("rho"));
        },
        1, lumped, type);
    const auto& mass = model.getMass();
    Real merr = 0.;
    for (UInt n = 0; n < N; ++n) merr = std::max(merr, std::fabs(lumped(n) - mass(n, 0)));
    expect_below("lumped rho - model mass", merr, 0.);
    write(out + "_lumped.bin", lumped);

    // a smooth displacement, its gradient, and the internal force of the
    // Material API against the model's own assembly
    auto& u = model.getDisplacement();
    for (UInt n = 0; n < N; ++n)
      for (UInt i = 0; i < dim; ++i) {
        Real s = 0.;
        for (UInt j = 0; j < dim; ++j) s += (1.0 + i + 0.5 * j) * X(n, j);
        u(n, i) = 1e-3 * std::sin(s);
      }
    Array<Real> gu;
    fem.gradientOnIntegrationPoints(u, gu, dim, type);
    write(out + "_u.bin", u);
    write(out + "_gradient.bin", gu);
    model.assembleInternalForces();
    Array<Real> f_model;
    f_model.copy(model.getInternalForce());
    auto& fint = model.getInternalForce();
    fint.clear();
    for (UInt k = 0; k < model.getNbMaterials(); ++k) model.getMaterial(k).computeAllStresses();
    for (UInt k = 0; k < model.getNbMaterials(); ++k) model.getMaterial(k).assembleInternalForces();
    const auto& f_mat = model.getInternalForce();
    Real ferr = 0., fmax = 0.;
    for (UInt n = 0; n < N; ++n)
      for (UInt i = 0; i < dim; ++i) {
        ferr = std::max(ferr, std::fabs(f_mat(n, i) - f_model(n, i)));
        fmax = std::max(fmax, std::fabs(f_model(n, i)));
      }
    expect_below("material-API internal force - model assembly", ferr, 0.);
    write(out + "_fint.bin", f_mat);
    std::cout << "max |f_int| " << fmax << std::endl;

    // computeStress on the gradients computeAllStresses left: same stress
    Material& m0 = model.getMaterial(0);
    Array<Real> s0;
    s0.copy(m0.getStress(type));
    m0.computeStress(type);
    const auto& s1 = m0.getStress(type);
    Real serr = 0.;
    for (UInt k = 0; k < s0.size() * s0.getNbComponent(); ++k)
      serr = std::max(serr, std::fabs(s1[k] - s0[k]));
    expect_below("computeStress(grad_u internal) - computeAllStresses", serr, 0.);
    std::cout << (failures ? "FAILED" : "OK") << std::endl;
  } catch (debug::Exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return failures ? 1 : 0;
}
