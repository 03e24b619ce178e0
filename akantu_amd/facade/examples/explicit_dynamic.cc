/**
 * explicit_dynamic.cc -- the canonical explicit driver on the akantu:: API of
 * libakantu_amd.so: read a mesh and a material file, take 0.8 of the stable
 * time step, start from a sine-times-Gaussian displacement pulse along x and
 * run central-difference steps, writing the energies of every step.
 *
 * Same loop as examples/explicit/explicit_dynamic.cc:38-106 of the reference
 * (initialize -> Mesh::read -> initFull(_explicit_lumped_mass) ->
 * getStableTimeStep -> pulse through getDisplacement() -> solveStep +
 * getEnergy per step), with the ParaView dumps of displacement, velocity,
 * acceleration and stress every k steps when k is given (the reference dumps
 * every step).
 *
 *   explicit_dynamic <mesh.msh> <material.dat> <dim> <steps> <fp64|mixed> <out prefix>
 *                    [dump every k steps]
 *
 * Writes <prefix>energy.csv (id,rtime,epot,ekin,tot), <prefix>dt.txt, and the
 * initial and final displacements as <prefix>u0.txt and <prefix>u.txt (one
 * node per line, %.17g).
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "akantu/solid_mechanics_model.hh"

using namespace akantu;

static void write_nodal(const std::string& path, const Array<Real>& a, UInt dim) {
  std::FILE* f = std::fopen(path.c_str(), "w");
  if (!f) AKANTU_EXCEPTION("cannot write " << path);
  for (UInt n = 0; n < a.size(); ++n) {
    for (UInt i = 0; i < dim; ++i) std::fprintf(f, i ? " %.17g" : "%.17g", a(n, i));
    std::fprintf(f, "\n");
  }
  std::fclose(f);
}

int main(int argc, char* argv[]) {
  if (argc < 7) {
    std::cerr << "usage: " << argv[0]
              << " mesh.msh material.dat dim steps fp64|mixed out_prefix\n";
    return 2;
  }
  const std::string mesh_file = argv[1], material_file = argv[2];
  const UInt dim = UInt(std::atoi(argv[3]));
  const UInt max_steps = UInt(std::atoi(argv[4]));
  const std::string precision = argv[5], prefix = argv[6];
  const UInt dump_every = argc > 7 ? UInt(std::atoi(argv[7])) : 0;
  try {
    initialize(material_file, argc, argv);
    const Real pulse_width = 2.;
    const Real A = 0.01;
    const Real time_factor = 0.8;

    Mesh mesh(dim);
    mesh.read(mesh_file);
    SolidMechanicsModel model(mesh);
    model.initFull(_analysis_method = _explicit_lumped_mass);
    model.setPrecision(precision == "mixed" ? Precision::mixed : Precision::fp64);

    const Real time_step = model.getStableTimeStep();
    std::cout << "Time Step = " << time_step * time_factor << "s (" << time_step << "s)"
              << std::endl;
    model.setTimeStep(time_step * time_factor);

    Array<Real>& displacement = model.getDisplacement();
    const Array<Real>& nodes = mesh.getNodes();
    for (UInt n = 0; n < mesh.getNbNodes(); ++n) {
      const Real x = nodes(n) - 2;
      const Real L = pulse_width;
      const Real k = 0.1 * 2 * M_PI * 3 / L;
      displacement(n) = A * std::sin(k * x) * std::exp(-(k * x) * (k * x) / (L * L));
    }

    write_nodal(prefix + "u0.txt", displacement, dim);

    if (dump_every) {
      model.setBaseName("explicit_dynamic", prefix + "paraview");
      model.addDumpField("displacement");
      model.addDumpField("velocity");
      model.addDumpField("acceleration");
      model.addDumpField("stress");
      model.dump();
    }

    std::ofstream energy(prefix + "energy.csv");
    energy.precision(17);
    energy << "id,rtime,epot,ekin,tot" << std::endl;
    for (UInt s = 1; s <= max_steps; ++s) {
      model.solveStep();
      const Real epot = model.getEnergy("potential");
      const Real ekin = model.getEnergy("kinetic");
      energy << s << "," << s * time_step << "," << epot << "," << ekin << "," << epot + ekin
             << std::endl;
      if (s % 100 == 0) std::cout << "passing step " << s << "/" << max_steps << std::endl;
      if (dump_every && s % dump_every == 0) model.dump();  // written in the background
    }
    model.flushDumps();

    std::ofstream dt_out(prefix + "dt.txt");
    dt_out.precision(17);
    dt_out << time_step * time_factor << std::endl;
    write_nodal(prefix + "u.txt", model.getDisplacement(), dim);
    finalize();
  } catch (const debug::Exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return EXIT_SUCCESS;
}
