/**
 * solid_mechanics_model.hh -- SolidMechanicsModel for the explicit lumped-mass
 * path, on one MI355X through libakx.so.
 *
 * Mirrors src/model/solid_mechanics/solid_mechanics_model.hh:58-543 for this
 * path: construction on a Mesh, initFull(_analysis_method =
 * _explicit_lumped_mass) (materials instantiated from the parsed input file),
 * getStableTimeStep / setTimeStep / solveStep, assembleInternalForces, the
 * nodal arrays getDisplacement / getVelocity / getAcceleration /
 * getExternalForce / getInternalForce / getBlockedDOFs / getMass /
 * getCurrentPosition / getIncrement, getEnergy("kinetic" | "potential" |
 * "external work"), applyBC(functor, group), getMaterial.
 *
 * The nodal arrays live on the GPU.  The host Array returned by an accessor is
 * refreshed from the device when the device copy is newer, and is sent back
 * before the next device operation if a non-const accessor handed it out
 * (the caller may have written into it) -- so drivers written against the
 * reference (write initial conditions through getVelocity(), read results
 * through getDisplacement()) work unchanged.  Host copies cost PCIe traffic:
 * a driver that only reads energies never moves nodal arrays.
 */
#pragma once

#include <future>
#include <memory>
#include <string>
#include <vector>

#include "aka_array.hh"
#include "aka_common.hh"
#include "boundary_condition_functor.hh"
#include "fe_engine.hh"
#include "material.hh"
#include "mesh.hh"

struct akx_ctx;

namespace akantu {

class SolidMechanicsModel {
 public:
  explicit SolidMechanicsModel(Mesh& mesh, UInt spatial_dimension = 0,
                               const ID& id = "solid_mechanics_model", int device = 0);
  ~SolidMechanicsModel();
  SolidMechanicsModel(const SolidMechanicsModel&) = delete;
  SolidMechanicsModel& operator=(const SolidMechanicsModel&) = delete;

  void initFull(const AnalysisMethodArg& method = AnalysisMethodArg{});

  Real getStableTimeStep();
  void setTimeStep(Real dt);
  Real getTimeStep() const { return time_step_; }

  /// one explicit central-difference step (SolidMechanicsModel::solveStep)
  void solveStep();
  /// nsteps steps without host round trips (extension)
  void solveStep(UInt nsteps);
  void assembleInternalForces();

  Array<Real>& getDisplacement();
  Array<Real>& getVelocity();
  Array<Real>& getAcceleration();
  Array<Real>& getExternalForce();
  Array<Real>& getForce() { return getExternalForce(); }
  /// writable as in the reference: Material::assembleInternalForces adds into it
  Array<Real>& getInternalForce();
  Array<bool>& getBlockedDOFs();
  const Array<Real>& getMass();
  const Array<Real>& getCurrentPosition();
  const Array<Real>& getIncrement();
  const Array<Real>& getPreviousDisplacement();

  Real getEnergy(const std::string& energy_id);
  Real getKineticEnergy() { return getEnergy("kinetic"); }
  Real getPotentialEnergy() { return getEnergy("potential"); }
  Real getExternalWork() { return getEnergy("external work"); }

  /// Dirichlet / Neumann functors over a node group of the mesh
  /// (BoundaryCondition::applyBC, boundary_condition_tmpl.hh:62-88)
  void applyBC(const BC::Dirichlet::FixedValue& f, const std::string& group);
  void applyBC(const BC::Dirichlet::FlagOnly& f, const std::string& group);
  void applyBC(const BC::Dirichlet::IncrementValue& f, const std::string& group);
  void applyBC(const BC::Neumann::FromTraction& f, const std::string& group);
  /// per-step IncrementValue applied on the device before every step
  /// (extension: the reference driver calls applyBC(IncrementValue) in its loop)
  void setDirichletIncrement(const std::string& group, SpatialDirection axis, Real increment);

  UInt getNbMaterials() const { return UInt(materials_.size()); }
  Material& getMaterial(UInt i);
  Material& getMaterial(const std::string& name);
  /// element -> material index (before initFull; default: every element -> 0)
  void setMaterialIndex(const Array<UInt>& element_material);
  /// add a material in code (besides the input-file sections)
  Material& registerNewMaterial(const std::string& type, const std::string& name);

  /// the model's FEEngine (SolidMechanicsModel::getFEEngine, model.hh)
  FEEngine& getFEEngine(const ID& name = "");

  Mesh& getMesh() { return mesh_; }
  UInt getSpatialDimension() const { return dim_; }

  /// ParaView output (Model::setBaseName / addDumpField / dump, src/model/model.cc):
  /// paraview/<base>_<NNNN>.vtu per dump and paraview/<base>.pvd listing them.
  /// Nodal fields: displacement, velocity, acceleration, external_force,
  /// internal_force, mass, blocked_dofs; per-element (averaged over the
  /// quadrature points): stress, grad_u, damage.  dump() snapshots the
  /// fields on the device and returns; a background thread waits for the
  /// copies and writes the files while the caller keeps stepping.
  void setBaseName(const std::string& name, const std::string& directory = "paraview");
  void addDumpField(const std::string& field);
  void dump();
  /// wait until every dump so far is on disk (rethrows a write error)
  void flushDumps();

  /// extensions: arithmetic of the element kernels, launch autotuning, kernel times
  void setPrecision(Precision p) { precision_ = p; }
  Precision getPrecision() const { return precision_; }
  void autotune(UInt nsteps = 10);
  void getLastStepTiming(Real& element_ms, Real& node_ms);

  // used by Material and FEEngine
  akx_ctx* handle() { return ctx_; }
  void downloadPointField(int field, Array<Real>& out, const Material& m);
  /// sends the host arrays a caller may have written, before a device operation
  void synchronizeHost() { pushDirty(); }
  /// changes whenever the displacement or the material state may have changed
  UInt stateVersion() const { return state_version_; }
  void internalForceChanged() { host_stale_[kFint] = true; }
  UInt materialIndex(const Material& m) const;
  bool isInitialized() const { return initialized_; }

 private:
  enum NodalId { kU, kV, kA, kFext, kFint, kBlocked, kMass, kPos, kInc, kPrev, kCount };
  void check(int rc, const char* what) const;
  void pushDirty();
  void pull(NodalId id);
  void instantiateMaterials();

  Mesh& mesh_;
  UInt dim_;
  ID id_;
  akx_ctx* ctx_ = nullptr;
  bool initialized_ = false;
  Real time_step_ = 0.;
  Precision precision_ = Precision::fp64;
  std::vector<std::unique_ptr<Material>> materials_;
  Array<UInt> element_material_;
  bool custom_assignment_ = false;

  Array<Real> u_, v_, a_, fext_, fint_, mass_, pos_, inc_, prev_;
  Array<bool> blocked_;
  std::string dump_base_ = "model", dump_dir_ = "paraview";
  std::vector<std::string> dump_fields_;
  std::vector<std::pair<Real, std::string>> dumps_;
  struct DumpBuffer {
    void* host = nullptr;  // page-locked (akx_host_alloc)
    size_t bytes = 0;
  };
  std::vector<DumpBuffer> dump_bufs_;  // one per dump field, reused
  std::future<void> dump_job_;         // the file writer of the last dump
  UInt steps_done_ = 0;
  UInt state_version_ = 0;
  std::unique_ptr<FEEngine> fe_engine_;
  bool host_stale_[kCount] = {};
  bool host_dirty_[kCount] = {};
};

}  // namespace akantu
