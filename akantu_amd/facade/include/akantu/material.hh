/**
 * material.hh -- constitutive laws of the explicit path as the façade sees
 * them: parameters (registerParam / getParam / setParam, material.cc:104-121),
 * the per-point fields a driver reads (getStress, getGradU, getDamage), and
 * the operators of material.hh:110-195 (computeStress, computeAllStresses,
 * assembleInternalForces, getPotentialEnergy).  The arithmetic runs in
 * libakx.so, in FP64 with the reference's operation order.
 *
 * Laws (MaterialFactory ids, `material <id> [ ... ]` input-file sections):
 *   elastic           MaterialElastic<dim>       (material_elastic.cc)
 *   neohookean        MaterialNeohookean<dim>    (material_neohookean.cc)
 *   neohookean_damage neo-Hookean + Marigo isotropic damage (the Marigo update of
 *                     material_marigo_inline_impl.hh:67-76 driven by the
 *                     neo-Hookean energy density; no counterpart in the reference)
 */
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <string>

#include "aka_array.hh"
#include "aka_common.hh"
#include "fe_engine.hh"

namespace akantu {

class SolidMechanicsModel;

class Material {
 public:
  Material(SolidMechanicsModel& model, const ID& id, int kind);
  virtual ~Material() = default;

  const ID& getID() const { return id_; }
  const std::string& getName() const { return name_; }
  int getKind() const { return kind_; }
  /// index of the material in the model (creation order)
  UInt getIndex() const;

  /// parameters: "name", "rho", "E", "nu", "Plane_Stress", "Sd", "Yd"
  Real getParam(const std::string& key) const;
  Real get(const std::string& key) const { return getParam(key); }
  void setParam(const std::string& key, Real value);
  void setName(const std::string& name) { name_ = name; }
  bool hasParam(const std::string& key) const { return params_.count(key) > 0; }

  /// random threshold `Yd = base weibull [scale, shape]` (parser_random.cc:71-83)
  void setRandomParam(const std::string& key, Real base, const std::string& distribution,
                      Real p0, Real p1);

  /// Lamé coefficients after initMaterial (material_elastic.cc:83-107)
  Real getLambda() const;
  Real getMu() const;
  Real getPushWaveSpeed() const;
  Real getCelerity() const { return getPushWaveSpeed(); }

  /// per-point fields of this material's elements (E_m*Q, d*d), column-major:
  /// the internals of the last computeStress / computeAllStresses, or, when
  /// the model has stepped since, the fields of the current displacement
  const Array<Real>& getStress(ElementType type);
  /// writable, as the "grad_u" internal is: computeStress reads it
  Array<Real>& getGradU(ElementType type);
  const Array<Real>& getDamage(ElementType type);
  Real getPotentialEnergy();
  Real getEnergy(const std::string& id);

  /// Material::computeStress (material.hh:110-113): stress of this material's
  /// points from its gradient internal (getGradU); updates the damage history
  void computeStress(ElementType type, GhostType ghost_type = _not_ghost);
  /// Material::computeAllStresses (material.hh:195, material.cc:254-279):
  /// gradient of the model's displacement on the points, then computeStress
  void computeAllStresses(GhostType ghost_type = _not_ghost);
  /// Material::assembleInternalForces (material.hh:186, material.cc:178-246,
  /// 637-722): adds -int B^T sigma of this material's elements, from its
  /// stress internal, to the model's internal force, in element order
  void assembleInternalForces(GhostType ghost_type = _not_ghost);

  /// element filter of this material (ascending element ids)
  const Array<UInt>& getElementFilter() const { return filter_; }

 private:
  friend class SolidMechanicsModel;
  SolidMechanicsModel& model_;
  ID id_;
  std::string name_;
  int kind_;
  std::map<std::string, Real> params_;
  struct RandomSpec {
    Real base, p0, p1;
    std::string dist;
  };
  std::map<std::string, RandomSpec> random_;
  Array<UInt> filter_;
  Array<Real> stress_, gradu_, damage_;
  // model state each host field was taken at: gradu_ / stress_ are the
  // internals of the last computeStress / computeAllStresses (or a download)
  // while the model has not moved since
  UInt gradu_version_ = ~0u, stress_version_ = ~0u;
};

/// MaterialFactory::registerAllocator / INSTANTIATE_MATERIAL (material.hh:59-60,687-691)
class MaterialFactory {
 public:
  using Allocator = std::function<std::unique_ptr<Material>(const ID&, SolidMechanicsModel&)>;
  static MaterialFactory& getInstance();
  void registerAllocator(const std::string& name, Allocator alloc);
  std::unique_ptr<Material> allocate(const std::string& name, const ID& id,
                                     SolidMechanicsModel& model) const;
  bool has(const std::string& name) const { return allocators_.count(name) > 0; }

 private:
  MaterialFactory();
  std::map<std::string, Allocator> allocators_;
};

/// Input-file sections `material <type> [ key = value ... ]` (parser.hh:88-443).
struct ParserSection {
  std::string type;  // material law id
  std::map<std::string, std::string> params;
};

class Parser {
 public:
  void parse(const std::string& filename);
  void parseString(const std::string& text);
  const std::vector<ParserSection>& getMaterialSections() const { return sections_; }
  void clear() { sections_.clear(); }
  long getSeed() const { return seed_; }
  void setSeed(long s) { seed_ = s; }

 private:
  std::vector<ParserSection> sections_;
  long seed_ = 5489;  // aka_extern.cc:100
};

Parser& getStaticParser();

}  // namespace akantu
