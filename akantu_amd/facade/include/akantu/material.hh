/**
 * material.hh -- constitutive laws of the explicit path as the façade sees
 * them: parameters (registerParam / getParam / setParam, material.cc:104-121)
 * and the per-point fields a driver reads (getStress, getGradU,
 * getPotentialEnergy).  The arithmetic runs in libakx.so.
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

namespace akantu {

class SolidMechanicsModel;

class Material {
 public:
  Material(SolidMechanicsModel& model, const ID& id, int kind);
  virtual ~Material() = default;

  const ID& getID() const { return id_; }
  const std::string& getName() const { return name_; }
  int getKind() const { return kind_; }

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

  /// per-point fields of this material's elements (E_m*Q, d*d), column-major
  const Array<Real>& getStress(ElementType type);
  const Array<Real>& getGradU(ElementType type);
  const Array<Real>& getDamage(ElementType type);
  Real getPotentialEnergy();
  Real getEnergy(const std::string& id);

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
