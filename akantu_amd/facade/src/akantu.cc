// akantu.cc -- the akantu:: façade over the C ABI of libakx.so.
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <future>
#include <limits>
#include <map>
#include <random>
#include <set>
#include <sstream>

#include "akantu/aka_random_generator.hh"
#include "akantu/solid_mechanics_model.hh"
#include "akx.h"

namespace akantu {

/* ------------------------------------------------------------------------ */
/* common                                                                    */
/* ------------------------------------------------------------------------ */
UInt getSpatialDimension(ElementType t) {
  return (t == _tetrahedron_4 || t == _hexahedron_8) ? 3 : 2;
}
UInt getNbNodesPerElement(ElementType t) {
  return t == _triangle_3 ? 3 : t == _hexahedron_8 ? 8 : 4;
}
UInt getNbIntegrationPoints(ElementType t) {
  return t == _quadrangle_4 ? 4 : t == _hexahedron_8 ? 8 : 1;
}

void drawRandomValues(std::default_random_engine& engine, const std::string& dist, Real base,
                      Real p0, Real p1, Real* out, size_t n) {
  if (dist == "weibull") {
    std::weibull_distribution<Real> w(p1, p0);  // [scale, shape] -> (shape, scale)
    for (size_t k = 0; k < n; ++k) out[k] = base + w(engine);
  } else if (dist == "uniform") {
    std::uniform_real_distribution<Real> u(p0, p1);
    for (size_t k = 0; k < n; ++k) out[k] = base + u(engine);
  } else {
    AKANTU_EXCEPTION("unsupported random distribution " << dist);
  }
}

Parser& getStaticParser() {
  static Parser p;
  return p;
}

void initialize(int& argc, char**& argv) {
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a.rfind("--aka_seed=", 0) == 0) getStaticParser().setSeed(std::stol(a.substr(11)));
    else if (a == "--aka_seed" && i + 1 < argc) getStaticParser().setSeed(std::stol(argv[++i]));
  }
}

void initialize(const std::string& input_file, int& argc, char**& argv) {
  if (!input_file.empty()) getStaticParser().parse(input_file);
  initialize(argc, argv);
}

void finalize() {}

/* ------------------------------------------------------------------------ */
/* Parser: `material <type> [ key = value ... ]` sections, `seed = N`         */
/* ------------------------------------------------------------------------ */
static std::string trim(const std::string& s) {
  size_t b = s.find_first_not_of(" \t\r\n"), e = s.find_last_not_of(" \t\r\n");
  return b == std::string::npos ? std::string() : s.substr(b, e - b + 1);
}

void Parser::parse(const std::string& filename) {
  std::ifstream f(filename);
  if (!f) AKANTU_EXCEPTION("cannot open input file " << filename);
  std::stringstream ss;
  ss << f.rdbuf();
  parseString(ss.str());
}

void Parser::parseString(const std::string& text) {
  // strip comments
  std::string clean;
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    auto h = line.find('#');
    if (h != std::string::npos) line = line.substr(0, h);
    clean += line + "\n";
  }
  size_t pos = 0;
  while (pos < clean.size()) {
    size_t m = clean.find("material", pos);
    // top-level `seed = N` before the next section
    std::string head = clean.substr(pos, m == std::string::npos ? std::string::npos : m - pos);
    std::istringstream hs(head);
    while (std::getline(hs, line)) {
      auto eq = line.find('=');
      if (eq != std::string::npos && trim(line.substr(0, eq)) == "seed")
        seed_ = std::stol(trim(line.substr(eq + 1)));
    }
    if (m == std::string::npos) break;
    size_t open = clean.find('[', m);
    if (open == std::string::npos) AKANTU_EXCEPTION("material section without '['");
    ParserSection sec;
    sec.type = trim(clean.substr(m + 8, open - m - 8));
    // matching ']' (random parameters have their own brackets)
    int depth = 0;
    size_t close = open;
    for (; close < clean.size(); ++close) {
      if (clean[close] == '[') ++depth;
      if (clean[close] == ']' && --depth == 0) break;
    }
    if (close >= clean.size()) AKANTU_EXCEPTION("material section without ']'");
    std::istringstream body(clean.substr(open + 1, close - open - 1));
    while (std::getline(body, line)) {
      auto eq = line.find('=');
      if (eq == std::string::npos) continue;
      sec.params[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
    }
    sections_.push_back(sec);
    pos = close + 1;
  }
}

/* ------------------------------------------------------------------------ */
/* Mesh                                                                      */
/* ------------------------------------------------------------------------ */
Mesh::Mesh(UInt spatial_dimension, const ID& id) : dim_(spatial_dimension), id_(id) {
  nodes_.resize(0, spatial_dimension);
}

void Mesh::setNodes(const Array<Real>& nodes) {
  if (nodes.getNbComponent() != dim_) AKANTU_EXCEPTION("nodes must have " << dim_ << " components");
  nodes_.copy(nodes);
}

void Mesh::setConnectivity(ElementType type, const Array<UInt>& conn) {
  if (akantu::getSpatialDimension(type) != dim_) AKANTU_EXCEPTION("element type of another dimension");
  if (conn.getNbComponent() != getNbNodesPerElement(type)) AKANTU_EXCEPTION("bad connectivity width");
  type_ = type;
  conn_.copy(conn);
}

UInt Mesh::getNbElement(ElementType type) const { return type == type_ ? conn_.size() : 0; }

const Array<UInt>& Mesh::getConnectivity(ElementType type) const {
  if (type != type_) AKANTU_EXCEPTION("mesh has no elements of type " << int(type));
  return conn_;
}

std::vector<UInt>& Mesh::createNodeGroup(const std::string& name) { return groups_[name]; }

const std::vector<UInt>& Mesh::getNodeGroup(const std::string& name) const {
  auto it = groups_.find(name);
  if (it == groups_.end()) AKANTU_EXCEPTION("no node group " << name);
  return it->second;
}

Real Mesh::getLowerBound(UInt axis) const {
  Real v = std::numeric_limits<Real>::max();
  for (UInt n = 0; n < nodes_.size(); ++n) v = std::min(v, nodes_(n, axis));
  return v;
}
Real Mesh::getUpperBound(UInt axis) const {
  Real v = -std::numeric_limits<Real>::max();
  for (UInt n = 0; n < nodes_.size(); ++n) v = std::max(v, nodes_(n, axis));
  return v;
}

void Mesh::createGroupsFromBoundingBox(Real tol) {
  const char* names[3] = {"x", "y", "z"};
  for (UInt k = 0; k < dim_; ++k) {
    Real lo = getLowerBound(k), hi = getUpperBound(k), span = std::max(hi - lo, Real(1));
    auto& gl = groups_[std::string(names[k]) + "_min"];
    auto& gh = groups_[std::string(names[k]) + "_max"];
    gl.clear();
    gh.clear();
    for (UInt n = 0; n < nodes_.size(); ++n) {
      if (std::abs(nodes_(n, k) - lo) <= tol * span) gl.push_back(n);
      if (std::abs(nodes_(n, k) - hi) <= tol * span) gh.push_back(n);
    }
  }
}

void Mesh::createBoundaryGroup() {
  static const std::vector<std::vector<int>> facets[4] = {
      {{0, 1}, {1, 2}, {2, 0}},
      {{0, 1}, {1, 2}, {2, 3}, {3, 0}},
      {{0, 2, 1}, {1, 2, 3}, {2, 0, 3}, {0, 1, 3}},
      {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}}};
  std::map<std::vector<UInt>, int> count;
  const auto& fl = facets[type_];
  for (UInt e = 0; e < conn_.size(); ++e)
    for (auto& f : fl) {
      std::vector<UInt> key;
      for (int a : f) key.push_back(conn_(e, a));
      std::sort(key.begin(), key.end());
      count[key]++;
    }
  std::set<UInt> nodes;
  for (auto& kv : count)
    if (kv.second == 1) nodes.insert(kv.first.begin(), kv.first.end());
  groups_["boundary"].assign(nodes.begin(), nodes.end());
}

void Mesh::read(const std::string& filename) {
  std::ifstream f(filename);
  if (!f) AKANTU_EXCEPTION("cannot open mesh file " << filename);
  std::string line;
  std::vector<long> ids;
  std::vector<std::array<Real, 3>> coords;
  struct El { int type, phys; std::vector<long> nodes; };
  std::vector<El> els;
  while (std::getline(f, line)) {
    line = trim(line);
    if (line == "$Nodes") {
      std::getline(f, line);
      long n = std::stol(line);
      for (long k = 0; k < n; ++k) {
        std::getline(f, line);
        std::istringstream s(line);
        long id;
        std::array<Real, 3> c{0, 0, 0};
        s >> id >> c[0] >> c[1] >> c[2];
        ids.push_back(id);
        coords.push_back(c);
      }
    } else if (line == "$Elements") {
      std::getline(f, line);
      long n = std::stol(line);
      for (long k = 0; k < n; ++k) {
        std::getline(f, line);
        std::istringstream s(line);
        long id, t, ntags;
        s >> id >> t >> ntags;
        std::vector<long> tags(ntags);
        for (auto& x : tags) s >> x;
        El el{int(t), ntags ? int(tags[0]) : 0, {}};
        long v;
        while (s >> v) el.nodes.push_back(v);
        els.push_back(el);
      }
    }
  }
  // gmsh codes (mesh_io_msh.cc:181-199): 2 tri3, 3 quad4, 4 tet4, 5 hex8
  std::map<int, ElementType> code = {{2, _triangle_3}, {3, _quadrangle_4}, {4, _tetrahedron_4},
                                     {5, _hexahedron_8}};
  std::map<long, UInt> id_map;
  for (size_t k = 0; k < ids.size(); ++k) id_map[ids[k]] = UInt(k);
  int found = -1;
  for (auto& el : els)
    if (code.count(el.type) && akantu::getSpatialDimension(code[el.type]) == dim_) {
      if (found >= 0 && found != el.type) AKANTU_EXCEPTION("mixed element types are not supported");
      found = el.type;
    }
  if (found < 0) AKANTU_EXCEPTION("no elements of dimension " << dim_ << " in " << filename);
  type_ = code[found];
  nodes_.resize(UInt(ids.size()), dim_);
  for (size_t k = 0; k < ids.size(); ++k)
    for (UInt j = 0; j < dim_; ++j) nodes_(UInt(k), j) = coords[k][j];
  const UInt nn = getNbNodesPerElement(type_);
  std::vector<UInt> flat;
  for (auto& el : els) {
    if (el.type == found)
      for (UInt a = 0; a < nn; ++a) flat.push_back(id_map.at(el.nodes[a]));
    if (el.phys) {
      auto& g = groups_["phys_" + std::to_string(el.phys)];
      for (long v : el.nodes) g.push_back(id_map.at(v));
    }
  }
  for (auto& kv : groups_) {
    std::sort(kv.second.begin(), kv.second.end());
    kv.second.erase(std::unique(kv.second.begin(), kv.second.end()), kv.second.end());
  }
  conn_.resize(UInt(flat.size() / nn), nn);
  std::copy(flat.begin(), flat.end(), conn_.storage());
}

/* ------------------------------------------------------------------------ */
/* Materials                                                                 */
/* ------------------------------------------------------------------------ */
Material::Material(SolidMechanicsModel& model, const ID& id, int kind)
    : model_(model), id_(id), name_(id), kind_(kind) {
  params_ = {{"rho", 0.}, {"E", 0.}, {"nu", 0.}, {"Plane_Stress", 0.}};
  if (kind == AKX_MATERIAL_NEOHOOKEAN_DAMAGE) {
    params_["Sd"] = 5000.;  // material_marigo.cc:44
    params_["Yd"] = 50.;
  }
}

Real Material::getParam(const std::string& key) const {
  auto it = params_.find(key);
  if (it == params_.end()) AKANTU_EXCEPTION("material " << name_ << " has no parameter " << key);
  return it->second;
}

void Material::setParam(const std::string& key, Real value) {
  if (!params_.count(key)) AKANTU_EXCEPTION("material " << name_ << " has no parameter " << key);
  params_[key] = value;
}

void Material::setRandomParam(const std::string& key, Real base, const std::string& dist, Real p0,
                              Real p1) {
  if (!params_.count(key)) AKANTU_EXCEPTION("material " << name_ << " has no parameter " << key);
  params_[key] = base;
  random_[key] = RandomSpec{base, p0, p1, dist};
}

Real Material::getLambda() const {
  const Real E = getParam("E"), nu = getParam("nu");
  if (model_.getSpatialDimension() == 2 && getParam("Plane_Stress") != 0 && kind_ == AKX_MATERIAL_ELASTIC)
    return nu * E / ((1 + nu) * (1 - nu));
  return nu * E / ((1 + nu) * (1 - 2 * nu));
}
Real Material::getMu() const { return getParam("E") / (2 * (1 + getParam("nu"))); }
Real Material::getPushWaveSpeed() const {
  return std::sqrt((getLambda() + 2 * getMu()) / getParam("rho"));
}

UInt Material::getIndex() const { return model_.materialIndex(*this); }

const Array<Real>& Material::getStress(ElementType) {
  model_.synchronizeHost();
  if (stress_version_ != model_.stateVersion()) {
    model_.downloadPointField(AKX_STRESS, stress_, *this);
    stress_version_ = model_.stateVersion();
  }
  return stress_;
}
Array<Real>& Material::getGradU(ElementType) {
  model_.synchronizeHost();
  if (gradu_version_ != model_.stateVersion()) {
    model_.downloadPointField(AKX_GRAD_U, gradu_, *this);
    gradu_version_ = model_.stateVersion();
  }
  return gradu_;
}
const Array<Real>& Material::getDamage(ElementType) {
  model_.downloadPointField(AKX_DAMAGE, damage_, *this);
  return damage_;
}
Real Material::getPotentialEnergy() {
  if (!model_.isInitialized()) AKANTU_EXCEPTION("initFull first");
  model_.synchronizeHost();
  Real e = 0.;
  int rc = akx_material_potential_energy(model_.handle(), int(getIndex()), &e);
  if (rc) AKANTU_EXCEPTION("akx_material_potential_energy: " << akx_last_error(model_.handle()));
  return e;
}

void Material::computeStress(ElementType type, GhostType ghost_type) {
  if (ghost_type != _not_ghost) return;  // no ghost elements on one GPU
  if (!model_.isInitialized()) AKANTU_EXCEPTION("initFull first");
  const UInt dd = model_.getSpatialDimension() * model_.getSpatialDimension();
  const UInt rows = filter_.size() * getNbIntegrationPoints(type);
  if (gradu_.size() != rows || gradu_.getNbComponent() != dd) getGradU(type);
  stress_.resize(rows, dd);
  model_.synchronizeHost();
  int rc = akx_material_compute_stress(model_.handle(), int(getIndex()), gradu_.storage(),
                                       stress_.storage());
  if (rc) AKANTU_EXCEPTION("akx_material_compute_stress: " << akx_last_error(model_.handle()));
  stress_version_ = model_.stateVersion();
}

void Material::computeAllStresses(GhostType ghost_type) {
  if (ghost_type != _not_ghost) return;
  if (!model_.isInitialized()) AKANTU_EXCEPTION("initFull first");
  const UInt d = model_.getSpatialDimension();
  const UInt rows = filter_.size() * getNbIntegrationPoints(model_.getMesh().getElementType());
  gradu_.resize(rows, d * d);
  stress_.resize(rows, d * d);
  model_.synchronizeHost();
  int rc = akx_material_compute_all_stresses(model_.handle(), int(getIndex()), gradu_.storage(),
                                             stress_.storage());
  if (rc)
    AKANTU_EXCEPTION("akx_material_compute_all_stresses: " << akx_last_error(model_.handle()));
  gradu_version_ = stress_version_ = model_.stateVersion();
}

void Material::assembleInternalForces(GhostType ghost_type) {
  if (ghost_type != _not_ghost) return;
  if (!model_.isInitialized()) AKANTU_EXCEPTION("initFull first");
  model_.synchronizeHost();
  int rc = akx_material_assemble_internal_forces(model_.handle(), int(getIndex()));
  if (rc)
    AKANTU_EXCEPTION("akx_material_assemble_internal_forces: " << akx_last_error(model_.handle()));
  model_.internalForceChanged();
}
Real Material::getEnergy(const std::string& id) {
  if (id == "potential") return getPotentialEnergy();
  return 0.;
}

MaterialFactory::MaterialFactory() {
  auto make = [](int kind) {
    return [kind](const ID& id, SolidMechanicsModel& m) {
      return std::unique_ptr<Material>(new Material(m, id, kind));
    };
  };
  allocators_["elastic"] = make(AKX_MATERIAL_ELASTIC);
  allocators_["neohookean"] = make(AKX_MATERIAL_NEOHOOKEAN);
  allocators_["neohookean_damage"] = make(AKX_MATERIAL_NEOHOOKEAN_DAMAGE);
}

MaterialFactory& MaterialFactory::getInstance() {
  static MaterialFactory f;
  return f;
}

void MaterialFactory::registerAllocator(const std::string& name, Allocator alloc) {
  allocators_[name] = std::move(alloc);
}

std::unique_ptr<Material> MaterialFactory::allocate(const std::string& name, const ID& id,
                                                    SolidMechanicsModel& model) const {
  auto it = allocators_.find(name);
  if (it == allocators_.end()) AKANTU_EXCEPTION("unknown material type " << name);
  return it->second(id, model);
}

/* ------------------------------------------------------------------------ */
/* SolidMechanicsModel                                                       */
/* ------------------------------------------------------------------------ */
void SolidMechanicsModel::check(int rc, const char* what) const {
  if (rc != AKX_OK)
    AKANTU_EXCEPTION(what << " failed (" << rc << "): " << akx_last_error(ctx_));
}

SolidMechanicsModel::SolidMechanicsModel(Mesh& mesh, UInt spatial_dimension, const ID& id,
                                         int device)
    : mesh_(mesh), dim_(spatial_dimension ? spatial_dimension : mesh.getSpatialDimension()),
      id_(id) {
  if (dim_ != mesh.getSpatialDimension()) AKANTU_EXCEPTION("dimension mismatch with the mesh");
  if (mesh.getElementType() == _not_defined) AKANTU_EXCEPTION("the mesh has no elements");
  if (akx_ctx_create(device, 0, &ctx_) != AKX_OK || !ctx_)
    AKANTU_EXCEPTION("no usable HIP device " << device << " for libakx");
  const UInt N = mesh.getNbNodes(), E = mesh.getNbElement();
  check(akx_mesh_set(ctx_, int(mesh.getElementType()), int(dim_), N, mesh.getNodes().storage(), E,
                     mesh.getConnectivity(mesh.getElementType()).storage()),
        "akx_mesh_set");
  for (Array<Real>* a : {&u_, &v_, &a_, &fext_, &fint_, &mass_, &pos_, &inc_, &prev_})
    a->resize(N, dim_), a->clear();
  blocked_.resize(N, dim_);
  blocked_.clear();
  element_material_.resize(E, 1);
  element_material_.clear();
}

SolidMechanicsModel::~SolidMechanicsModel() {
  try {
    flushDumps();
  } catch (...) {
  }
  for (auto& b : dump_bufs_) akx_host_free(b.host);
  if (ctx_) akx_ctx_destroy(ctx_);
}

Material& SolidMechanicsModel::registerNewMaterial(const std::string& type,
                                                   const std::string& name) {
  auto m = MaterialFactory::getInstance().allocate(type, id_ + ":" + std::to_string(materials_.size()),
                                                   *this);
  m->setName(name);
  materials_.push_back(std::move(m));
  return *materials_.back();
}

void SolidMechanicsModel::instantiateMaterials() {
  // SolidMechanicsModel::instantiateMaterials (solid_mechanics_model_material.cc:44-107)
  for (const auto& sec : getStaticParser().getMaterialSections()) {
    Material& m = registerNewMaterial(sec.type, sec.type);
    for (const auto& kv : sec.params) {
      std::string key = kv.first, val = kv.second;
      if (key == "name") { m.setName(val); continue; }
      if (key == "finite_deformation") continue;
      std::string lower = val;
      std::transform(lower.begin(), lower.end(), lower.begin(), ::tolower);
      if (lower == "true" || lower == "false") {
        m.setParam(key, lower == "true" ? 1. : 0.);
        continue;
      }
      auto br = val.find('[');
      if (br != std::string::npos) {
        // `base dist [p0, p1]`
        std::istringstream s(val.substr(0, br));
        Real base;
        std::string dist;
        s >> base >> dist;
        std::string args = val.substr(br + 1, val.find(']') - br - 1);
        std::replace(args.begin(), args.end(), ',', ' ');
        std::istringstream as(args);
        Real p0 = 0, p1 = 0;
        as >> p0 >> p1;
        m.setRandomParam(key, base, dist, p0, p1);
        continue;
      }
      m.setParam(key, std::stod(val));
    }
  }
}

void SolidMechanicsModel::setMaterialIndex(const Array<UInt>& em) {
  if (em.size() != mesh_.getNbElement()) AKANTU_EXCEPTION("one material index per element");
  element_material_.copy(em);
  custom_assignment_ = true;
}

void SolidMechanicsModel::initFull(const AnalysisMethodArg& method) {
  if (method.value != _explicit_lumped_mass)
    AKANTU_EXCEPTION("only _explicit_lumped_mass is implemented on the GPU path");
  if (materials_.empty()) instantiateMaterials();
  if (materials_.empty()) AKANTU_EXCEPTION("no material defined (input file or registerNewMaterial)");
  const UInt E = mesh_.getNbElement(), Q = getNbIntegrationPoints(mesh_.getElementType());
  for (auto& m : materials_) m->filter_.resize(0, 1);
  std::vector<std::vector<UInt>> filt(materials_.size());
  for (UInt e = 0; e < E; ++e) {
    UInt k = element_material_(e);
    if (k >= materials_.size()) AKANTU_EXCEPTION("element " << e << " has no material " << k);
    filt[k].push_back(e);
  }
  // std::default_random_engine seeded as the reference's RandomGenerator
  // (aka_random_generator.hh:45-96); draws in element-filter then point order.
  std::default_random_engine gen(static_cast<unsigned>(getStaticParser().getSeed()));
  for (size_t k = 0; k < materials_.size(); ++k) {
    Material& m = *materials_[k];
    m.filter_.resize(UInt(filt[k].size()), 1);
    std::copy(filt[k].begin(), filt[k].end(), m.filter_.storage());
    int mid = -1;
    check(akx_material_add(ctx_, m.kind_, m.getParam("E"), m.getParam("nu"), m.getParam("rho"),
                           int(m.getParam("Plane_Stress") != 0), &mid),
          "akx_material_add");
    if (m.kind_ == AKX_MATERIAL_NEOHOOKEAN_DAMAGE) {
      std::vector<Real> yd(size_t(E) * Q, 0.);
      auto it = m.random_.find("Yd");
      for (UInt e : filt[k])
        for (UInt q = 0; q < Q; ++q) {
          Real v = m.getParam("Yd");
          if (it != m.random_.end()) {
            const auto& r = it->second;
            drawRandomValues(gen, r.dist, r.base, r.p0, r.p1, &v, 1);
          }
          yd[size_t(e) * Q + q] = v;
        }
      check(akx_material_damage(ctx_, mid, m.getParam("Sd"), yd.data()), "akx_material_damage");
    }
  }
  std::vector<int32_t> em(E);
  for (UInt e = 0; e < E; ++e) em[e] = int32_t(element_material_(e));
  check(akx_assign_materials(ctx_, em.data()), "akx_assign_materials");
  check(akx_init(ctx_), "akx_init");
  initialized_ = true;
  ++state_version_;
  for (bool& s : host_stale_) s = true;
  host_stale_[kFext] = host_stale_[kBlocked] = false;
  for (bool& d : host_dirty_) d = false;
}

static int field_of(int id) {
  static const int map[] = {AKX_DISPLACEMENT,   AKX_VELOCITY, AKX_ACCELERATION, AKX_EXTERNAL_FORCE,
                            AKX_INTERNAL_FORCE, AKX_BLOCKED_DOFS, AKX_MASS, AKX_CURRENT_POSITION,
                            AKX_DISPLACEMENT_INCREMENT, AKX_PREVIOUS_DISPLACEMENT};
  return map[id];
}

void SolidMechanicsModel::pull(NodalId id) {
  if (!host_stale_[id]) return;
  const size_t nd = size_t(mesh_.getNbNodes()) * dim_;
  if (id == kBlocked) {
    std::vector<uint8_t> b(nd);
    check(akx_field_download(ctx_, AKX_BLOCKED_DOFS, b.data(), nd), "akx_field_download");
    for (size_t k = 0; k < nd; ++k) blocked_[k] = b[k] != 0;
  } else {
    Array<Real>* arr[] = {&u_, &v_, &a_, &fext_, &fint_, nullptr, &mass_, &pos_, &inc_, &prev_};
    check(akx_field_download(ctx_, field_of(id), arr[id]->storage(), nd * 8), "akx_field_download");
  }
  host_stale_[id] = false;
}

void SolidMechanicsModel::pushDirty() {
  const size_t nd = size_t(mesh_.getNbNodes()) * dim_;
  Array<Real>* arr[] = {&u_, &v_, &a_, &fext_, &fint_};
  for (int id : {kU, kV, kA, kFext, kFint}) {
    if (!host_dirty_[id]) continue;
    check(akx_field_upload(ctx_, field_of(id), arr[id]->storage(), nd * 8), "akx_field_upload");
    host_dirty_[id] = false;
    if (id == kU) ++state_version_;
  }
  if (host_dirty_[kBlocked]) {
    std::vector<uint8_t> b(nd);
    for (size_t k = 0; k < nd; ++k) b[k] = blocked_[k] ? 1 : 0;
    check(akx_field_upload(ctx_, AKX_BLOCKED_DOFS, b.data(), nd), "akx_field_upload");
    host_dirty_[kBlocked] = false;
  }
  // a new displacement changes the current position
  host_stale_[kPos] = true;
}

Array<Real>& SolidMechanicsModel::getDisplacement() { pull(kU); host_dirty_[kU] = true; return u_; }
Array<Real>& SolidMechanicsModel::getVelocity() { pull(kV); host_dirty_[kV] = true; return v_; }
Array<Real>& SolidMechanicsModel::getAcceleration() { pull(kA); host_dirty_[kA] = true; return a_; }
Array<Real>& SolidMechanicsModel::getExternalForce() {
  pull(kFext);
  host_dirty_[kFext] = true;
  return fext_;
}
Array<bool>& SolidMechanicsModel::getBlockedDOFs() {
  pull(kBlocked);
  host_dirty_[kBlocked] = true;
  return blocked_;
}
Array<Real>& SolidMechanicsModel::getInternalForce() {
  pull(kFint);
  host_dirty_[kFint] = true;
  return fint_;
}
const Array<Real>& SolidMechanicsModel::getMass() {
  if (!initialized_) AKANTU_EXCEPTION("initFull first");
  pull(kMass);
  return mass_;
}
const Array<Real>& SolidMechanicsModel::getCurrentPosition() {
  pushDirty();
  host_stale_[kPos] = true;
  pull(kPos);
  return pos_;
}
const Array<Real>& SolidMechanicsModel::getIncrement() { pull(kInc); return inc_; }
const Array<Real>& SolidMechanicsModel::getPreviousDisplacement() { pull(kPrev); return prev_; }

Real SolidMechanicsModel::getStableTimeStep() {
  if (!initialized_) AKANTU_EXCEPTION("initFull first");
  pushDirty();
  Real dt = 0;
  check(akx_stable_time_step(ctx_, &dt), "akx_stable_time_step");
  return dt;
}

void SolidMechanicsModel::setTimeStep(Real dt) {
  check(akx_set_time_step(ctx_, dt), "akx_set_time_step");
  time_step_ = dt;
}

void SolidMechanicsModel::solveStep() { solveStep(1); }

void SolidMechanicsModel::solveStep(UInt nsteps) {
  if (!initialized_) AKANTU_EXCEPTION("initFull first");
  pushDirty();
  check(akx_solve_step(ctx_, int(nsteps), int(precision_)), "akx_solve_step");
  steps_done_ += nsteps;
  ++state_version_;
  for (int id : {kU, kV, kA, kFint, kPos, kInc, kPrev}) host_stale_[id] = true;
}

void SolidMechanicsModel::assembleInternalForces() {
  if (!initialized_) AKANTU_EXCEPTION("initFull first");
  pushDirty();
  check(akx_assemble_internal_forces(ctx_, int(precision_)), "akx_assemble_internal_forces");
  host_stale_[kFint] = true;
  ++state_version_;
}

Real SolidMechanicsModel::getEnergy(const std::string& id) {
  if (!initialized_) AKANTU_EXCEPTION("initFull first");
  pushDirty();
  int kind = id == "kinetic" ? AKX_ENERGY_KINETIC
             : id == "potential" ? AKX_ENERGY_POTENTIAL
             : id == "external work" ? AKX_ENERGY_EXTERNAL_WORK : -1;
  if (kind < 0) AKANTU_EXCEPTION("unknown energy " << id);
  Real v = 0;
  check(akx_energy(ctx_, kind, &v), "akx_energy");
  return v;
}

void SolidMechanicsModel::applyBC(const BC::Dirichlet::FixedValue& f, const std::string& group) {
  auto& u = getDisplacement();
  auto& b = getBlockedDOFs();
  for (UInt n : mesh_.getNodeGroup(group)) {
    b(n, f.axis) = true;
    u(n, f.axis) = f.value;
  }
}

void SolidMechanicsModel::applyBC(const BC::Dirichlet::FlagOnly& f, const std::string& group) {
  auto& b = getBlockedDOFs();
  for (UInt n : mesh_.getNodeGroup(group)) b(n, f.axis) = true;
}

void SolidMechanicsModel::applyBC(const BC::Dirichlet::IncrementValue& f,
                                  const std::string& group) {
  auto& u = getDisplacement();
  auto& b = getBlockedDOFs();
  for (UInt n : mesh_.getNodeGroup(group)) {
    b(n, f.axis) = true;
    u(n, f.axis) += f.value;
  }
}

void SolidMechanicsModel::applyBC(const BC::Neumann::FromTraction& f, const std::string& group) {
  if (f.traction.size() != dim_) AKANTU_EXCEPTION("traction must have " << dim_ << " components");
  auto& fe = getExternalForce();
  for (UInt n : mesh_.getNodeGroup(group))
    for (UInt i = 0; i < dim_; ++i) fe(n, i) += f.traction[i];
}

void SolidMechanicsModel::setDirichletIncrement(const std::string& group, SpatialDirection axis,
                                                Real increment) {
  pushDirty();
  const auto& g = mesh_.getNodeGroup(group);
  std::vector<int64_t> nodes(g.begin(), g.end());
  check(akx_set_dirichlet_increment(ctx_, int64_t(nodes.size()), nodes.data(), int(axis), increment),
        "akx_set_dirichlet_increment");
  host_stale_[kBlocked] = true;
}

UInt SolidMechanicsModel::materialIndex(const Material& m) const {
  for (size_t k = 0; k < materials_.size(); ++k)
    if (materials_[k].get() == &m) return UInt(k);
  AKANTU_EXCEPTION("material " << m.getName() << " does not belong to this model");
}

FEEngine& SolidMechanicsModel::getFEEngine(const ID&) {
  if (!fe_engine_) fe_engine_.reset(new FEEngine(*this));
  return *fe_engine_;
}

Material& SolidMechanicsModel::getMaterial(UInt i) {
  if (i >= materials_.size()) AKANTU_EXCEPTION("no material " << i);
  return *materials_[i];
}

Material& SolidMechanicsModel::getMaterial(const std::string& name) {
  for (auto& m : materials_)
    if (m->getName() == name) return *m;
  AKANTU_EXCEPTION("no material named " << name);
}

void SolidMechanicsModel::autotune(UInt nsteps) {
  if (!initialized_) AKANTU_EXCEPTION("initFull first");
  pushDirty();
  check(akx_autotune(ctx_, int(precision_), int(nsteps), nullptr, nullptr), "akx_autotune");
}

void SolidMechanicsModel::getLastStepTiming(Real& element_ms, Real& node_ms) {
  check(akx_last_step_timing(ctx_, &element_ms, &node_ms), "akx_last_step_timing");
}

/* ------------------------------------------------------------------------ */
/* ParaView output                                                           */
/* ------------------------------------------------------------------------ */
void SolidMechanicsModel::setBaseName(const std::string& name, const std::string& directory) {
  flushDumps();
  dump_base_ = name;
  dump_dir_ = directory;
  dumps_.clear();
}

void SolidMechanicsModel::addDumpField(const std::string& field) {
  static const char* known[] = {"displacement", "velocity", "acceleration", "external_force",
                                "internal_force", "mass", "blocked_dofs", "stress", "grad_u",
                                "damage"};
  bool ok = false;
  for (const char* k : known) ok = ok || field == k;
  if (!ok) AKANTU_EXCEPTION("unknown dump field " << field);
  if (std::find(dump_fields_.begin(), dump_fields_.end(), field) == dump_fields_.end())
    dump_fields_.push_back(field);
}

namespace {
// VTK inline binary data: base64 of a UInt64 byte count followed by the raw
// little-endian values, one stream per DataArray (header_type="UInt64").
std::string vtk_base64(const std::vector<unsigned char>& raw) {
  static const char* T = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  std::vector<unsigned char> b(8 + raw.size());
  const uint64_t n = raw.size();
  std::memcpy(b.data(), &n, 8);
  std::memcpy(b.data() + 8, raw.data(), raw.size());
  std::string out;
  out.reserve((b.size() + 2) / 3 * 4);
  size_t i = 0;
  for (; i + 2 < b.size(); i += 3) {
    const uint32_t v = uint32_t(b[i]) << 16 | uint32_t(b[i + 1]) << 8 | b[i + 2];
    out += T[v >> 18 & 63]; out += T[v >> 12 & 63]; out += T[v >> 6 & 63]; out += T[v & 63];
  }
  if (i < b.size()) {
    const bool two = i + 1 < b.size();
    const uint32_t v = uint32_t(b[i]) << 16 | (two ? uint32_t(b[i + 1]) << 8 : 0u);
    out += T[v >> 18 & 63]; out += T[v >> 12 & 63];
    out += two ? T[v >> 6 & 63] : '=';
    out += '=';
  }
  return out;
}

template <typename V>
void vtu_binary(std::ostream& o, const char* type, const std::string& name, UInt ncomp,
                const std::vector<V>& v) {
  std::vector<unsigned char> raw(v.size() * sizeof(V));
  std::memcpy(raw.data(), v.data(), raw.size());
  o << "<DataArray type=\"" << type << "\" Name=\"" << name << "\" NumberOfComponents=\""
    << ncomp << "\" format=\"binary\">\n" << vtk_base64(raw) << "\n</DataArray>\n";
}

void vtu_array(std::ostream& o, const std::string& name, const Real* v, size_t count,
               UInt ncomp, UInt pad_to) {
  std::vector<double> p(count * pad_to, 0.);
  for (size_t k = 0; k < count; ++k)
    for (UInt c = 0; c < ncomp; ++c) p[k * pad_to + c] = v[k * ncomp + c];
  vtu_binary(o, "Float64", name, pad_to, p);
}
}  // namespace

void SolidMechanicsModel::flushDumps() {
  if (dump_job_.valid()) dump_job_.get();
}

void SolidMechanicsModel::dump() {
  if (!initialized_) AKANTU_EXCEPTION("initFull first");
  flushDumps();  // the previous dump's buffers are free again
  pushDirty();
  const ElementType t = mesh_.getElementType();
  const UInt N = mesh_.getNbNodes(), E = mesh_.getNbElement(), n = getNbNodesPerElement(t);
  const UInt Q = getNbIntegrationPoints(t), d = dim_;
  static const std::map<ElementType, int> vtk = {
      {_triangle_3, 5}, {_quadrangle_4, 9}, {_tetrahedron_4, 10}, {_hexahedron_8, 12}};
  char num[16];
  std::snprintf(num, sizeof(num), "%04u", unsigned(dumps_.size()));
  const std::string file = dump_base_ + "_" + num + ".vtu";
  // snapshots of every field, enqueued behind the steps so far; the copies
  // to the page-locked buffers overlap the steps that follow
  struct Item {
    std::string name;
    bool point, bytes_u8;
    UInt comp;
    const void* host;
    uint64_t ticket;
  };
  std::vector<Item> items;
  dump_bufs_.resize(dump_fields_.size());
  for (size_t k = 0; k < dump_fields_.size(); ++k) {
    const std::string& f = dump_fields_[k];
    int id = -1;
    bool point = false;
    if (f == "displacement") id = AKX_DISPLACEMENT;
    else if (f == "velocity") id = AKX_VELOCITY;
    else if (f == "acceleration") id = AKX_ACCELERATION;
    else if (f == "external_force") id = AKX_EXTERNAL_FORCE;
    else if (f == "internal_force") id = AKX_INTERNAL_FORCE;
    else if (f == "mass") id = AKX_MASS;
    else if (f == "blocked_dofs") id = AKX_BLOCKED_DOFS;
    else if (f == "stress") { id = AKX_STRESS; point = true; }
    else if (f == "grad_u") { id = AKX_GRAD_U; point = true; }
    else if (f == "damage") { id = AKX_DAMAGE; point = true; }
    const bool u8 = id == AKX_BLOCKED_DOFS;
    const UInt comp = id == AKX_DAMAGE ? 1 : point ? d * d : d;
    const size_t bytes = (point ? size_t(E) * Q : size_t(N)) * comp * (u8 ? 1 : 8);
    DumpBuffer& b = dump_bufs_[k];
    if (b.bytes < bytes) {
      akx_host_free(b.host);
      b = DumpBuffer{};
      check(akx_host_alloc(bytes, &b.host), "akx_host_alloc");
      b.bytes = bytes;
    }
    uint64_t ticket = 0;
    check(akx_field_download_async(ctx_, id, b.host, bytes, &ticket), "akx_field_download_async");
    items.push_back({f, point, u8, comp, b.host, ticket});
  }
  dumps_.push_back({Real(steps_done_) * time_step_, file});
  const auto entries = dumps_;
  const std::string dir = dump_dir_, pvd_name = dump_base_ + ".pvd";
  akx_ctx* ctx = ctx_;
  const Mesh* mesh = &mesh_;
  const int cell = vtk.at(t);
  dump_job_ = std::async(std::launch::async, [=]() {
    for (const auto& it : items)
      if (akx_download_wait(ctx, it.ticket) != AKX_OK)
        AKANTU_EXCEPTION("akx_download_wait failed for " << it.name);
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    std::ofstream o(dir + "/" + file);
    if (!o) AKANTU_EXCEPTION("cannot write " << dir << "/" << file);
    o.precision(17);
    o << "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
         "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n<UnstructuredGrid>\n"
         "<Piece NumberOfPoints=\""
      << N << "\" NumberOfCells=\"" << E << "\">\n<Points>\n";
    vtu_array(o, "positions", mesh->getNodes().storage(), N, d, 3);
    o << "</Points>\n<Cells>\n";
    const auto& conn = mesh->getConnectivity(t);
    std::vector<int64_t> cv(size_t(E) * n), ov(E);
    for (UInt e = 0; e < E; ++e) {
      for (UInt a = 0; a < n; ++a) cv[size_t(e) * n + a] = int64_t(conn(e, a));
      ov[e] = int64_t(e + 1) * n;
    }
    vtu_binary(o, "Int64", "connectivity", 1, cv);
    vtu_binary(o, "Int64", "offsets", 1, ov);
    vtu_binary(o, "UInt8", "types", 1, std::vector<uint8_t>(E, uint8_t(cell)));
    o << "</Cells>\n<PointData>\n";
    for (const auto& it : items) {
      if (it.point) continue;
      if (it.bytes_u8) {
        const auto* b = static_cast<const uint8_t*>(it.host);
        std::vector<Real> v(size_t(N) * d);
        for (size_t k = 0; k < v.size(); ++k) v[k] = b[k] ? 1. : 0.;
        vtu_array(o, it.name, v.data(), N, d, d == 2 ? 3 : d);
      } else {
        vtu_array(o, it.name, static_cast<const Real*>(it.host), N, d, d == 2 ? 3 : d);
      }
    }
    o << "</PointData>\n<CellData>\n";
    for (const auto& it : items) {
      if (!it.point) continue;
      const auto* all = static_cast<const Real*>(it.host);
      std::vector<Real> avg(size_t(E) * it.comp, 0.);
      for (UInt e = 0; e < E; ++e)
        for (UInt q = 0; q < Q; ++q)
          for (UInt c = 0; c < it.comp; ++c)
            avg[size_t(e) * it.comp + c] += all[(size_t(e) * Q + q) * it.comp + c] / Q;
      vtu_array(o, it.name, avg.data(), E, it.comp, it.comp);
    }
    o << "</CellData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n";
    std::ofstream pvd(dir + "/" + pvd_name);
    pvd.precision(17);
    pvd << "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\">\n<Collection>\n";
    for (const auto& dd : entries)
      pvd << "<DataSet timestep=\"" << dd.first << "\" group=\"\" part=\"0\" file=\"" << dd.second
          << "\"/>\n";
    pvd << "</Collection>\n</VTKFile>\n";
  });
}

void SolidMechanicsModel::downloadPointField(int field, Array<Real>& out, const Material& m) {
  if (!initialized_) AKANTU_EXCEPTION("initFull first");
  pushDirty();
  const ElementType t = mesh_.getElementType();
  const UInt E = mesh_.getNbElement(), Q = getNbIntegrationPoints(t);
  const UInt comp = field == AKX_DAMAGE ? 1 : dim_ * dim_;
  std::vector<Real> all(size_t(E) * Q * comp);
  check(akx_field_download(ctx_, field, all.data(), all.size() * 8), "akx_field_download");
  const auto& f = m.getElementFilter();
  out.resize(f.size() * Q, comp);
  for (UInt k = 0; k < f.size(); ++k)
    std::memcpy(out.storage() + size_t(k) * Q * comp, all.data() + size_t(f(k)) * Q * comp,
                sizeof(Real) * Q * comp);
}

/* ------------------------------------------------------------------------ */
/* FEEngine (fe_engine.hh:75-251) over the akx_fe_* operators                */
/* ------------------------------------------------------------------------ */
const Array<UInt> empty_filter;

namespace {
const uint32_t* filter_ptr(const Array<UInt>& f) {
  static_assert(sizeof(UInt) == sizeof(uint32_t), "element ids are 32-bit");
  return &f == &empty_filter ? nullptr : reinterpret_cast<const uint32_t*>(f.storage());
}
int64_t filter_size(const Array<UInt>& f) { return &f == &empty_filter ? 0 : int64_t(f.size()); }
}  // namespace

void FEEngine::checkType(const ElementType& type, const GhostType& ghost_type) const {
  if (!model_.isInitialized()) AKANTU_EXCEPTION("FEEngine: initFull the model first");
  if (type != model_.getMesh().getElementType())
    AKANTU_EXCEPTION("FEEngine: the mesh has no elements of type " << int(type));
  if (ghost_type != _not_ghost) AKANTU_EXCEPTION("FEEngine: no ghost elements on one GPU");
}

void FEEngine::initShapeFunctions(const GhostType& ghost_type) {
  checkType(model_.getMesh().getElementType(), ghost_type);
}

UInt FEEngine::getNbIntegrationPoints(const ElementType& type, const GhostType&) const {
  return akantu::getNbIntegrationPoints(type);
}

const Array<Real>& FEEngine::getShapesDerivatives(const ElementType& type,
                                                  const GhostType& ghost_type, UInt) {
  checkType(type, ghost_type);
  const UInt rows = model_.getMesh().getNbElement() * getNbIntegrationPoints(type);
  const UInt comp = getSpatialDimension(type) * getNbNodesPerElement(type);
  shapes_derivatives_.resize(rows, comp);
  int rc = akx_field_download(model_.handle(), AKX_SHAPES_DERIVATIVES, shapes_derivatives_.storage(),
                              size_t(rows) * comp * 8);
  if (rc) AKANTU_EXCEPTION("akx_field_download: " << akx_last_error(model_.handle()));
  return shapes_derivatives_;
}

const Array<Real>& FEEngine::getShapes(const ElementType& type, const GhostType& ghost_type,
                                       UInt) {
  checkType(type, ghost_type);
  // N_a at the points: the interpolation of each node's indicator field
  const UInt E = model_.getMesh().getNbElement(), Q = getNbIntegrationPoints(type);
  const UInt n = getNbNodesPerElement(type), N = model_.getMesh().getNbNodes();
  const auto& conn = model_.getMesh().getConnectivity(type);
  shapes_.resize(E * Q, n);
  Array<Real> ind(N, 1), nq;
  Array<UInt> one(1, 1);
  for (UInt e = 0; e < E; ++e) {
    one(0) = e;
    for (UInt a = 0; a < n; ++a) {
      ind.clear();
      ind(conn(e, a)) = 1.;
      interpolateOnIntegrationPoints(ind, nq, 1, type, ghost_type, one);
      for (UInt q = 0; q < Q; ++q) shapes_(e * Q + q, a) = nq(q);
    }
  }
  return shapes_;
}

#define AKX_FE_CHECK(call)                                                           \
  do {                                                                               \
    int _rc = (call);                                                                \
    if (_rc) AKANTU_EXCEPTION(#call << ": " << akx_last_error(model_.handle()));    \
  } while (0)

void FEEngine::gradientOnIntegrationPoints(const Array<Real>& u, Array<Real>& nablauq,
                                           UInt nb_dof, const ElementType& type,
                                           const GhostType& ghost_type,
                                           const Array<UInt>& filter) const {
  checkType(type, ghost_type);
  const UInt ne = filter_size(filter) ? filter.size() : model_.getMesh().getNbElement();
  nablauq.resize(ne * getNbIntegrationPoints(type), nb_dof * getSpatialDimension(type));
  AKX_FE_CHECK(akx_fe_gradient_on_integration_points(model_.handle(), u.storage(), int(nb_dof),
                                                     filter_ptr(filter), filter_size(filter),
                                                     nablauq.storage()));
}

void FEEngine::interpolateOnIntegrationPoints(const Array<Real>& u, Array<Real>& uq, UInt nb_dof,
                                              const ElementType& type,
                                              const GhostType& ghost_type,
                                              const Array<UInt>& filter) const {
  checkType(type, ghost_type);
  const UInt ne = filter_size(filter) ? filter.size() : model_.getMesh().getNbElement();
  uq.resize(ne * getNbIntegrationPoints(type), nb_dof);
  AKX_FE_CHECK(akx_fe_interpolate_on_integration_points(model_.handle(), u.storage(), int(nb_dof),
                                                        filter_ptr(filter), filter_size(filter),
                                                        uq.storage()));
}

void FEEngine::computeBtD(const Array<Real>& Ds, Array<Real>& BtDs, const ElementType& type,
                          const GhostType& ghost_type, const Array<UInt>& filter) const {
  checkType(type, ghost_type);
  const UInt d = getSpatialDimension(type), rows = Ds.getNbComponent() / d;
  BtDs.resize(Ds.size(), rows * getNbNodesPerElement(type));
  AKX_FE_CHECK(akx_fe_compute_btd(model_.handle(), Ds.storage(), int(rows), filter_ptr(filter),
                                  filter_size(filter), BtDs.storage()));
}

void FEEngine::computeNtb(const Array<Real>& bs, Array<Real>& Ntbs, const ElementType& type,
                          const GhostType& ghost_type, const Array<UInt>& filter) const {
  checkType(type, ghost_type);
  Ntbs.resize(bs.size(), bs.getNbComponent() * getNbNodesPerElement(type));
  AKX_FE_CHECK(akx_fe_compute_ntb(model_.handle(), bs.storage(), int(bs.getNbComponent()),
                                  filter_ptr(filter), filter_size(filter), Ntbs.storage()));
}

void FEEngine::integrate(const Array<Real>& f, Array<Real>& intf, UInt nb_dof,
                         const ElementType& type, const GhostType& ghost_type,
                         const Array<UInt>& filter) const {
  checkType(type, ghost_type);
  intf.resize(f.size() / getNbIntegrationPoints(type), nb_dof);
  AKX_FE_CHECK(akx_fe_integrate(model_.handle(), f.storage(), int(nb_dof), filter_ptr(filter),
                                filter_size(filter), intf.storage()));
}

Real FEEngine::integrate(const Array<Real>& f, const ElementType& type,
                         const GhostType& ghost_type, const Array<UInt>& filter) const {
  checkType(type, ghost_type);
  Real v = 0.;
  AKX_FE_CHECK(akx_fe_integrate_scalar(model_.handle(), f.storage(), filter_ptr(filter),
                                       filter_size(filter), &v));
  return v;
}

void FEEngine::integrateOnIntegrationPoints(const Array<Real>& f, Array<Real>& intf, UInt nb_dof,
                                            const ElementType& type, const GhostType& ghost_type,
                                            const Array<UInt>& filter) const {
  checkType(type, ghost_type);
  intf.resize(f.size(), nb_dof);
  AKX_FE_CHECK(akx_fe_integrate_on_integration_points(model_.handle(), f.storage(), int(nb_dof),
                                                      filter_ptr(filter), filter_size(filter),
                                                      intf.storage()));
}

void FEEngine::assembleFieldLumped(
    const std::function<void(Matrix<Real>&, const Element&)>& field_funct, UInt nb_dof,
    Array<Real>& lumped, ElementType type, const GhostType& ghost_type) const {
  checkType(type, ghost_type);
  const UInt E = model_.getMesh().getNbElement(), Q = getNbIntegrationPoints(type);
  // fillField (fe_engine_tmpl.hh): one nb_dof x Q matrix per element
  Array<Real> field(E * Q, nb_dof);
  for (UInt e = 0; e < E; ++e) {
    Matrix<Real> m(field.storage() + size_t(e) * Q * nb_dof, nb_dof, Q);
    field_funct(m, Element{type, e, ghost_type});
  }
  if (lumped.size() != model_.getMesh().getNbNodes() || lumped.getNbComponent() != nb_dof)
    lumped.resize(model_.getMesh().getNbNodes(), nb_dof);
  AKX_FE_CHECK(akx_fe_assemble_field_lumped(model_.handle(), field.storage(), int(nb_dof),
                                            lumped.storage()));
}
#undef AKX_FE_CHECK

}  // namespace akantu

extern "C" int akantu_random_values(long seed, const char* dist, double base, double p0,
                                    double p1, int64_t n, double* out) {
  std::default_random_engine engine(static_cast<unsigned>(seed));
  try {
    akantu::drawRandomValues(engine, dist ? dist : "", base, p0, p1, out, size_t(n));
  } catch (akantu::debug::Exception&) {
    return -1;
  }
  return 0;
}
