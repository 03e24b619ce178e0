/**
 * aka_common.hh -- basic types of the akantu:: façade over libakx.so.
 *
 * Mirrors the names the reference exposes to a driver of the explicit path:
 * Real / UInt (src/common/aka_common.hh, aka_config.hh.in:36-38), ElementType
 * (aka_element_classes_info.hh.in), AnalysisMethod (model_options.hh),
 * SpatialDirection (_x, _y, _z), debug::Exception (aka_error.hh:100), and the
 * named argument `_analysis_method = ...` of Model::initFull (model.cc:61-95).
 */
#pragma once

#include <cstdint>
#include <exception>
#include <sstream>
#include <string>

namespace akantu {

using Real = double;
using UInt = unsigned int;
using Int = int;
using ID = std::string;

enum ElementType {
  _not_defined = -1,
  _triangle_3 = 0,
  _quadrangle_4 = 1,
  _tetrahedron_4 = 2,
  _hexahedron_8 = 3
};

enum AnalysisMethod { _static = 0, _implicit_dynamic = 1, _explicit_lumped_mass = 2 };

enum SpatialDirection { _x = 0, _y = 1, _z = 2 };

/// Arithmetic of the element kernels (extension: the reference has one Real type).
enum class Precision { fp64 = 0, mixed = 1 };

namespace debug {
/// Error type thrown by every failing call (debug::Exception in the reference).
class Exception : public std::exception {
 public:
  explicit Exception(std::string info) : info_(std::move(info)) {}
  const char* what() const noexcept override { return info_.c_str(); }

 private:
  std::string info_;
};
}  // namespace debug

#define AKANTU_EXCEPTION(info)                      \
  do {                                              \
    std::stringstream _s;                           \
    _s << info;                                     \
    throw ::akantu::debug::Exception(_s.str());     \
  } while (0)

/// `model.initFull(_analysis_method = _explicit_lumped_mass)`
struct AnalysisMethodArg {
  AnalysisMethod value = _explicit_lumped_mass;
};
struct AnalysisMethodTag {
  AnalysisMethodArg operator=(AnalysisMethod m) const { return AnalysisMethodArg{m}; }
};
static const AnalysisMethodTag _analysis_method{};

UInt getSpatialDimension(ElementType type);
UInt getNbNodesPerElement(ElementType type);
UInt getNbIntegrationPoints(ElementType type);

/// akantu::initialize: parses the input file (material sections) and the
/// command line option --aka_seed (src/common/aka_common.cc:85-113).
void initialize(const std::string& input_file, int& argc, char**& argv);
void initialize(int& argc, char**& argv);
void finalize();

}  // namespace akantu
