/**
 * boundary_condition_functor.hh -- the boundary-condition functors a driver of
 * the explicit path uses (src/model/common/boundary_condition/
 * boundary_condition_functor.hh:78-205): Dirichlet FixedValue, FlagOnly,
 * IncrementValue, and a Neumann traction on a node group.
 */
#pragma once

#include <vector>

#include "aka_common.hh"

namespace akantu {
namespace BC {
namespace Dirichlet {

/// blocks `axis` and sets the displacement to `value`
struct FixedValue {
  FixedValue(Real val, SpatialDirection ax = _x) : value(val), axis(ax) {}
  Real value;
  SpatialDirection axis;
};

/// blocks `axis` only
struct FlagOnly {
  explicit FlagOnly(SpatialDirection ax = _x) : axis(ax) {}
  SpatialDirection axis;
};

/// blocks `axis` and adds `value` to the displacement
struct IncrementValue {
  IncrementValue(Real val, SpatialDirection ax = _x) : value(val), axis(ax) {}
  void setIncrement(Real val) { value = val; }
  Real value;
  SpatialDirection axis;
};

}  // namespace Dirichlet

namespace Neumann {

/// nodal force `traction` added to the external force of every node of the
/// group (the lumped form of a constant traction)
struct FromTraction {
  explicit FromTraction(std::vector<Real> t) : traction(std::move(t)) {}
  std::vector<Real> traction;
};

}  // namespace Neumann
}  // namespace BC
}  // namespace akantu
