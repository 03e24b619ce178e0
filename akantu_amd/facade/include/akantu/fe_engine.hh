/**
 * fe_engine.hh -- FEEngine of the façade: the operators of
 * src/fe_engine/fe_engine.hh:75-251 for the mesh of a SolidMechanicsModel
 * (FEEngineTemplate<IntegratorGauss, ShapeLagrange> of its element type),
 * computed on the GPU by libakx.so in FP64 with the reference's operation
 * order.
 *
 * Arrays keep the reference's layouts: nodal fields (nb_nodes, nb_dof),
 * per-point fields (nb_elements * nb_points, components) with the element's
 * points consecutive, per-point matrices column-major.  `filter_elements`
 * lists element ids (empty_filter: every element); per-element and per-point
 * arrays then follow that list.  One element type per mesh, no ghosts: the
 * GhostType arguments accept _not_ghost only.
 */
#pragma once

#include <functional>

#include "aka_array.hh"
#include "aka_common.hh"

namespace akantu {

class SolidMechanicsModel;

enum GhostType { _not_ghost = 0, _ghost = 1 };

/// an element of the mesh (src/mesh/element.hh:40-70)
struct Element {
  ElementType type;
  UInt element;
  GhostType ghost_type;
};

/// column-major dense matrix view of a block of values (aka_types.hh:879)
template <typename T>
class Matrix {
 public:
  Matrix(T* data, UInt rows, UInt cols) : data_(data), rows_(rows), cols_(cols) {}
  T& operator()(UInt i, UInt j) { return data_[i + size_t(j) * rows_]; }
  const T& operator()(UInt i, UInt j) const { return data_[i + size_t(j) * rows_]; }
  UInt rows() const { return rows_; }
  UInt cols() const { return cols_; }
  void set(const T& v) { std::fill(data_, data_ + size_t(rows_) * cols_, v); }

 private:
  T* data_;
  UInt rows_, cols_;
};

/// "no filter": every element of the type (fe_engine.hh:37, empty_filter)
extern const Array<UInt> empty_filter;

class FEEngine {
 public:
  explicit FEEngine(SolidMechanicsModel& model) : model_(model) {}

  /// shape functions are precomputed by SolidMechanicsModel::initFull
  void initShapeFunctions(const GhostType& ghost_type = _not_ghost);

  UInt getNbIntegrationPoints(const ElementType& type,
                              const GhostType& ghost_type = _not_ghost) const;
  /// dN/dX of every point, (nb_elements * nb_points, d * n), column-major d x n
  const Array<Real>& getShapesDerivatives(const ElementType& type,
                                          const GhostType& ghost_type = _not_ghost,
                                          UInt id = 0);
  /// N of every point, (nb_elements * nb_points, n)
  const Array<Real>& getShapes(const ElementType& type, const GhostType& ghost_type = _not_ghost,
                               UInt id = 0);

  void gradientOnIntegrationPoints(const Array<Real>& u, Array<Real>& nablauq,
                                   UInt nb_degree_of_freedom, const ElementType& type,
                                   const GhostType& ghost_type = _not_ghost,
                                   const Array<UInt>& filter_elements = empty_filter) const;
  void interpolateOnIntegrationPoints(const Array<Real>& u, Array<Real>& uq,
                                      UInt nb_degree_of_freedom, const ElementType& type,
                                      const GhostType& ghost_type = _not_ghost,
                                      const Array<UInt>& filter_elements = empty_filter) const;
  void computeBtD(const Array<Real>& Ds, Array<Real>& BtDs, const ElementType& type,
                  const GhostType& ghost_type = _not_ghost,
                  const Array<UInt>& filter_elements = empty_filter) const;
  void computeNtb(const Array<Real>& bs, Array<Real>& Ntbs, const ElementType& type,
                  const GhostType& ghost_type = _not_ghost,
                  const Array<UInt>& filter_elements = empty_filter) const;
  void integrate(const Array<Real>& f, Array<Real>& intf, UInt nb_degree_of_freedom,
                 const ElementType& type, const GhostType& ghost_type = _not_ghost,
                 const Array<UInt>& filter_elements = empty_filter) const;
  Real integrate(const Array<Real>& f, const ElementType& type,
                 const GhostType& ghost_type = _not_ghost,
                 const Array<UInt>& filter_elements = empty_filter) const;
  void integrateOnIntegrationPoints(const Array<Real>& f, Array<Real>& intf,
                                    UInt nb_degree_of_freedom, const ElementType& type,
                                    const GhostType& ghost_type = _not_ghost,
                                    const Array<UInt>& filter_elements = empty_filter) const;

  /// assembles the lumped version of int N^t field N (row sum,
  /// fe_engine_template_tmpl_field.hh:126-199) into `lumped` (nb_nodes,
  /// nb_dof), adding to its values; field_funct fills the nb_dof x nb_points
  /// matrix of an element (fe_engine_tmpl.hh, fillField).  The reference
  /// assembles into a DOFManager's lumped matrix; this one takes the array.
  void assembleFieldLumped(const std::function<void(Matrix<Real>&, const Element&)>& field_funct,
                           UInt nb_degree_of_freedom, Array<Real>& lumped, ElementType type,
                           const GhostType& ghost_type = _not_ghost) const;

 private:
  void checkType(const ElementType& type, const GhostType& ghost_type) const;
  SolidMechanicsModel& model_;
  Array<Real> shapes_derivatives_, shapes_;
};

}  // namespace akantu
