/**
 * mesh.hh -- Mesh of one element type: nodes (N x d), connectivity (E x n),
 * node groups.  Mirrors the parts of src/mesh/mesh.hh:100-611 a driver of the
 * explicit path uses, plus Mesh::read for gmsh 2.x files
 * (src/io/mesh_io/mesh_io_msh.cc) and a MeshAccessor-like filler for meshes
 * built in code.
 */
#pragma once

#include <map>
#include <string>
#include <vector>

#include "aka_array.hh"
#include "aka_common.hh"

namespace akantu {

class Mesh {
 public:
  explicit Mesh(UInt spatial_dimension, const ID& id = "mesh");

  /// gmsh 2.x ASCII reader (elements of this mesh's dimension; physical tags
  /// become node groups "phys_<tag>").
  void read(const std::string& filename);

  /// fill from arrays (what test_model_solver_dynamic.cc does via MeshAccessor)
  void setNodes(const Array<Real>& nodes);
  void setConnectivity(ElementType type, const Array<UInt>& conn);

  UInt getSpatialDimension() const { return dim_; }
  UInt getNbNodes() const { return nodes_.size(); }
  UInt getNbElement(ElementType type) const;
  UInt getNbElement() const { return conn_.size(); }
  ElementType getElementType() const { return type_; }
  const Array<Real>& getNodes() const { return nodes_; }
  const Array<UInt>& getConnectivity(ElementType type) const;

  /// node groups (NodeGroup of the reference, src/mesh/node_group.hh)
  std::vector<UInt>& createNodeGroup(const std::string& name);
  const std::vector<UInt>& getNodeGroup(const std::string& name) const;
  bool hasNodeGroup(const std::string& name) const { return groups_.count(name) > 0; }
  /// groups "x_min", "x_max", "y_min", ... of the nodes on the bounding box
  /// faces (the geometric part of Mesh::createBoundaryGroupFromGeometry).
  void createGroupsFromBoundingBox(Real tolerance = 1e-10);
  /// all nodes on the boundary facets, group "boundary"
  void createBoundaryGroup();

  /// lower / upper corner of the bounding box
  Real getLowerBound(UInt axis) const;
  Real getUpperBound(UInt axis) const;

 private:
  UInt dim_;
  ID id_;
  ElementType type_ = _not_defined;
  Array<Real> nodes_;
  Array<UInt> conn_;
  std::map<std::string, std::vector<UInt>> groups_;
};

}  // namespace akantu
