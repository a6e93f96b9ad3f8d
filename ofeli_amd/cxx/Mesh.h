// Mesh.h -- the mesh data the solve path needs (counterpart: include/mesh/Mesh.h).
//
// Structured constructors with the reference's node order, coordinates and
// boundary codes (Mesh.cpp:547-676 for rectangles; Mesh.cpp:678-726 node
// layout for boxes, with a positively oriented Kuhn tetrahedral split), DOF
// numbering with imposed-DOF elimination (Mesh.cpp:1483-1527,
// removeImposedDOF Mesh.h:420) and the boundary sides of rectangles.
#pragma once

#include <cstdint>
#include <vector>

#include "OFELI_Config.h"

namespace OFELI {

class Mesh {
 public:
  Mesh() = default;

  // Mesh(xmin,xmax,ymin,ymax,nx,ny,cx0,cxN,cy0,cyN,opt,nb_dof): opt is
  // TRIANGLE or QUADRILATERAL.  Negative codes create boundary sides.
  Mesh(real_t xmin, real_t xmax, real_t ymin, real_t ymax, size_t nx, size_t ny, int cx0,
       int cxN, int cy0, int cyN, int opt = TRIANGLE, size_t nb_dof = 1)
      : _dim(2), _nb_dof((int)nb_dof), _nx((int)nx), _ny((int)ny) {
    const bool quad = opt == QUADRILATERAL;
    if (!quad && opt != TRIANGLE) throw OFELIException("Mesh::Mesh: illegal option");
    _nen = quad ? 4 : 3;
    const size_t nn = (nx + 1) * (ny + 1), ne = quad ? nx * ny : 2 * nx * ny;
    _coord.resize(2 * nn);
    _code.resize(nn * nb_dof);
    _elem.resize(ne * _nen);
    hip_check(ofh_mesh2d_structured(xmin, xmax, ymin, ymax, (int)nx, (int)ny, cx0, cxN, cy0, cyN,
                                    quad, (int)nb_dof, _coord.data(), _code.data(), _elem.data()),
              "Mesh::Mesh");
    // boundary sides with negative codes (Mesh.cpp:617-669)
    auto add_side = [&](int64_t a, int64_t b, int c) {
      _sides.push_back((int32_t)a);
      _sides.push_back((int32_t)b);
      _side_code.push_back(c);
    };
    int64_t n = 1;
    if (cx0 < 0)
      for (size_t j = 1; j <= ny; ++j, ++n) add_side(n, n + 1, -cx0);
    n = (int64_t)nx * (ny + 1) + 1;
    if (cxN < 0)
      for (size_t j = 1; j <= ny; ++j, ++n) add_side(n, n + 1, -cxN);
    n = 1;
    if (cy0 < 0)
      for (size_t i = 1; i <= nx; ++i, n += ny + 1) add_side(n, n + ny + 1, -cy0);
    n = ny + 1;
    if (cyN < 0)
      for (size_t i = 1; i <= nx; ++i, n += ny + 1) add_side(n, n + ny + 1, -cyN);
    NumberEquations();
  }

  // Box [xmin,xmax]x[ymin,ymax]x[zmin,zmax], nx x ny x nz cells, TETRAHEDRON.
  Mesh(real_t xmin, real_t xmax, real_t ymin, real_t ymax, real_t zmin, real_t zmax, size_t nx,
       size_t ny, size_t nz, int cx0, int cxN, int cy0, int cyN, int cz0, int czN,
       int opt = TETRAHEDRON, size_t nb_dof = 1)
      : _dim(3), _nb_dof(1), _nen(4) {
    if (opt != TETRAHEDRON || nb_dof != 1)
      throw OFELIException("Mesh::Mesh: boxes are split into tetrahedra with one DOF per node");
    const size_t nn = (nx + 1) * (ny + 1) * (nz + 1), ne = 6 * nx * ny * nz;
    _coord.resize(3 * nn);
    _code.resize(nn);
    _elem.resize(4 * ne);
    hip_check(ofh_mesh3d_kuhn(xmin, xmax, ymin, ymax, zmin, zmax, (int)nx, (int)ny, (int)nz, cx0,
                              cxN, cy0, cyN, cz0, czN, _coord.data(), _code.data(), _elem.data()),
              "Mesh::Mesh");
    NumberEquations();
  }

  size_t getDim() const { return _dim; }
  size_t getNbNodes() const { return _coord.size() / (_dim ? _dim : 1); }
  size_t getNbElements() const { return _nen ? _elem.size() / _nen : 0; }
  size_t getNbDOF() const { return getNbNodes() * _nb_dof; }
  size_t getNbEq() const { return _nb_eq; }
  size_t getNbSides() const { return _side_code.size(); }
  int nbDOFPerNode() const { return _nb_dof; }
  int nbElementNodes() const { return _nen; }

  // Mesh::removeImposedDOF (Mesh.h:420): imposed DOFs get label 0
  void removeImposedDOF() {
    _no_imposed_dof = true;
    NumberEquations();
  }

  // Mesh::NumberEquations (Mesh.cpp:1483-1527)
  size_t NumberEquations() {
    _dof.assign(_code.size(), 0);
    if (_no_imposed_dof) {
      int neq = 0;
      hip_check(ofh_number_equations((int)getNbNodes(), _nb_dof, _code.data(), _dof.data(), &neq),
                "Mesh::NumberEquations");
      _nb_eq = neq;
    } else {
      for (size_t t = 0; t < _dof.size(); ++t) _dof[t] = (int32_t)(t + 1);
      _nb_eq = _dof.size();
    }
    return _nb_eq;
  }

  // node n (1-based), DOF k (1-based)
  int getCode(size_t n, size_t k = 1) const { return _code[(n - 1) * _nb_dof + (k - 1)]; }
  void setCode(size_t n, size_t k, int c) {
    _code[(n - 1) * _nb_dof + (k - 1)] = c;
    NumberEquations();
  }
  size_t getDOF(size_t n, size_t k = 1) const { return _dof[(n - 1) * _nb_dof + (k - 1)]; }
  real_t getCoord(size_t n, size_t c) const { return _coord[(n - 1) * _dim + c]; }

  const std::vector<real_t>& coords() const { return _coord; }
  const std::vector<int32_t>& elements() const { return _elem; }
  const std::vector<int32_t>& codes() const { return _code; }
  std::vector<int32_t>& codes() { return _code; }
  const std::vector<int32_t>& dofs() const { return _dof; }
  const std::vector<int32_t>& sides() const { return _sides; }
  const std::vector<int>& sideCodes() const { return _side_code; }
  bool imposedDOFRemoved() const { return _no_imposed_dof; }
  // changes whenever codes change, so cached device plans can be rebuilt
  uint64_t revision() const { return _rev; }
  void touch() { ++_rev; }

 private:
  int _dim = 0, _nb_dof = 1, _nen = 0, _nx = 0, _ny = 0;
  size_t _nb_eq = 0;
  bool _no_imposed_dof = false;
  uint64_t _rev = 0;
  std::vector<real_t> _coord;
  std::vector<int32_t> _code, _elem, _dof, _sides;
  std::vector<int> _side_code;
};

}  // namespace OFELI
