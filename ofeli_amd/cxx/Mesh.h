// Mesh.h -- the mesh data the solve path needs (counterpart: include/mesh/Mesh.h).
//
// Structured constructors with the reference's node order, coordinates and
// boundary codes (Mesh.cpp:547-676 for rectangles; Mesh.cpp:678-726 node
// layout for boxes, with a positively oriented Kuhn tetrahedral split), DOF
// numbering with imposed-DOF elimination (Mesh.cpp:1483-1527,
// removeImposedDOF Mesh.h:420), the boundary sides of rectangles, and the
// OFELI XML mesh files (.m): Mesh(file) / get(file) and put(file) / save(file).
#pragma once

#include <cstdint>
#include <string>
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

  // Mesh(file) and Mesh::get(file) for the OFELI XML format (Mesh.cpp:88-100,
  // 2639-2670): coordinates rounded to single precision as the reference
  // reads them; equations numbered as after Mesh::get.
  explicit Mesh(const std::string& file) { get(file); }
  void get(const std::string& file) {
    ofh_mesh h = nullptr;
    hip_check(ofh_mesh_read(file.c_str(), &h), "Mesh::get");
    int dim = 0, nb_dof = 0, nn = 0, ne = 0, es = 0, nen = 0, ns = 0, ss = 0, nsn = 0, nm = 0;
    ofh_mesh_info(h, &dim, &nb_dof, &nn, &ne, &es, &nen, &ns, &ss, &nsn, &nm);
    _dim = dim;
    _nb_dof = nb_dof;
    _nen = nen;
    _el_shape = es;
    _sd_shape = ss;
    _coord.assign((size_t)nn * dim, 0.0);
    _code.assign((size_t)nn * nb_dof, 0);
    _elem.assign((size_t)ne * nen, 0);
    _elem_code.assign((size_t)ne, 0);
    _sides.assign((size_t)ns * nsn, 0);
    _side_codes.assign((size_t)ns * nb_dof, 0);
    const int rc = ofh_mesh_get(h, _coord.data(), (int64_t)_coord.size(), _code.data(),
                                (int64_t)_code.size(), _elem.data(), (int64_t)_elem.size(),
                                _elem_code.data(), (int64_t)_elem_code.size(), _sides.data(),
                                (int64_t)_sides.size(), _side_codes.data(),
                                (int64_t)_side_codes.size());
    if (rc != OFH_OK) ofh_mesh_destroy(h);
    hip_check(rc, "Mesh::get");
    _materials.clear();
    for (int i = 0; i < nm; ++i) {
      int c = 0;
      char name[256];
      ofh_mesh_material(h, i, &c, name, (int)sizeof(name));
      _materials.emplace_back(c, name);
    }
    ofh_mesh_destroy(h);
    _side_code.resize(ns);
    for (int s = 0; s < ns; ++s) _side_code[s] = _side_codes[(size_t)s * nb_dof];
    _no_imposed_dof = false;
    NumberEquations();
    touch();
  }
  // Mesh::put (Mesh.cpp:2734-2800) and Mesh::save(file) for ".m" names
  void put(const std::string& file) const {
    std::vector<int32_t> ec = _elem_code, sc = _side_codes;
    if (ec.size() != getNbElements()) ec.assign(getNbElements(), 1);
    if (sc.size() != getNbSides() * _nb_dof) {
      sc.assign(getNbSides() * _nb_dof, 0);
      for (size_t s = 0; s < getNbSides(); ++s) sc[s * _nb_dof] = _side_code[s];
    }
    std::vector<int32_t> mc;
    std::vector<const char*> mn;
    for (auto& m : _materials) {
      mc.push_back(m.first);
      mn.push_back(m.second.c_str());
    }
    const int nsn = getNbSides() ? (int)(_sides.size() / getNbSides()) : 0;
    hip_check(ofh_mesh_write(file.c_str(), _dim, _nb_dof, (int)getNbNodes(), _coord.data(),
                             _code.data(), (int)getNbElements(), elementShape(), _nen,
                             _elem.data(), ec.data(), (int)getNbSides(), sideShape(), nsn,
                             _sides.data(), sc.data(), (int)mc.size(), mc.data(), mn.data()),
              "Mesh::put");
  }
  void save(const std::string& file) const {
    if (file.size() < 2 || file.compare(file.size() - 2, 2, ".m") != 0)
      throw OFELIException("Mesh::save: only the OFELI format (.m) is provided");
    put(file);
  }
  // ElementShape values (OFELI_Config.h:272-283)
  int elementShape() const {
    if (_el_shape) return _el_shape;
    if (_nen == 2) return 2;
    if (_dim == 2) return _nen == 4 ? 4 : 3;
    return 5;
  }
  int sideShape() const { return _sd_shape ? _sd_shape : (_dim == 3 ? 3 : 2); }
  int getElementCode(size_t e) const {
    return e - 1 < _elem_code.size() ? _elem_code[e - 1] : 1;
  }
  const std::vector<std::pair<int, std::string>>& materials() const { return _materials; }

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
  int _el_shape = 0, _sd_shape = 0;  // from a mesh file (0: by dimension)
  std::vector<int32_t> _elem_code, _side_codes;
  std::vector<std::pair<int, std::string>> _materials;
  size_t _nb_eq = 0;
  bool _no_imposed_dof = false;
  uint64_t _rev = 0;
  std::vector<real_t> _coord;
  std::vector<int32_t> _code, _elem, _dof, _sides;
  std::vector<int> _side_code;
};

}  // namespace OFELI
