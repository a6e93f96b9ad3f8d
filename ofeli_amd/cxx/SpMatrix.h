// SpMatrix.h -- SpMatrix<T_> of the drop-in layer (counterpart:
// include/linear_algebra/SpMatrix.h:78-611, SpMatrix_impl.h).
//
// CSR storage as the reference keeps it on the host (row_ptr 0-based,
// col_ind 1-based, sorted per row) plus a device copy (ofh_csr).  Values
// live where they were last written: host writes through set/add/
// operator()/get() invalidate the device copy; GPU assembly invalidates the
// host copy, which is downloaded on the next host read.  Every product and
// solve runs on the GPU.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

#include "Mesh.h"
#include "OFELI_Config.h"
#include "Vect.h"

namespace OFELI {

class LinearSolver;

template <class T_>
class SpMatrix {
  static_assert(sizeof(T_) == sizeof(double), "the GPU path stores FP64 values");

 public:
  SpMatrix() = default;

  // SpMatrix(mesh, dof=0, is_diagonal=0) (SpMatrix_impl.h:108-119): the
  // NodeGraph pattern on the GPU; dof != 0 selects NodeGraphScal's pattern
  // as setMesh does
  explicit SpMatrix(Mesh& mesh, size_t dof = 0, int is_diagonal = 0) {
    _diag_only = is_diagonal != 0;
    setMesh(mesh, dof);
  }

  // SpMatrix(row_ptr, col_ind) / (row_ptr, col_ind, a): col_ind 1-based
  SpMatrix(const std::vector<size_t>& row_ptr, const std::vector<size_t>& col_ind)
      : _row_ptr(row_ptr), _col_ind(col_ind), _a(col_ind.size(), T_(0)) {
    _nrows = row_ptr.size() - 1;
    _where = HOST;
  }
  SpMatrix(const std::vector<size_t>& row_ptr, const std::vector<size_t>& col_ind,
           const std::vector<T_>& a)
      : _row_ptr(row_ptr), _col_ind(col_ind), _a(a) {
    _nrows = row_ptr.size() - 1;
    _where = HOST;
  }

  SpMatrix(const SpMatrix& m)
      : _row_ptr(m._row_ptr), _col_ind(m._col_ind), _nrows(m._nrows), _diag_only(m._diag_only),
        _solver(m._solver), _prec(m._prec), _max_it(m._max_it), _toler(m._toler) {
    m.host_read();
    _a = m._a;
    _where = HOST;
  }
  SpMatrix& operator=(const SpMatrix&) = delete;

  ~SpMatrix() { release(); }

  // SpMatrix::setMesh (SpMatrix_impl.h:391-413).  dof == 0: NodeGraph over
  // the mesh's equations (imposed DOFs eliminated), nb_eq rows.  dof != 0:
  // NodeGraphScal (GraphOfMatrix.cpp:532-546), one row per node numbered as
  // the node (Matrix_impl.h:148-150), every element node pair coupled, no
  // elimination: built as NodeGraph of one DOF per node with no imposed DOF.
  void setMesh(Mesh& mesh, size_t dof = 0) {
    release();
    std::vector<int32_t> rp, col;
    const int dim = (int)mesh.getDim();
    ofh_asm plan = nullptr;
    const std::vector<int32_t> free_nodes(dof ? mesh.getNbNodes() : 0, 0);
    hip_check(ofh_asm_create(hip_context(), dim, (int)mesh.getNbNodes(), mesh.coords().data(),
                             (int)mesh.getNbElements(), mesh.nbElementNodes(),
                             mesh.elements().data(), dof ? 1 : mesh.nbDOFPerNode(),
                             dof ? free_nodes.data() : mesh.codes().data(), &plan),
              "SpMatrix::setMesh");
    _plan = plan;
    hip_check(ofh_asm_matrix(plan, &_csr), "SpMatrix::setMesh");
    int n = 0;
    int64_t nnz = 0;
    hip_check(ofh_csr_info(_csr, &n, &nnz, nullptr), "SpMatrix::setMesh");
    rp.resize(n + 1);
    col.resize(nnz);
    hip_check(ofh_csr_get_pattern(_csr, rp.data(), col.data()), "SpMatrix::setMesh");
    _nrows = n;
    _row_ptr.assign(rp.begin(), rp.end());
    _col_ind.resize(nnz);
    for (int64_t k = 0; k < nnz; ++k) _col_ind[k] = (size_t)col[k] + 1;
    _a.assign(nnz, T_(0));
    _where = BOTH;
    _pattern_on_device = true;
  }

  size_t getNbRows() const { return _nrows; }
  size_t getNbColumns() const { return _nrows; }
  size_t size() const { return _nrows; }
  size_t getLength() const { return _col_ind.size(); }
  size_t getColInd(size_t i) const { return _col_ind[i - 1]; }
  size_t getRowPtr(size_t i) const { return _row_ptr[i - 1]; }
  bool isDiagonal() const { return _diag_only; }

  // entry access, 1-based (SpMatrix_impl.h:668-685, 798-804)
  T_ operator()(size_t i, size_t j) const {
    const int k = col_index(i, j);
    if (k < 0) return T_(0);
    host_read();
    return _a[_row_ptr[i - 1] + k];
  }
  T_ operator()(size_t i) const {
    host_read();
    return _a[i - 1];
  }
  void set(size_t i, size_t j, const T_& v) {
    const int k = col_index(i, j);
    if (k < 0)
      throw OFELIException("In SpMatrix::set(i,j,x): Index pair (" + std::to_string(i) + "," +
                           std::to_string(j) + ") not compatible with sparse storage.");
    host_write();
    _a[_row_ptr[i - 1] + k] = v;
  }
  void add(size_t i, size_t j, const T_& v) {
    const int k = col_index(i, j);
    if (k < 0)
      throw OFELIException("In SpMatrix::add(i,j,x): Index pair (" + std::to_string(i) + "," +
                           std::to_string(j) + ") not compatible with sparse storage.");
    host_write();
    _a[_row_ptr[i - 1] + k] += v;
  }
  void clear() {
    host_write();
    std::fill(_a.begin(), _a.end(), T_(0));
  }
  void operator=(const T_& x) {
    host_write();
    std::fill(_a.begin(), _a.end(), x);
  }
  // mutable host pointer (SpMatrix_impl.h:980-984): later writes through it
  // are re-uploaded before every GPU use
  T_* get() {
    host_write();
    _exposed = true;
    return _a.data();
  }

  // y = A x (Mult, SpMatrix_impl.h:728-734) and y += A x (MultAdd, :737-745)
  void Mult(const Vect<T_>& x, Vect<T_>& y) const {
    y.setSize(_nrows);
    gpu_spmv(x, y, false, 1.0);
  }
  void MultAdd(const Vect<T_>& x, Vect<T_>& y) const { gpu_spmv(x, y, true, 1.0); }
  void MultAdd(T_ a, const Vect<T_>& x, Vect<T_>& y) const { gpu_spmv(x, y, true, a); }
  Vect<T_> operator*(const Vect<T_>& x) const {
    Vect<T_> y(_nrows);
    Mult(x, y);
    return y;
  }
  // SpMatrix::TMult (SpMatrix_impl.h:761-767): y = A^T x.  y is overwritten;
  // the reference adds into y and indexes the previous entry's column.
  void TMult(const Vect<T_>& x, Vect<T_>& y) const {
    y.setSize(_nrows);
    ofh_csr A = device();
    DeviceBuffer dx, dy;
    dx.upload(x.data(), _nrows);
    dy.resize(_nrows);
    hip_check(ofh_spmv_t_f64(A, dx.get(), dy.get()), "SpMatrix::TMult");
    dy.download(y.data(), _nrows);
  }

  // SpMatrix::setSolver / solve (SpMatrix_impl.h:956-978); unlike the
  // reference, the max_it and toler given here are honoured.
  void setSolver(Iteration solver = CG_SOLVER, Preconditioner prec = DIAG_PREC,
                 int max_it = 1000, real_t toler = 1.e-8) {
    _solver = solver;
    _prec = prec;
    _max_it = max_it;
    _toler = toler;
  }
  int solve(const Vect<T_>& b, Vect<T_>& x, bool fact = false);
  int solve(Vect<T_>& b, bool fact = false) {
    Vect<T_> x(b.size());
    int ret = solve(b, x, fact);
    b = x;
    return ret;
  }

  // SpMatrix::SSORSolve (SpMatrix_impl.h:913-942)
  void SSORSolve(const Vect<T_>& b, Vect<T_>& x) const {
    DeviceBuffer db, dx;
    db.upload(b.data(), _nrows);
    dx.resize(_nrows);
    hip_check(ofh_ssor_solve(device(), db.get(), dx.get()), "SpMatrix::SSORSolve(b,x)");
    x.setSize(_nrows);
    dx.download(x.data(), _nrows);
  }

  // ---- device side (used by the solver functions and the equations) ----
  ofh_csr device() const {
    if (!_csr) {
      std::vector<int32_t> rp(_row_ptr.begin(), _row_ptr.end()), col(_col_ind.size());
      for (size_t k = 0; k < col.size(); ++k) col[k] = (int32_t)_col_ind[k] - 1;
      ofh_csr c = nullptr;
      hip_check(ofh_csr_create(hip_context(), (int)_nrows, (int64_t)col.size(), rp.data(),
                               col.data(), _a.data(), &c),
                "SpMatrix");
      _csr = c;
      _owns_csr = true;
      _where = BOTH;
    }
    if (_where == HOST || _exposed) {
      hip_check(ofh_csr_set_values(_csr, _a.data()), "SpMatrix");
      _where = BOTH;
    }
    return _csr;
  }
  ofh_asm plan() const { return _plan; }
  // the values were produced on the device (GPU assembly)
  void deviceWritten() const {
    _where = DEVICE;
    _exposed = false;
  }

 private:
  enum Where { HOST, DEVICE, BOTH };

  int col_index(size_t i, size_t j) const {  // SpMatrix_impl.h:1008-1015
    if (i < 1 || i > _nrows) return -1;
    const size_t a = _row_ptr[i - 1], e = _row_ptr[i];
    auto it = std::lower_bound(_col_ind.begin() + a, _col_ind.begin() + e, j);
    if (it == _col_ind.begin() + e || *it != j) return -1;
    return (int)(it - (_col_ind.begin() + a));
  }
  void host_read() const {
    if (_where == DEVICE) {
      hip_check(ofh_csr_get_values(_csr, const_cast<T_*>(_a.data())), "SpMatrix");
      _where = BOTH;
    }
  }
  void host_write() {
    host_read();
    _where = HOST;
  }
  void gpu_spmv(const Vect<T_>& x, Vect<T_>& y, bool add, double a) const {
    ofh_csr A = device();
    DeviceBuffer dx, dy;
    dx.upload(x.data(), _nrows);
    if (add) dy.upload(y.data(), _nrows);
    else dy.resize(_nrows);
    if (add) hip_check(ofh_spmv_add_f64(A, a, dx.get(), dy.get()), "SpMatrix::MultAdd");
    else hip_check(ofh_spmv_f64(A, dx.get(), dy.get()), "SpMatrix::Mult");
    dy.download(y.data(), _nrows);
  }
  void release() {
    if (_plan) ofh_asm_destroy(_plan);  // owns its CSR
    else if (_csr && _owns_csr) ofh_csr_destroy(_csr);
    _plan = nullptr;
    _csr = nullptr;
    _owns_csr = false;
  }

  std::vector<size_t> _row_ptr, _col_ind;
  mutable std::vector<T_> _a;
  size_t _nrows = 0;
  bool _diag_only = false;
  mutable Where _where = HOST;
  mutable bool _exposed = false;
  mutable ofh_csr _csr = nullptr;
  mutable bool _owns_csr = false;
  bool _pattern_on_device = false;
  ofh_asm _plan = nullptr;
  Iteration _solver = CG_SOLVER;
  Preconditioner _prec = DIAG_PREC;
  int _max_it = 1000;
  real_t _toler = 1.e-8;
};

}  // namespace OFELI
