// Vect.h -- Vect<T_> of the drop-in layer (counterpart: include/linear_algebra/Vect.h).
//
// A std::vector<T_> with OFELI's 1-based operator(), 0-based operator[], the
// dot product operator, getNorm2, arithmetic and insertBC.  Host storage is
// authoritative: solver calls upload their vector operands and download the
// results, so code written against OFELI's Vect keeps working unchanged.
#pragma once

#include <cmath>
#include <vector>

#include "Mesh.h"
#include "OFELI_Config.h"

namespace OFELI {

template <class T_>
class Vect : public std::vector<T_> {
  using base = std::vector<T_>;

 public:
  Vect() = default;
  explicit Vect(size_t n) : base(n, T_(0)) {}
  Vect(size_t nx, size_t ny) : base(nx * ny, T_(0)), _nx(nx), _ny(ny) {}
  // Vect(mesh): one entry per DOF of the mesh (Vect_impl.h:88-100)
  explicit Vect(const Mesh& m) : base(m.getNbDOF(), T_(0)) {}

  // Vect::setSize clears (Vect_impl.h:539-548)
  void setSize(size_t n) {
    base::assign(n, T_(0));
  }
  void clear() { std::fill(base::begin(), base::end(), T_(0)); }
  size_t size() const { return base::size(); }

  T_& operator()(size_t i) { return (*this)[i - 1]; }
  T_ operator()(size_t i) const { return (*this)[i - 1]; }
  T_& operator()(size_t i, size_t j) { return (*this)[(i - 1) * (_ny ? _ny : 1) + j - 1]; }
  T_ operator()(size_t i, size_t j) const { return (*this)[(i - 1) * (_ny ? _ny : 1) + j - 1]; }

  Vect& operator=(const T_& a) {
    std::fill(base::begin(), base::end(), a);
    return *this;
  }
  // dot product, Vect_impl.h:2218-2225
  T_ operator,(const Vect& v) const {
    T_ p = 0;
    for (size_t i = 0; i < size(); ++i) p += (*this)[i] * v[i];
    return p;
  }
  // Vect_impl.h:660-668
  real_t getNorm2() const {
    real_t s = 0;
    for (size_t i = 0; i < size(); ++i) s += (*this)[i] * (*this)[i];
    return std::sqrt(s);
  }
  Vect& operator+=(const Vect& v) {
    for (size_t i = 0; i < size(); ++i) (*this)[i] += v[i];
    return *this;
  }
  Vect& operator-=(const Vect& v) {
    for (size_t i = 0; i < size(); ++i) (*this)[i] -= v[i];
    return *this;
  }
  Vect& operator*=(const T_& a) {
    for (auto& x : *this) x *= a;
    return *this;
  }
  friend Vect operator+(const Vect& a, const Vect& b) {
    Vect c(a);
    return c += b;
  }
  friend Vect operator-(const Vect& a, const Vect& b) {
    Vect c(a);
    return c -= b;
  }
  friend Vect operator*(const T_& s, const Vect& a) {
    Vect c(a);
    return c *= s;
  }

  // Vect::insertBC(mesh, v, bc) (Vect_impl.h:1097-1116): expand the free-DOF
  // vector v to all DOFs, imposed ones taking bc.
  void insertBC(const Mesh& m, const Vect& v, const Vect& bc) {
    base::assign(m.getNbDOF(), T_(0));
    size_t j = 0;
    const auto& code = m.codes();
    for (size_t i = 0; i < code.size(); ++i) (*this)[i] = code[i] == 0 ? v[j++] : bc[i];
  }
  void insertBC(const Mesh& m, const Vect& v) {
    base::assign(m.getNbDOF(), T_(0));
    size_t j = 0;
    const auto& code = m.codes();
    for (size_t i = 0; i < code.size(); ++i) (*this)[i] = code[i] == 0 ? v[j++] : T_(0);
  }

 private:
  size_t _nx = 0, _ny = 0;
};

}  // namespace OFELI
