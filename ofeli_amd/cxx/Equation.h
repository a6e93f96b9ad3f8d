// Equation.h -- the equation classes of the path (counterparts:
// include/equations/Equa.h, Equation_impl.h, laplace/Laplace2DT3.h,
// laplace/Laplace3DT4.h, therm/DC2DT3.h, therm/DC3DT4.h, solid/Elas2DQ4.h).
//
// run() follows Equa::run (Equa.cpp:381-421): build the system (element
// loop on the GPU, ofh_asm_build), solve it with the selected iterative
// method on the GPU (zero initial guess, as _uu.setSize clears it), and
// expand the solution with the Dirichlet values (insertBC).  Per-element
// coefficients (conductivity, Young's modulus) are a documented extension:
// the reference only offers expression strings or material codes.
#pragma once

#include <memory>

#include "Mesh.h"
#include "Solvers.h"
#include "SpMatrix.h"
#include "Vect.h"

namespace OFELI {

enum class EType { SOLUTION = 11, INITIAL = 11, BOUNDARY_CONDITION = 12, DIRICHLET = 12,
                   BODY_FORCE = 13, SOURCE = 13, BOUNDARY_FORCE = 14, FLUX = 14,
                   TRACTION = 15, NEUMANN = 15 };

class Equa {
 public:
  Equa(Mesh& ms, Vect<real_t>& u, int kind, Iteration s, Preconditioner p)
      : _mesh(&ms), _u(&u), _kind(kind) {
    ms.removeImposedDOF();  // Equa::setMesh (Equa.cpp:334-348)
    _ls.setSolver(s, p);
  }
  virtual ~Equa() = default;

  // Equa::setSolver (Equa.cpp:467-485; default preconditioner Equa.h:318-319)
  void setSolver(Iteration s, Preconditioner p = IDENT_PREC) {
    if (s == DIRECT_SOLVER)
      throw OFELIException("In Equa::setSolver: the GPU path solves iteratively");
    _ls.setSolver(s, p);
  }
  LinearSolver& getLinearSolver() { return _ls; }
  void setTolerance(real_t t) { _toler_equa = t; }  // as Equa.cpp:723-726 (not forwarded)

  void setInput(EType opt, Vect<real_t>& v) {
    switch (opt) {
      case EType::SOLUTION: _u = &v; break;
      case EType::BOUNDARY_CONDITION: _bc = &v; break;
      case EType::SOURCE: _bf = &v; break;
      case EType::FLUX:
      case EType::TRACTION: _sf = &v; break;
    }
  }
  void setBoundaryCondition(Vect<real_t>& u) { _bc = &u; }
  void setDirichlet(Vect<real_t>& u) { _bc = &u; }
  void setSolution(Vect<real_t>& u) { _u = &u; }
  void setSource(Vect<real_t>& f) { _bf = &f; }
  void setBodyForce(Vect<real_t>& f) { _bf = &f; }
  void setTraction(Vect<real_t>& f) { _sf = &f; }
  // per-element coefficient: k (diffusion kinds) or E (Elas2DQ4)
  void setElementCoefficient(const Vect<real_t>& c) { _coef = &c; _coef_dirty = true; }

  // Equa::run (Equa.cpp:381-421), stationary
  int run() { return solve_step(false, nullptr); }

  SpMatrix<real_t>& getMatrix() { ensure_matrix(); return *_A; }
  const Vect<real_t>& getRHS() const { return _b; }
  const ofh_solve_info& getInfo() const { return _ls.getInfo(); }

 protected:
  virtual void params(std::vector<double>& prm) const { prm.clear(); }
  virtual const double* previous_solution() { return nullptr; }

  void ensure_matrix() {
    if (!_A) _A = std::make_unique<SpMatrix<real_t>>(*_mesh);
  }

  // build + solve + insertBC; warm: start from the current solution
  int solve_step(bool warm, const Vect<real_t>* guess) {
    ensure_matrix();
    const size_t neq = _mesh->getNbEq(), ndof = _mesh->getNbDOF();
    ofh_asm P = _A->plan();
    if (_coef && _coef_dirty) {
      _dcoef.upload(_coef->data(), _coef->size());
      _coef_dirty = false;
    }
    if (_bf) _df.upload(_bf->data(), _bf->size());
    if (_bc) _dbc.upload(_bc->data(), _bc->size());
    _db.resize(neq);
    std::vector<double> prm;
    params(prm);
    const double* up = previous_solution();
    const bool matrix = !_matrix_built || !_matrix_reusable;
    hip_check(ofh_asm_build(P, _kind, _coef ? _dcoef.get() : nullptr, _bf ? _df.get() : nullptr,
                            _bc ? _dbc.get() : nullptr, up, prm.empty() ? nullptr : prm.data(),
                            (int)prm.size(), matrix ? 1 : 0, _db.get()),
              "Equa::build");
    if (matrix) _A->deviceWritten();
    _matrix_built = true;
    if (_sf && !_mesh->sides().empty()) {
      _dsf.upload(_sf->data(), _sf->size());
      hip_check(ofh_asm_traction(P, (int)_mesh->getNbSides(), _mesh->sides().data(), _dsf.get(),
                                 _db.get()),
                "Equa::build");
    }
    // initial guess: zero (Equa::run) or the free DOFs of the current solution
    _dx.resize(neq);
    std::vector<double> x0(neq, 0.0);
    const Vect<real_t>* g = guess ? guess : (warm ? _u : nullptr);
    if (g && g->size() == ndof) {
      const auto& dof = _mesh->dofs();
      for (size_t t = 0; t < ndof; ++t)
        if (dof[t]) x0[dof[t] - 1] = (*g)[t];
    }
    _dx.upload(x0.data(), neq);
    ofh_solve_info info{};
    hip_check(ofh_linear_solve(_A->device(), _ls.getSolver(), _ls.getPreconditioner(), _db.get(),
                               _dx.get(), max_it(), toler(), &info),
              "Equa::solveLinearSystem");
    _last = info;
    // _u->insertBC(mesh, _uu, bc)
    _du.resize(ndof);
    hip_check(ofh_asm_insert_bc(P, _dx.get(), _bc ? _dbc.get() : nullptr, _du.get()),
              "Vect::insertBC");
    _u->setSize(ndof);
    _du.download(_u->data(), ndof);
    _b.setSize(neq);
    _db.download(_b.data(), neq);
    return info.nb_it < 0 ? -info.nb_it : 0;
  }

  int max_it() const { return _ls_max_it; }
  real_t toler() const { return _ls_toler; }

 public:
  // LinearSolver settings used by run() (getLinearSolver().setMaxIter/
  // setTolerance in the reference); mirrored here for the device call.
  void setMaxIter(int m) { _ls_max_it = m; _ls.setMaxIter(m); }
  void setLinearTolerance(real_t t) { _ls_toler = t; _ls.setTolerance(t); }
  const ofh_solve_info& lastSolve() const { return _last; }

 protected:
  Mesh* _mesh;
  Vect<real_t>* _u;
  Vect<real_t>* _bc = nullptr;
  Vect<real_t>* _bf = nullptr;
  Vect<real_t>* _sf = nullptr;
  const Vect<real_t>* _coef = nullptr;
  bool _coef_dirty = false;
  int _kind;
  LinearSolver _ls;
  int _ls_max_it = 1000;
  real_t _ls_toler = 1.4901161193847656e-08;
  real_t _toler_equa = 1.e-8;
  std::unique_ptr<SpMatrix<real_t>> _A;
  Vect<real_t> _b;
  DeviceBuffer _dcoef, _df, _dbc, _db, _dx, _du, _dsf;
  bool _matrix_built = false, _matrix_reusable = false;
  ofh_solve_info _last{};
};

// Laplace2DT3 (Laplace2DT3.cpp:64-143): -Delta u = f with P1 triangles
class Laplace2DT3 : public Equa {
 public:
  // the reference's default is DIRECT_SOLVER (Laplace2DT3.cpp:69); this path
  // solves iteratively, with the preconditioner of the other classes
  Laplace2DT3(Mesh& ms, Vect<real_t>& u) : Equa(ms, u, OFH_LAPLACE2DT3, CG_SOLVER, DILU_PREC) {}
};

// Laplace3DT4 (Laplace3DT4.cpp:62-129): -Delta u = f with P1 tetrahedra;
// with setElementCoefficient(k) it assembles DC3DT4's Diffusion(k_e).
class Laplace3DT4 : public Equa {
 public:
  Laplace3DT4(Mesh& ms, Vect<real_t>& u) : Equa(ms, u, OFH_LAPLACE3DT4, CG_SOLVER, DILU_PREC) {}  // Laplace3DT4.cpp:70
};

// DC2DT3 (DC2DT3.cpp:64-253): steady diffusion with per-element k, or the
// backward-Euler transient of Equa_Therm::build (Equa_Therm.h:162-218).
class DC2DT3 : public Equa {
 public:
  DC2DT3(Mesh& ms, Vect<real_t>& u) : Equa(ms, u, OFH_DC2DT3, CG_SOLVER, DILU_PREC) {}  // DC2DT3.cpp:72
  void setTransient(TimeScheme s) {
    if (s != BACKWARD_EULER && s != NONE)
      throw OFELIException("DC2DT3::setTransient: the GPU path provides BACKWARD_EULER");
    _kind = s == BACKWARD_EULER ? OFH_DC2DT3_BE : OFH_DC2DT3;
    _matrix_built = false;
    _matrix_reusable = _kind == OFH_DC2DT3_BE;  // dt, k, rho, cp constant: A is step-invariant
  }
  void setTimeStep(real_t dt) { _dt = dt; _matrix_built = false; }
  void set_rho(real_t r) { _rho = r; _matrix_built = false; }
  void set_Cp(real_t c) { _cp = c; _matrix_built = false; }
  // one backward-Euler step from the current solution u^n (Equa::run with
  // TRANSIENT_ONE_STEP, Equa.cpp:412-419); warm = true starts CG from u^n
  // as the TimeStepping driver does (TimeStepping.cpp:207, 452)
  int runOneTimeStep(bool warm = false, const Vect<real_t>* guess = nullptr) {
    if (_kind != OFH_DC2DT3_BE) throw OFELIException("DC2DT3::runOneTimeStep: call setTransient first");
    _un.upload(_u->data(), _u->size());
    return solve_step(warm, guess);
  }

 protected:
  void params(std::vector<double>& prm) const override {
    prm.clear();
    if (_kind == OFH_DC2DT3_BE) prm = {_dt, _rho, _cp};
  }
  const double* previous_solution() override {
    return _kind == OFH_DC2DT3_BE ? _un.get() : nullptr;
  }

 private:
  real_t _dt = 0.1, _rho = 1.0, _cp = 1.0;  // material/Generic.md: rho = cp = 1
  DeviceBuffer _un;
};

// Elas2DQ4 (Elas2DQ4.cpp:54-298): plane elasticity with Q4 elements.
// Reference parity (default): Elas2DQ4::set always ends with PlaneStrain()
// and eMat = 2*Deviator + Dilatation.  setReferenceParity(false) makes
// PlaneStress()/PlaneStrain() take effect with eMat = Deviator + Dilatation.
class Elas2DQ4 : public Equa {
 public:
  Elas2DQ4(Mesh& ms, Vect<real_t>& u) : Equa(ms, u, OFH_ELAS2DQ4, CG_SOLVER, DILU_PREC) {  // Elas2DQ4.cpp:61
    setYoung(1000.0);  // material/Generic.md: E = 1000, nu = 0.2
  }
  void PlaneStress() { _stress = true; }
  void PlaneStrain() { _stress = false; }
  void setReferenceParity(bool on) { _parity = on; _matrix_built = false; }
  void setPoisson(real_t nu) { _nu = nu; }
  // constant Young's modulus (the reference reads it from the material)
  void setYoung(real_t E) {
    _E.setSize(_mesh->getNbElements());
    _E = E;
    setElementCoefficient(_E);
  }

 protected:
  void params(std::vector<double>& prm) const override {
    prm = {_nu, _parity ? 0.0 : (_stress ? 1.0 : 2.0)};
  }

 private:
  real_t _nu = 0.2;  // material/Generic.md
  bool _stress = false, _parity = true;
  Vect<real_t> _E;
};

}  // namespace OFELI
