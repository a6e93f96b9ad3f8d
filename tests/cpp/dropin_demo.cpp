// dropin_demo.cpp -- OFELI-style programs written against the drop-in layer.
//
// Each case is the kind of code an OFELI user writes (laplace_demo1-like
// harmonic Dirichlet problem, a LinearSolver on a hand-built SpMatrix, a
// transient DC2DT3 loop, an Elas2DQ4 cantilever).  The program prints one
// JSON object per case; tests/test_gpu_dropin.py checks them against the
// oracle.  Build: see tests/cpp/build.sh.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "OFELI.h"

using namespace OFELI;

static void print_vec(const char* name, const Vect<real_t>& v) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? "," : "", v[i]);
  std::printf("]");
}

int main(int argc, char** argv) {
  std::printf("[\n");
  // 1. Laplace2DT3 with harmonic Dirichlet data (demos/laplace/laplace_demo1.cpp)
  {
    Mesh ms(0., 1., 0., 1., 16, 16, 1, 1, 1, 1, TRIANGLE, 1);
    Vect<real_t> u(ms), bc(ms), f(ms);
    for (size_t n = 1; n <= ms.getNbNodes(); ++n) {
      real_t x = ms.getCoord(n, 0), y = ms.getCoord(n, 1);
      bc(n) = std::exp(x) * std::sin(y) + std::exp(y) * std::cos(x);
    }
    Laplace2DT3 eq(ms, u);
    eq.setInput(EType::BOUNDARY_CONDITION, bc);
    eq.setSolver(CG_SOLVER, DIAG_PREC);
    eq.setLinearTolerance(1e-12);
    eq.setMaxIter(2000);
    int ret = eq.run();
    real_t err = 0;
    for (size_t n = 1; n <= ms.getNbNodes(); ++n) err = std::max(err, std::fabs(u(n) - bc(n)));
    std::printf("{\"case\": \"laplace2dt3\", \"ret\": %d, \"it\": %d, \"maxerr\": %.6e, ", ret,
                eq.lastSolve().nb_it, err);
    print_vec("u", u);
    std::printf("},\n");
  }
  // 2. LinearSolver on a SpMatrix built from (row_ptr, col_ind, a)
  {
    // 1D Laplacian, n = 50 (SpMatrix::Laplace1D pattern)
    const size_t n = 50;
    std::vector<size_t> rp(1, 0), ci;
    std::vector<real_t> a;
    for (size_t i = 1; i <= n; ++i) {
      if (i > 1) { ci.push_back(i - 1); a.push_back(-1.0); }
      ci.push_back(i); a.push_back(2.0);
      if (i < n) { ci.push_back(i + 1); a.push_back(-1.0); }
      rp.push_back(ci.size());
    }
    SpMatrix<real_t> A(rp, ci, a);
    Vect<real_t> b(n), x(n);
    for (size_t i = 1; i <= n; ++i) b(i) = 1.0 / n;
    LinearSolver ls(A, b, x);
    ls.setTolerance(1e-12);
    int r1 = ls.solve(CG_SOLVER, DIAG_PREC);
    int it1 = ls.getNbIter();
    Vect<real_t> y = A * x;
    real_t res = (y - b).getNorm2() / b.getNorm2();
    Vect<real_t> x2(n), x3(n);
    int it2 = BiCGStab(A, SSOR_PREC, b, x2, 1000, 1e-12);
    int it3 = GMRes(A, DIAG_PREC, b, x3, 25, 1000, 1e-12);
    std::printf("{\"case\": \"linearsolver\", \"ret\": %d, \"it_cg\": %d, \"res\": %.6e, "
                "\"it_bicgstab\": %d, \"it_gmres\": %d, \"diff23\": %.6e, ",
                r1, it1, res, it2, it3, (x2 - x3).getNorm2() / x3.getNorm2());
    print_vec("x", x);
    std::printf("},\n");
    // entry access, 1-based
    A.set(3, 3, 5.0);
    std::printf("{\"case\": \"access\", \"a33\": %g, \"a34\": %g, \"a310\": %g},\n", A(3, 3),
                A(3, 4), A(3, 10));
  }
  // 3. DC2DT3 backward Euler, 5 steps, warm-started as TimeStepping does
  {
    Mesh ms(0., 1., 0., 1., 12, 12, 1, 1, 1, 1, TRIANGLE, 1);
    Vect<real_t> u(ms), f(ms), k(ms.getNbElements());
    for (size_t n = 1; n <= ms.getNbNodes(); ++n) f(n) = 1.0 + ms.getCoord(n, 0);
    for (size_t e = 1; e <= ms.getNbElements(); ++e) k(e) = 1.0 + 0.5 * std::sin(0.1 * e);
    DC2DT3 eq(ms, u);
    eq.setSource(f);
    eq.setElementCoefficient(k);
    eq.setTransient(BACKWARD_EULER);
    eq.setTimeStep(0.01);
    eq.setLinearTolerance(1e-12);
    int its = 0;
    for (int s = 0; s < 5; ++s) {
      eq.runOneTimeStep(true);
      its += eq.lastSolve().iterations;
    }
    std::printf("{\"case\": \"dc2dt3_be\", \"iterations\": %d, ", its);
    print_vec("u", u);
    std::printf("},\n");
  }
  // 4. Elas2DQ4 cantilever, clamped left edge, traction on the right edge
  {
    Mesh ms(0., 4., 0., 1., 8, 2, 1, -1, 0, 0, QUADRILATERAL, 2);
    for (size_t n = 1; n <= ms.getNbNodes(); ++n) ms.setCode(n, 2, ms.getCode(n, 1));
    Vect<real_t> u(ms), tr(2 * ms.getNbSides()), E(ms.getNbElements());
    for (size_t s = 0; s < ms.getNbSides(); ++s) { tr[2 * s] = 0.0; tr[2 * s + 1] = -0.01; }
    for (size_t e = 1; e <= ms.getNbElements(); ++e) E(e) = 1.0 + 0.1 * (e % 3);
    Elas2DQ4 eq(ms, u);
    eq.setElementCoefficient(E);
    eq.setPoisson(0.3);
    eq.setTraction(tr);
    eq.setSolver(CG_SOLVER, SSOR_PREC);
    eq.setLinearTolerance(1e-12);
    int ret = eq.run();
    std::printf("{\"case\": \"elas2dq4\", \"ret\": %d, \"it\": %d, ", ret, eq.lastSolve().nb_it);
    print_vec("u", u);
    std::printf("},\n");
  }
  // 4b. SpMatrix(mesh, dof): dof != 0 gives NodeGraphScal's node pattern,
  // dof == 0 NodeGraph's equation pattern (SpMatrix_impl.h:391-413)
  {
    Mesh ms(0., 1., 0., 1., 5, 4, 1, 1, 1, 1, TRIANGLE, 1);
    SpMatrix<real_t> As(ms, 1), Ae(ms);
    std::printf("{\"case\": \"graph_scal\", \"n\": %zu, \"n_eq\": %zu, \"row_ptr\": [",
                As.getNbRows(), Ae.getNbRows());
    for (size_t i = 1; i <= As.getNbRows() + 1; ++i)
      std::printf("%s%zu", i > 1 ? "," : "", As.getRowPtr(i));
    std::printf("], \"col_ind\": [");
    for (size_t k = 1; k <= As.getLength(); ++k)
      std::printf("%s%zu", k > 1 ? "," : "", As.getColInd(k));
    std::printf("]},\n");
  }
  // 5. unsymmetric SpMatrix: BiCG, CGS, SpMatrix::TMult, Prec::TransSolve
  {
    const size_t n = 40;
    std::vector<size_t> rp(1, 0), ci;
    std::vector<real_t> a;
    for (size_t i = 1; i <= n; ++i) {
      if (i > 1) { ci.push_back(i - 1); a.push_back(-1.3); }
      ci.push_back(i); a.push_back(2.5);
      if (i < n) { ci.push_back(i + 1); a.push_back(-0.7); }
      rp.push_back(ci.size());
    }
    SpMatrix<real_t> A(rp, ci, a);
    Vect<real_t> b(n), x1(n), x2(n), v(n), t(n), z(n);
    for (size_t i = 1; i <= n; ++i) {
      b(i) = 1.0 + 0.1 * i;
      v(i) = std::sin(0.3 * i);
    }
    int itb = BiCG(A, DIAG_PREC, b, x1, 1000, 1e-12);
    int itc = CGS(A, IDENT_PREC, b, x2, 5, 1e-12);
    A.TMult(v, t);
    Prec<real_t> P(A, ILU_PREC);
    P.TransSolve(v, z);
    std::printf("{\"case\": \"unsym\", \"it_bicg\": %d, \"it_cgs\": %d, ", itb, itc);
    print_vec("x_bicg", x1);
    std::printf(", ");
    print_vec("x_cgs", x2);
    std::printf(", ");
    print_vec("tmult", t);
    std::printf(", ");
    print_vec("transsolve", z);
    std::printf("}%s\n", argc > 1 ? "," : "");
  }
  // 6. laplace_demo1 on a mesh file (argv[1]: the reference's laplace/square.m)
  if (argc > 1) {
    Mesh ms(argv[1]);
    Vect<real_t> u(ms), bc(ms);
    for (size_t n = 1; n <= ms.getNbNodes(); ++n) {
      real_t x = ms.getCoord(n, 0), y = ms.getCoord(n, 1);
      bc(n) = std::exp(x) * std::sin(y) + std::exp(y) * std::cos(x);
    }
    Laplace2DT3 eq(ms, u);
    eq.setInput(EType::BOUNDARY_CONDITION, bc);
    eq.setSolver(CG_SOLVER, DIAG_PREC);
    eq.setLinearTolerance(1e-12);
    int ret = eq.run();
    if (argc > 2) ms.put(argv[2]);
    std::printf("{\"case\": \"mesh_file\", \"ret\": %d, \"it\": %d, \"nodes\": %zu, "
                "\"elements\": %zu, \"sides\": %zu, ", ret, eq.lastSolve().nb_it,
                ms.getNbNodes(), ms.getNbElements(), ms.getNbSides());
    print_vec("u", u);
    std::printf("}\n");
  }
  std::printf("]\n");
  return 0;
}
