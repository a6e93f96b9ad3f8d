"""Pins of the CPU oracle (CPU only).

The reference's tests contain no expected values, so each oracle function is
checked against what the mathematics fixes: the golden fixtures of
tests/golden (textbook element formulas evaluated with dense numpy algebra),
closed-form nonzero counts, stencils, invariants, dense solves and
brute-force patterns.  Each check is chosen so that a plausible slip (a
dropped term, a transposed operand, a wrong index) fails it.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fe_fixtures.npz"))


def dense(rp, ci, a):
    return O.to_dense(rp, ci, a)


# ---------------------------------------------------------------- meshes
def test_mesh2d_layout_and_orientation():
    m = O.mesh2d(5, 3, box=(0.0, 1.0, 0.0, 0.6))
    assert m.nb_nodes == 24 and m.nb_elements == 30
    # node id = i*(ny+1)+j+1, x outer
    assert np.allclose(m.coord[7], [0.2, 0.6]) and np.allclose(m.coord[4], [0.2, 0.0])
    for en in m.elem:
        x = m.coord[en - 1]
        det = (x[1, 0] - x[0, 0]) * (x[2, 1] - x[0, 1]) - (x[1, 1] - x[0, 1]) * (x[2, 0] - x[0, 0])
        assert det > 0
    # boundary codes on DOF 1 only where the reference puts them
    onb = (np.isclose(m.coord[:, 0], 0) | np.isclose(m.coord[:, 0], 1) | np.isclose(m.coord[:, 1], 0)
           | np.isclose(m.coord[:, 1], 0.6))
    assert np.array_equal(m.code[:, 0] != 0, onb)


def test_kuhn_tets_positive_volume_and_tiling():
    m = O.mesh3d_kuhn(3, 2, 2, box=(0, 1.5, 0, 1, 0, 2))
    vols = []
    for en in m.elem:
        rc, vol, _ = O.tetra4(m.coord[en - 1])
        assert rc == 0 and vol > 0
        vols.append(vol)
    assert np.isclose(sum(vols), 1.5 * 1 * 2, rtol=1e-13)
    assert np.allclose(vols, GOLD["p1tet_vol"][0] * (0.5 * 0.5 * 1.0) / (1 / 27), rtol=1e-12)


def test_number_equations_eliminates_coded_dofs():
    code = np.array([[0, 1], [1, 0], [0, 0]], np.int32)
    dof, neq = O.number_equations(code)
    assert neq == 4 and dof.tolist() == [[1, 0], [0, 2], [3, 4]]


# ---------------------------------------------------------------- pattern
def brute_pattern(m, dof, neq):
    P = np.zeros((neq, neq), bool)
    for en in m.elem:
        d = np.array([dof[v - 1, c] for v in en for c in range(m.nb_dof)])
        d = d[d > 0] - 1
        P[np.ix_(d, d)] = True
    rp = np.concatenate([[0], np.cumsum(P.sum(1))])
    ci = np.nonzero(P)[1] + 1
    return rp, ci


@pytest.mark.parametrize("case", ["tri", "quad2", "tet"])
def test_node_graph_brute_force(case):
    if case == "tri":
        m = O.mesh2d(7, 5)
    elif case == "quad2":
        m = O.mesh2d(6, 3, quad=True, nb_dof=2, codes=(1, 0, 0, 0))
        m.code[:, 1] = m.code[:, 0]
    else:
        m = O.mesh3d_kuhn(4, 3, 3)
    dof, neq = O.number_equations(m.code)
    rp, ci = O.node_graph(m.elem, m.nb_dof, dof, neq)
    brp, bci = brute_pattern(m, dof, neq)
    assert np.array_equal(rp, brp) and np.array_equal(ci, bci)


def test_node_graph_closed_form_counts():
    # 2D P1, n x n cells, interior m = n-1 per side: m^2 + 4m(m-1) + 2(m-1)^2
    for n in (4, 9, 128):
        m = O.mesh2d(n, n)
        dof, neq = O.number_equations(m.code)
        rp, _ = O.node_graph(m.elem, 1, dof, neq)
        k = n - 1
        assert rp[-1] == k * k + 4 * k * (k - 1) + 2 * (k - 1) ** 2
    assert k * k == 16129 and rp[-1] == 111889  # the 128 x 128 Poisson system
    # Kuhn: m^3 + 6 m^2 (m-1) + 6 m (m-1)^2 + 2 (m-1)^3
    for n in (3, 7):
        m = O.mesh3d_kuhn(n, n, n)
        dof, neq = O.number_equations(m.code)
        rp, _ = O.node_graph(m.elem, 1, dof, neq)
        k = n - 1
        assert rp[-1] == k ** 3 + 6 * k * k * (k - 1) + 6 * k * (k - 1) ** 2 + 2 * (k - 1) ** 3
    # Q4, 2 DOF per node, A x B free nodes: 4 (3A-2)(3B-2)
    m = O.mesh2d(8, 3, quad=True, nb_dof=2, codes=(1, 0, 0, 0))
    m.code[:, 1] = m.code[:, 0]
    dof, neq = O.number_equations(m.code)
    rp, _ = O.node_graph(m.elem, 2, dof, neq)
    assert rp[-1] == 4 * (3 * 8 - 2) * (3 * 4 - 2)


# ---------------------------------------------------------------- shapes
def test_gauss_two_points():
    x, w = O.gauss(2)
    # Newton's iteration (Gauss.cpp:37-54) lands within one ulp of the exact rule
    assert np.all(np.abs(x - [-1 / np.sqrt(3), 1 / np.sqrt(3)]) <= np.spacing(0.6))
    assert np.all(np.abs(w - 1.0) <= np.spacing(1.0))
    x3, w3 = O.gauss(3)
    assert np.allclose(x3, [-np.sqrt(0.6), 0, np.sqrt(0.6)], atol=1e-15)
    assert np.allclose(w3, [5 / 9, 8 / 9, 5 / 9], atol=1e-15)


def test_triangle_and_tetra_geometry():
    rc, area, d = O.triang3([[0, 0], [2, 0], [0, 1]])
    assert rc == 0 and area == 1.0
    # gradients of barycentric coordinates sum to zero and reproduce x, y
    assert np.allclose(d.sum(0), 0) and np.allclose(d.T @ np.array([0, 2, 0]), [1, 0])
    assert O.triang3([[0, 0], [0, 1], [2, 0]])[0] == -1  # negative orientation rejected
    rc, vol, d = O.tetra4([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 3]])
    assert rc == 0 and np.isclose(vol, 0.5)
    assert np.allclose(d.T @ np.array([0, 1, 0, 0]), [1, 0, 0])
    assert np.allclose(d.T @ np.array([0, 0, 0, 3]), [0, 0, 1])


# ---------------------------------------------------------------- elements
def assemble_case(kind, m, **kw):
    dof, neq = O.number_equations(m.code)
    rp, ci = O.node_graph(m.elem, m.nb_dof, dof, neq)
    a, b = O.assemble(kind, m, dof, neq, rp, ci, **kw)
    return rp, ci, a, b, dof


def test_p1_triangle_matches_golden():
    m = O.mesh2d(6, 5)
    rp, ci, a, b, _ = assemble_case(O.DC2DT3, m, coef=GOLD["p1tri_k"], f=GOLD["p1tri_f"])
    assert np.array_equal(rp, GOLD["p1tri_rp"]) and np.array_equal(ci - 1, GOLD["p1tri_col"])
    K = dense(rp, ci, a)
    assert np.allclose(K, GOLD["p1tri_K"], rtol=0, atol=1e-13 * np.abs(K).max())
    assert np.allclose(b, GOLD["p1tri_F"], rtol=1e-14)
    ret, x, _ = O.cg(rp, ci, a, b, max_it=1000, toler=1e-14)
    assert ret > 0 and np.allclose(x, GOLD["p1tri_x"], rtol=1e-11, atol=0)


def test_p1_triangle_is_five_point_stencil():
    n = 10
    m = O.mesh2d(n, n)
    rp, ci, a, b, _ = assemble_case(O.LAPLACE2DT3, m)
    K = dense(rp, ci, a)
    N = n - 1
    T = 2 * np.eye(N) - np.eye(N, k=1) - np.eye(N, k=-1)
    L5 = np.kron(T, np.eye(N)) + np.kron(np.eye(N), T)
    assert np.allclose(K, L5, rtol=0, atol=1e-14)


def test_p1_kuhn_is_seven_point_stencil_times_h():
    n = 5
    m = O.mesh3d_kuhn(n, n, n)
    rp, ci, a, b, _ = assemble_case(O.LAPLACE3DT4, m)
    N = n - 1
    T = 2 * np.eye(N) - np.eye(N, k=1) - np.eye(N, k=-1)
    I = np.eye(N)
    L7 = np.kron(np.kron(T, I), I) + np.kron(np.kron(I, T), I) + np.kron(np.kron(I, I), T)
    assert np.allclose(dense(rp, ci, a), L7 / n, rtol=0, atol=1e-14)
    # DC3DT4 form with k = 1 gives the same matrix
    rp2, ci2, a2, _, _ = assemble_case(O.LAPLACE3DT4, m, coef=np.ones(m.nb_elements))
    assert np.allclose(a2, a, rtol=1e-15, atol=1e-17)


def test_p1_tet_matches_golden():
    m = O.mesh3d_kuhn(3, 3, 3)
    rp, ci, a, b, _ = assemble_case(O.LAPLACE3DT4, m)
    assert np.allclose(dense(rp, ci, a), GOLD["p1tet_K"], rtol=0, atol=1e-14)


def test_zero_row_sums_and_exact_symmetry():
    m = O.mesh2d(6, 6, codes=(0, 0, 0, 0))
    k = np.exp(np.random.default_rng(0).standard_normal(m.nb_elements))
    rp, ci, a, _, _ = assemble_case(O.DC2DT3, m, coef=k)
    K = dense(rp, ci, a)
    assert np.allclose(K.sum(1), 0, atol=1e-13 * np.abs(K).max())
    assert np.array_equal(K, K.T)  # bitwise: same element terms, same order
    m3 = O.mesh3d_kuhn(3, 3, 3, codes=(0,) * 6)
    rp, ci, a, _, _ = assemble_case(O.LAPLACE3DT4, m3, coef=np.exp(np.arange(m3.nb_elements) % 3))
    K = dense(rp, ci, a)
    assert np.allclose(K.sum(1), 0, atol=1e-13 * np.abs(K).max()) and np.array_equal(K, K.T)


@pytest.mark.parametrize("mode", [0, 1])
def test_q4_elasticity_matches_golden(mode):
    m = O.mesh2d(4, 2, quad=True, nb_dof=2, codes=(1, 0, 0, 0), box=(0.0, 4.0, 0.0, 1.0))
    m.code[:, 1] = m.code[:, 0]
    rp, ci, a, b, _ = assemble_case(O.ELAS2DQ4, m, coef=GOLD["q4_E"], prm=[0.3, mode])
    K = dense(rp, ci, a)
    G = GOLD["q4_K_mode%d" % mode]
    assert np.allclose(K, G, rtol=0, atol=1e-13 * np.abs(G).max())
    # symmetric to rounding only: a.y*b.x and a.x*b.y round differently
    assert np.allclose(K, K.T, rtol=0, atol=1e-15 * np.abs(K).max())


@pytest.mark.parametrize("mode", [0, 1])
def test_q4_rigid_modes_in_kernel(mode):
    x = np.array([[0.1, 0.0], [1.3, 0.2], [1.1, 0.9], [0.0, 1.2]])
    rc, K = O.element(O.ELAS2DQ4, x, 2.5, [0.3, mode])
    assert rc == 0
    tx = np.tile([1.0, 0.0], 4)
    ty = np.tile([0.0, 1.0], 4)
    rot = np.ravel(np.column_stack([-x[:, 1], x[:, 0]]))
    for v in (tx, ty, rot):
        assert np.abs(K @ v).max() <= 1e-13 * np.abs(K).max()
    w = np.linalg.eigvalsh(K)
    assert (w > -1e-12 * w.max()).all() and (w > 1e-10 * w.max()).sum() == 5


def test_backward_euler_element():
    x = np.array([[0.0, 0.0], [0.2, 0.0], [0.0, 0.1]])
    rc, K = O.element(O.DC2DT3_BE, x, 1.7, [1e-3, 2.0, 3.0])
    rc2, K0 = O.element(O.DC2DT3, x, 1.7)
    area = 0.01
    assert np.allclose(K - K0, np.eye(3) * area / 3 * 2.0 * 3.0 / 1e-3, rtol=1e-14)


def test_dirichlet_lifting_and_manufactured_convergence():
    """Laplace2DT3 with the harmonic exp(x)sin(y)+exp(y)cos(x) of
    demos/laplace/proj2.dat as Dirichlet data: nodal error O(h^2)."""
    errs = []
    for n in (8, 16, 32):
        m = O.mesh2d(n, n)
        ue = np.exp(m.coord[:, 0]) * np.sin(m.coord[:, 1]) + np.exp(m.coord[:, 1]) * np.cos(m.coord[:, 0])
        rp, ci, a, b, dof = assemble_case(O.LAPLACE2DT3, m, bc=ue)
        ret, v, _ = O.cg(rp, ci, a, b, max_it=5000, toler=1e-13)
        u = O.insert_bc(m.code, v, ue)
        assert np.array_equal(u[m.code[:, 0] != 0], ue[m.code[:, 0] != 0])
        errs.append(np.sqrt(np.mean((u - ue) ** 2)))
    assert 3.3 < errs[0] / errs[1] < 4.7 and 3.3 < errs[1] / errs[2] < 4.7


# ---------------------------------------------------------------- solvers
def spd_system(n=40, seed=0):
    m = O.mesh2d(n, n)
    k = np.exp(np.random.default_rng(seed).standard_normal(m.nb_elements))
    f = 1 + np.random.default_rng(seed + 1).random(m.nb_nodes)
    rp, ci, a, b, _ = assemble_case(O.DC2DT3, m, coef=k, f=f)
    return rp, ci, a, b


@pytest.mark.parametrize("prec", [O.IDENT, O.DIAG, O.SSOR])
def test_cg_solves_dense_reference(prec):
    rp, ci, a, b = spd_system(12)
    ret, x, res = O.cg(rp, ci, a, b, prec=prec, max_it=2000, toler=1e-13)
    xd = np.linalg.solve(dense(rp, ci, a), b)
    assert ret > 0 and res <= 1e-13
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-11


def test_cg_exact_termination_small_system():
    rng = np.random.default_rng(3)
    Q = np.linalg.qr(rng.standard_normal((8, 8)))[0]
    A = Q @ np.diag(np.arange(1.0, 9.0)) @ Q.T
    rp = np.arange(0, 65, 8, dtype=np.int64)
    ci = np.tile(np.arange(1, 9, dtype=np.int64), 8)
    ret, x, _ = O.cg(rp, ci, A.ravel(), np.ones(8), prec=O.IDENT, max_it=50, toler=1e-12)
    assert 0 < ret <= 9


def test_cg_return_conventions():
    rp, ci, a, b = spd_system(10)
    ret, _, _ = O.cg(rp, ci, a, b, max_it=3, toler=1e-14)
    assert ret == -4  # -(max_it+1), CG.h:154-156
    xd = np.linalg.solve(dense(rp, ci, a), b)
    ret, _, _ = O.cg(rp, ci, a, b, x0=xd, toler=1e-8)
    assert ret == 0


def test_diag_prec_and_sgs_definition():
    rp, ci, a, b = spd_system(6)
    A = dense(rp, ci, a)
    D = np.diag(np.diag(A))
    L = np.tril(A, -1)
    U = np.triu(A, 1)
    M = (D + L) @ np.linalg.inv(D) @ (D + U)
    r = np.random.default_rng(2).standard_normal(A.shape[0])
    z = O.ssor_solve(rp, ci, a, r)
    assert np.allclose(z, np.linalg.solve(M, r), rtol=1e-12, atol=1e-14)
    # a transposed operand (D+U) D^-1 (D+L) would not match on this nonsymmetric grid order
    Mt = (D + U) @ np.linalg.inv(D) @ (D + L)
    assert not np.allclose(z, np.linalg.solve(Mt, r), rtol=1e-6)


@pytest.mark.parametrize("prec", [O.IDENT, O.DIAG, O.SSOR])
def test_bicgstab_and_gmres_dense_reference(prec):
    rp, ci, a, b = spd_system(10)
    xd = np.linalg.solve(dense(rp, ci, a), b)
    ret, x, _ = O.bicgstab(rp, ci, a, b, prec=prec, max_it=2000, toler=1e-12)
    assert ret > 0 and np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9
    ret, x, _ = O.gmres(rp, ci, a, b, 15, prec=prec, max_it=2000, toler=1e-12)
    assert ret > 0 and np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9


def test_gmres_full_space_terminates():
    """GMRES(m) with m >= n reaches the exact solution within n steps."""
    rng = np.random.default_rng(4)
    n = 12
    A = rng.standard_normal((n, n)) + n * np.eye(n)
    rp = np.arange(0, n * n + 1, n, dtype=np.int64)
    ci = np.tile(np.arange(1, n + 1, dtype=np.int64), n)
    b = rng.standard_normal(n)
    ret, x, _ = O.gmres(rp, ci, A.ravel(), b, n, prec=O.IDENT, max_it=n, toler=1e-12)
    assert 0 < ret <= n
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9)


def test_bicgstab_breakdown_code():
    """rho = (r~, r) = 0 at the first step returns 2 (BiCGStab.h:117-120):
    r0 != 0 but r~ = r0, so rho = ||r0||^2 > 0 -- force rho = 0 via a
    skew operator: A = [[0,1],[-1,0]], b = (1,0): r~.v = 0 -> alpha = inf."""
    rp = np.array([0, 1, 2], np.int64)
    ci = np.array([2, 1], np.int64)
    a = np.array([1.0, -1.0])
    ret, x, _ = O.bicgstab(rp, ci, a, np.array([1.0, 0.0]), prec=O.IDENT, max_it=5, toler=1e-12)
    assert ret in (2, 3) or ret < 0


def test_traction_rhs_is_edge_integral():
    """Constant traction per side: each end node of a side receives half the
    side length times the traction (2-point Gauss is exact for it); the total
    equals traction x loaded length; clamped DOFs receive nothing."""
    m = O.mesh2d(6, 3, quad=True, nb_dof=2, codes=(1, -2, 0, 0), box=(0.0, 4.0, 0.0, 1.0))
    m.code[:, 1] = m.code[:, 0]
    dof, neq = O.number_equations(m.code)
    right = m.sides[m.side_code == 2]  # negative codes make sides (Mesh.cpp:617-669)
    assert right.shape[0] == 3 and np.allclose(m.coord[right - 1, 0], 4.0)
    t = np.tile([0.25, -1.5], (3, 1))
    b = O.traction_rhs(m, right, t.ravel(), dof, np.zeros(neq))
    assert np.isclose(b[0::2].sum(), 0.25 * 1.0) and np.isclose(b[1::2].sum(), -1.5 * 1.0)
    ys = m.coord[:, 1]
    for node in np.unique(right):
        h = 1.0 / 3.0
        share = h if 0.0 < ys[node - 1] < 1.0 else h / 2
        assert np.isclose(b[dof[node - 1, 1] - 1], -1.5 * share)
        assert np.isclose(b[dof[node - 1, 0] - 1], 0.25 * share)
    # a traction on the clamped edge changes nothing
    left = np.array([[1, 2], [2, 3], [3, 4]])  # nodes of x = 0
    assert np.allclose(m.coord[left - 1, 0], 0.0)
    b2 = O.traction_rhs(m, left, np.tile([1.0, 1.0], len(left)), dof, np.zeros(neq))
    assert not b2.any()
