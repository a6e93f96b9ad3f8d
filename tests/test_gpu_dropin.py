"""The C++ drop-in layer (ofeli_amd/cxx, OFELI namespace) on the GPU: the
OFELI-style programs of tests/cpp/dropin_demo.cpp against the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "dropin_demo")
SQUARE = os.path.join(HERE, "golden", "meshes", "square.m")


@pytest.fixture(scope="module")
def cases(tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/dropin_demo not built (run tests/cpp/build.sh)")
    out = str(tmp_path_factory.mktemp("dropin") / "square_put.m")
    r = subprocess.run([EXE, SQUARE, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    cs = {c["case"]: c for c in json.loads(r.stdout)}
    cs["mesh_file"]["put"] = out
    return cs


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b)


def test_laplace2dt3_program(cases):
    c = cases["laplace2dt3"]
    m = O.mesh2d(16, 16)
    ue = np.exp(m.coord[:, 0]) * np.sin(m.coord[:, 1]) + np.exp(m.coord[:, 1]) * np.cos(m.coord[:, 0])
    dof, neq = O.number_equations(m.code)
    rp, ci = O.node_graph(m.elem, 1, dof, neq)
    a, b = O.assemble(O.LAPLACE2DT3, m, dof, neq, rp, ci, bc=ue)
    ret, v, _ = O.cg(rp, ci, a, b, max_it=2000, toler=1e-12)
    u = O.insert_bc(m.code, v, ue)
    assert c["ret"] == 0 and c["it"] == ret
    assert rel(c["u"], u) < 1e-11
    assert c["maxerr"] < 1e-3


def test_linear_solver_program(cases):
    c = cases["linearsolver"]
    n = 50
    A = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    x = np.linalg.solve(A, np.full(n, 1.0 / n))
    assert c["ret"] == 0 and c["it_cg"] > 0 and c["res"] < 1e-11
    assert rel(c["x"], x) < 1e-10
    # GMRES(25) reaches the tolerance, as the oracle does (GMRES(10) stagnates
    # on this problem in both)
    assert c["it_bicgstab"] > 0 and c["it_gmres"] > 0 and c["diff23"] < 1e-9
    a = cases["access"]
    assert a["a33"] == 5.0 and a["a34"] == -1.0 and a["a310"] == 0.0


def test_dc2dt3_backward_euler_program(cases):
    c = cases["dc2dt3_be"]
    m = O.mesh2d(12, 12)
    f = 1.0 + m.coord[:, 0]
    k = 1.0 + 0.5 * np.sin(0.1 * np.arange(1, m.nb_elements + 1))
    dof, neq = O.number_equations(m.code)
    rp, ci = O.node_graph(m.elem, 1, dof, neq)
    u = np.zeros(m.nb_nodes)
    for _ in range(5):
        a, b = O.assemble(O.DC2DT3_BE, m, dof, neq, rp, ci, f=f, coef=k, prm=[0.01, 1.0, 1.0],
                          u_prev=u)
        x0 = u[m.code[:, 0] == 0]
        ret, v, _ = O.cg(rp, ci, a, b, x0=x0, max_it=1000, toler=1e-12)
        u = O.insert_bc(m.code, v)
    assert rel(c["u"], u) < 1e-10


def test_elas2dq4_program(cases):
    c = cases["elas2dq4"]
    m = O.mesh2d(8, 2, quad=True, nb_dof=2, codes=(1, -1, 0, 0), box=(0.0, 4.0, 0.0, 1.0))
    m.code[:, 1] = m.code[:, 0]
    E = 1.0 + 0.1 * (np.arange(1, m.nb_elements + 1) % 3)
    dof, neq = O.number_equations(m.code)
    rp, ci = O.node_graph(m.elem, 2, dof, neq)
    a, b = O.assemble(O.ELAS2DQ4, m, dof, neq, rp, ci, coef=E, prm=[0.3, 0],
                      bc=np.zeros(2 * m.nb_nodes))
    tr = np.tile([0.0, -0.01], (m.sides.shape[0], 1))
    b = O.traction_rhs(m, m.sides, tr, dof, b)
    ret, v, _ = O.cg(rp, ci, a, b, prec=O.SSOR, max_it=1000, toler=1e-12)
    u = O.insert_bc(m.code, v)
    assert c["ret"] == 0 and abs(c["it"] - ret) <= 1
    assert rel(c["u"], u) < 1e-10


def test_unsymmetric_program(cases):
    """BiCG and CGS through the OFELI free functions, SpMatrix::TMult and
    Prec::TransSolve on a 40-point unsymmetric tridiagonal matrix."""
    c = cases["unsym"]
    n = 40
    rows, cols, vals = [], [], []
    for i in range(n):
        for j, v in ((i - 1, -1.3), (i, 2.5), (i + 1, -0.7)):
            if 0 <= j < n:
                rows.append(i), cols.append(j), vals.append(v)
    rp = np.searchsorted(rows, np.arange(n + 1)).astype(np.int64)
    ci = np.array(cols, np.int64) + 1
    a = np.array(vals)
    b = 1.0 + 0.1 * np.arange(1, n + 1)
    v = np.sin(0.3 * np.arange(1, n + 1))
    ret, x, _ = O.bicg(rp, ci, a, b, prec=O.DIAG, max_it=1000, toler=1e-12)
    assert ret > 0 and abs(c["it_bicg"] - ret) <= 1
    assert rel(c["x_bicg"], x) < 1e-9
    ret, x, _ = O.cgs(rp, ci, a, b, prec=O.IDENT, max_it=5, toler=1e-12)
    assert c["it_cgs"] == ret == -5
    assert rel(c["x_cgs"], x) < 1e-10
    assert np.array_equal(c["tmult"], O.spmv_t(rp, ci, a, v))
    assert np.array_equal(c["transsolve"], O.prec_trans_apply(rp, ci, a, O.ILU, v))


def test_mesh_file_program(cases):
    """laplace_demo1's pattern on the reference's own laplace/square.m read
    through Mesh(file): the solution equals the oracle's on the oracle's
    reading of the file, and Mesh::put writes what the oracle writes."""
    from oracle import mesh_io as MIO
    c = cases["mesh_file"]
    r = MIO.read(SQUARE)
    m = O.Mesh(r["coord"], r["code"], r["elem"].astype(np.int64), 1)
    x, y = m.coord[:, 0], m.coord[:, 1]
    ue = np.exp(x) * np.sin(y) + np.exp(y) * np.cos(x)
    dof, neq = O.number_equations(m.code)
    rp, ci = O.node_graph(m.elem, 1, dof, neq)
    a, b = O.assemble(O.LAPLACE2DT3, m, dof, neq, rp, ci, bc=ue)
    ret, v, _ = O.cg(rp, ci, a, b, max_it=2000, toler=1e-12)
    u = O.insert_bc(m.code, v, ue)
    assert (c["nodes"], c["elements"], c["sides"]) == (142, 242, 40)
    assert c["ret"] == 0 and abs(c["it"] - ret) <= 1
    assert rel(c["u"], u) < 1e-10
    MIO.write(c["put"] + ".oracle", r)
    assert open(c["put"], "rb").read() == open(c["put"] + ".oracle", "rb").read()


def test_spmatrix_mesh_dof_pattern(cases):
    """SpMatrix(mesh, dof != 0) has NodeGraphScal's pattern (one row per
    node, no elimination; GraphOfMatrix.cpp:532-546 as SpMatrix_impl.h:399-402
    selects it), SpMatrix(mesh) NodeGraph's over the equations."""
    c = cases["graph_scal"]
    m = O.mesh2d(5, 4)
    rp, ci = O.node_graph_scal(m.elem, m.nb_nodes)
    assert c["n"] == m.nb_nodes == 30
    assert np.array_equal(np.array(c["row_ptr"]), rp)
    assert np.array_equal(np.array(c["col_ind"]), ci)
    dof, neq = O.number_equations(m.code)
    assert c["n_eq"] == neq == 12
