"""Meshes of the reference's demos, read from their .m files by the library's
reader, assembled on the GPU and solved: bitwise the oracle's assembly on the
oracle's reading of the same file (laplace/square.m is the mesh of the
reference's laplace_demo1, solid/lelas2d/beam.m of its 2D elasticity demo)."""
import numpy as np
import pytest

from oracle import mesh_io as MIO
from oracle import oracle as O
from tests.test_gpu_assembly import check_plan, dev, host
from tests.test_mesh_io import path

pytestmark = pytest.mark.gpu


def both(name, nb_dof=None):
    """(library reading, oracle mesh object of the oracle's reading)."""
    import ofeli_amd as oa
    c = oa.read_mesh(path(name))
    r = MIO.read(path(name))
    assert np.array_equal(c.coord, r["coord"]) and np.array_equal(c.elem, r["elem"])
    code = r["code"] if nb_dof is None else np.ascontiguousarray(r["code"][:, :nb_dof])
    m = O.Mesh(r["coord"], code, r["elem"].astype(np.int64), code.shape[1])
    return c, m


def test_laplace_on_square_m(ctx):
    """laplace_demo1 on square.m: harmonic Dirichlet data on the coded
    boundary nodes, assembled bitwise, then CG to the oracle's solution."""
    import ofeli_amd as oa
    import torch
    c, m = both("square")
    x, y = m.coord[:, 0], m.coord[:, 1]
    bc = np.exp(x) * np.sin(y) + np.exp(y) * np.cos(x)
    P, a, b, dof = check_plan(ctx, m, O.LAPLACE2DT3, bc=bc, f=np.zeros(m.nb_nodes))
    rp, ci = O.node_graph(m.elem, 1, dof, P.nb_eq)
    ret, xo, _ = O.cg(rp, ci, a, b, max_it=2000, toler=1e-12)
    xg = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
    info = oa.cg(P.A, dev(b), xg, max_it=2000, toler=1e-12)
    assert ret > 0 and abs(info.nb_it - ret) <= 1
    assert np.linalg.norm(host(xg) - xo) <= 1e-10 * np.linalg.norm(xo)
    u = O.insert_bc(m.code, host(xg), bc)
    assert np.abs(u - bc).max() < 5e-3  # P1 interpolation of a harmonic function


def test_elas2dq4_on_beam_m(ctx):
    c, m = both("beam_q4")
    rng = np.random.default_rng(7)
    E = np.exp(0.3 * rng.standard_normal(m.nb_elements))
    check_plan(ctx, m, O.ELAS2DQ4, coef=E, prm=[0.3, 1], bc=np.zeros(2 * m.nb_nodes))


def test_laplace3dt4_on_tet_beam_m(ctx):
    """The 3D beam's tetrahedra with the first DOF's codes as a scalar problem."""
    c, m = both("beam_tet", nb_dof=1)
    rng = np.random.default_rng(8)
    check_plan(ctx, m, O.LAPLACE3DT4, f=1 + rng.random(m.nb_nodes),
               coef=np.exp(rng.standard_normal(m.nb_elements)), bc=rng.standard_normal(m.nb_nodes))
