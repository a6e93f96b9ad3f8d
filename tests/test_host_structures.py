"""Host-side structures against the oracle and closed forms (CPU only).

* The product's structured mesh generators (the C-ABI's host entry points,
  used by the 3D Poisson and cantilever workloads) equal the oracle's arrays
  exactly.
* NodeGraphScal, the pattern of SpMatrix(mesh, dof != 0), pinned by brute
  force.
* The Elas2DQ4 body-force right-hand side pinned by the Q4 consistent mass
  matrix (2x2 Gauss is exact for bilinear x bilinear on rectangles).
* The mesh-file reader rejects files whose records disagree in width.
"""
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("dims", [(20, 20, 20), (23, 17, 13), (1, 2, 3)])
def test_product_kuhn_generator_equals_oracle(dims):
    from ofeli_amd import api
    coord, code, elem = api.mesh3d_kuhn(*dims)
    m = O.mesh3d_kuhn(*dims)
    assert np.array_equal(coord, m.coord)
    assert np.array_equal(code, m.code)
    assert np.array_equal(elem.astype(np.int64), m.elem)


@pytest.mark.parametrize("codes", [(1, 1, 1, 1), (1, 0, 0, 0), (2, -1, 3, 0)])
def test_product_kuhn_generator_codes_and_box(codes):
    from ofeli_amd import api
    c6 = codes + codes[:2]
    box = (0.0, 2.0, -1.0, 1.0, 0.5, 1.0)
    coord, code, elem = api.mesh3d_kuhn(4, 3, 5, codes=c6, box=box)
    m = O.mesh3d_kuhn(4, 3, 5, codes=c6, box=box)
    assert np.array_equal(coord, m.coord) and np.array_equal(code, m.code)
    assert np.array_equal(elem.astype(np.int64), m.elem)


@pytest.mark.parametrize("quad,nb_dof,codes,box", [
    (True, 2, (1, 0, 0, 0), (0.0, 4.0, 0.0, 1.0)),
    (True, 1, (1, 1, 1, 1), (0.0, 1.0, 0.0, 1.0)),
    (False, 1, (1, 2, 3, 4), (0.0, 1.0, -0.5, 0.5)),
    (False, 2, (0, 1, 0, 1), (0.0, 3.0, 0.0, 1.0)),
])
def test_product_2d_generator_equals_oracle(quad, nb_dof, codes, box):
    from ofeli_amd import api
    coord, code, elem = api.mesh2d(32, 8, quad=quad, nb_dof=nb_dof, codes=codes, box=box)
    m = O.mesh2d(32, 8, quad=quad, nb_dof=nb_dof, codes=codes, box=box)
    assert np.array_equal(coord, m.coord)
    assert np.array_equal(code, m.code)
    assert np.array_equal(elem.astype(np.int64), m.elem)


def test_cantilever_workload_mesh_equals_oracle():
    """The config-4 generator path (2 DOF quads, left edge clamped in both
    DOFs) at a reduced size, through the numbering and the pattern."""
    from ofeli_amd import workloads as W
    d = W.config4(64, 8, ncases=2)
    m = O.mesh2d(64, 8, quad=True, nb_dof=2, codes=(1, 0, 0, 0), box=(0.0, 4.0, 0.0, 1.0))
    m.code[:, 1] = m.code[:, 0]
    assert np.array_equal(d["coord"], m.coord) and np.array_equal(d["code"], m.code)
    assert np.array_equal(d["elem"].astype(np.int64), m.elem)
    dof, neq = O.number_equations(m.code)
    assert neq == 2 * 64 * 9


# ---------------------------------------------------------------- NodeGraphScal
def test_node_graph_scal_brute_force():
    m = O.mesh2d(6, 4)
    rp, ci = O.node_graph_scal(m.elem, m.nb_nodes)
    P = np.zeros((m.nb_nodes, m.nb_nodes), bool)
    for en in m.elem:
        P[np.ix_(en - 1, en - 1)] = True
    assert np.array_equal(rp, np.concatenate([[0], np.cumsum(P.sum(1))]))
    assert np.array_equal(ci, np.nonzero(P)[1] + 1)


def test_node_graph_scal_ignores_codes_unlike_node_graph():
    """Boundary nodes keep their rows: the pattern is NodeGraph's with no
    imposed DOF, over all nodes, and larger than the eliminated one."""
    m = O.mesh3d_kuhn(3, 2, 2)
    rp, ci = O.node_graph_scal(m.elem, m.nb_nodes)
    free = np.zeros_like(m.code)
    dof, neq = O.number_equations(free)
    rp2, ci2 = O.node_graph(m.elem, 1, dof, neq)
    assert neq == m.nb_nodes
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2)
    dofe, neqe = O.number_equations(m.code)
    rpe, _ = O.node_graph(m.elem, 1, dofe, neqe)
    assert rpe[-1] < rp[-1]


# ---------------------------------------------------------------- Q4 body force
def q4_mass(hx, hy):
    return hx * hy / 36.0 * np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]])


def test_q4_body_force_is_consistent_mass_times_nodal_force():
    """One rectangular element: eRHS_x = M f_x, eRHS_y = M f_y with the Q4
    consistent mass matrix; a node shift (node j's force with node j-1's
    shape function) fails this check."""
    hx, hy = 0.75, 0.5
    m = O.mesh2d(1, 1, quad=True, nb_dof=2, codes=(0, 0, 0, 0), box=(0.0, hx, 0.0, hy))
    dof, neq = O.number_equations(m.code)
    rp, ci = O.node_graph(m.elem, 2, dof, neq)
    rng = np.random.default_rng(4)
    f = rng.standard_normal(2 * m.nb_nodes)
    _, b = O.assemble(O.ELAS2DQ4, m, dof, neq, rp, ci, f=f, coef=np.ones(1), prm=[0.3, 1])
    en = m.elem[0] - 1
    M = q4_mass(hx, hy)
    fx, fy = f[2 * en], f[2 * en + 1]
    ex = np.zeros(2 * m.nb_nodes)
    ex[2 * en] = M @ fx
    ex[2 * en + 1] = M @ fy
    assert np.allclose(b, ex, rtol=1e-13, atol=1e-15)


def test_q4_constant_body_force_integrates_to_total_load():
    """A constant force f on a clamped-free strip: the right-hand side of the
    free DOFs plus the eliminated ones is f times the area; interior nodes of
    a uniform grid receive f h^2 (four quarter cells)."""
    nx, ny = 8, 4
    m = O.mesh2d(nx, ny, quad=True, nb_dof=2, codes=(0, 0, 0, 0), box=(0.0, 2.0, 0.0, 1.0))
    dof, neq = O.number_equations(m.code)
    rp, ci = O.node_graph(m.elem, 2, dof, neq)
    f = np.tile([0.3, -1.0], m.nb_nodes)
    _, b = O.assemble(O.ELAS2DQ4, m, dof, neq, rp, ci, f=f, coef=np.ones(m.nb_elements),
                      prm=[0.3, 1])
    assert np.isclose(b[0::2].sum(), 0.3 * 2.0, rtol=1e-13)
    assert np.isclose(b[1::2].sum(), -1.0 * 2.0, rtol=1e-13)
    h2 = (2.0 / nx) * (1.0 / ny)
    inner = [i * (ny + 1) + j for i in range(1, nx) for j in range(1, ny)]
    assert np.allclose(b[1::2][inner], -h2, rtol=1e-13)


# ---------------------------------------------------------------- mesh files
SQUARE = os.path.join(HERE, "golden", "meshes", "square.m")


def _mixed_sections(text):
    """An extra <Elements nodes="4"> section before the triangles."""
    extra = '<Elements shape="quadrilateral" nodes="4">\n 1 2 3 4 1\n</Elements>\n'
    return text.replace("<Elements", extra + "<Elements", 1)


@pytest.mark.parametrize("mutate", [
    _mixed_sections,
    lambda t: t.replace("</Mesh>", '</Mesh>\n<Mesh dim="3" nb_dof="1">\n</Mesh>', 1),
    lambda t: re.sub(r'(<Sides[^>]*nodes=")2(")', r'\g<1>3\2', t, count=1)
    if 'nodes="2"' in t else t.replace("<Sides>", '<Sides nodes="3">', 1),
])
def test_reader_rejects_records_of_mixed_widths(mutate, tmp_path):
    import ofeli_amd as oa
    text = open(SQUARE, encoding="latin-1").read()
    bad = mutate(text)
    assert bad != text
    p = tmp_path / "bad.m"
    p.write_text(bad, encoding="latin-1")
    with pytest.raises(oa.OfhError):
        oa.read_mesh(str(p))


def test_mesh_get_checks_buffer_lengths():
    import ctypes as C
    import ofeli_amd as oa
    L = oa.lib()
    h = C.c_void_p()
    oa.check(L.ofh_mesh_read(SQUARE.encode(), C.byref(h)))
    try:
        v = [C.c_int() for _ in range(10)]
        oa.check(L.ofh_mesh_info(h, *[C.byref(x) for x in v]))
        nn, ne, nen = v[2].value, v[3].value, v[5].value
        elem = np.zeros(ne * nen - 1, np.int32)  # one entry short
        rc = L.ofh_mesh_get(h, None, 0, None, 0, elem.ctypes.data_as(C.POINTER(C.c_int32)),
                            elem.size, None, 0, None, 0, None, 0)
        assert rc == -1
        elem = np.zeros(ne * nen, np.int32)
        assert L.ofh_mesh_get(h, None, 0, None, 0, elem.ctypes.data_as(C.POINTER(C.c_int32)),
                              elem.size, None, 0, None, 0, None, 0) == 0
        assert elem.min() >= 1 and elem.max() <= nn
    finally:
        L.ofh_mesh_destroy(h)
