"""GPU assembly parity: CSR pattern, matrix values and right-hand side of every
element kind are bitwise equal to the oracle's element-order assembly."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _t():
    import torch
    return torch


def dev(a):
    return _t().as_tensor(np.ascontiguousarray(a, np.float64), device="cuda")


def host(t):
    _t().cuda.synchronize()
    return t.cpu().numpy()


def check_plan(ctx, m, kind, f=None, coef=None, bc=None, prm=None, u_prev=None):
    import ofeli_amd as oa
    P = oa.AssemblyPlan(ctx, m.coord, m.elem.astype(np.int32), m.code)
    dof, neq = O.number_equations(m.code)
    assert P.nb_eq == neq
    assert np.array_equal(P.dof(), dof)
    rp, ci = O.node_graph(m.elem, m.nb_dof, dof, neq)
    grp, gci = P.A.pattern()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci - 1)
    a, b = O.assemble(kind, m, dof, neq, rp, ci, f=f, coef=coef, bc=bc, prm=prm, u_prev=u_prev)
    gb = _t().zeros(neq, dtype=_t().float64, device="cuda")
    P.build(kind, coef=None if coef is None else dev(coef), f=None if f is None else dev(f),
            bc=None if bc is None else dev(bc), u_prev=None if u_prev is None else dev(u_prev),
            prm=() if prm is None else tuple(prm), matrix=True, b=gb)
    assert np.array_equal(P.A.values(), a)
    assert np.array_equal(host(gb), b)
    return P, a, b, dof


def test_laplace2dt3_harmonic_bc(ctx):
    """Nonzero Dirichlet data exercise updateBC (Equation_impl.h:160-176)."""
    m = O.mesh2d(12, 9, codes=(1, 1, 1, 1), box=(0.0, 1.0, 0.0, 0.7))
    x, y = m.coord[:, 0], m.coord[:, 1]
    bc = np.exp(x) * np.sin(y) + np.exp(y) * np.cos(x)
    f = np.zeros(m.nb_nodes) + 0.25
    check_plan(ctx, m, O.LAPLACE2DT3, f=f, bc=bc)


def test_dc2dt3_lognormal(ctx):
    m = O.mesh2d(20, 17)
    rng = np.random.default_rng(0)
    k = np.exp(0.75 * rng.standard_normal(m.nb_elements))
    check_plan(ctx, m, O.DC2DT3, f=1 + rng.random(m.nb_nodes), coef=k)


def test_dc2dt3_backward_euler(ctx):
    m = O.mesh2d(16, 16)
    rng = np.random.default_rng(1)
    k = np.exp(rng.standard_normal(m.nb_elements))
    check_plan(ctx, m, O.DC2DT3_BE, f=rng.random(m.nb_nodes), coef=k, prm=[1e-3, 1.0, 1.0],
               u_prev=rng.standard_normal(m.nb_nodes), bc=np.zeros(m.nb_nodes))


@pytest.mark.parametrize("with_k", [False, True])
def test_p1_tetra(ctx, with_k):
    m = O.mesh3d_kuhn(6, 5, 4)
    rng = np.random.default_rng(2)
    k = np.exp(rng.standard_normal(m.nb_elements)) if with_k else None
    check_plan(ctx, m, O.LAPLACE3DT4, f=1 + rng.random(m.nb_nodes), coef=k,
               bc=rng.standard_normal(m.nb_nodes))


def test_p1_tetra_jittered(ctx):
    """Kuhn mesh of [-1, 1.3] x [-0.7, 0.9] x [-1.1, 0.2] with every interior
    node moved by up to 0.15 of a cell: general tetrahedra, coordinates of
    both signs, bitwise."""
    m = O.mesh3d_kuhn(7, 6, 5, box=(-1.0, 1.3, -0.7, 0.9, -1.1, 0.2))
    rng = np.random.default_rng(21)
    h = np.array([2.3 / 7, 1.6 / 6, 1.3 / 5])
    inner = m.code[:, 0] == 0
    m.coord[inner] += 0.15 * h * rng.uniform(-1.0, 1.0, (int(inner.sum()), 3))
    k = np.exp(rng.standard_normal(m.nb_elements))
    check_plan(ctx, m, O.LAPLACE3DT4, f=rng.standard_normal(m.nb_nodes), coef=k,
               bc=rng.standard_normal(m.nb_nodes))


def _oriented(coord, elem):
    """Swap the last two nodes of every element with a negative Jacobian."""
    e = elem.copy()
    x = coord[e - 1]
    if e.shape[1] == 3:
        det = (x[:, 1, 0] - x[:, 0, 0]) * (x[:, 2, 1] - x[:, 0, 1]) - \
              (x[:, 1, 1] - x[:, 0, 1]) * (x[:, 2, 0] - x[:, 0, 0])
    else:
        det = np.linalg.det(x[:, 1:, :] - x[:, :1, :])
    neg = det < 0
    e[neg, -2], e[neg, -1] = elem[neg, -1], elem[neg, -2]
    return e


def _fan2d(centre, n, r=1.0):
    """n triangles around a centre node: the centre row holds n + 1 nonzeros."""
    t = 2 * np.pi * (np.arange(n) + 0.3) / n
    coord = np.vstack([centre, np.c_[centre[0] + r * np.cos(t), centre[1] + r * np.sin(t)]])
    elem = np.array([[1, 2 + k, 2 + (k + 1) % n] for k in range(n)], np.int64)
    return coord, elem


def _double_cone(centre, n, r=1.0, hgt=0.8):
    """2n tetrahedra around a centre node between a ring of n nodes and two
    apexes: the centre row holds n + 3 nonzeros, each apex row n + 2."""
    t = 2 * np.pi * (np.arange(n) + 0.2) / n
    ring = np.c_[centre[0] + r * np.cos(t), centre[1] + r * np.sin(t), np.full(n, centre[2])]
    coord = np.vstack([centre, centre + [0, 0, hgt], centre - [0, 0, 0.9 * hgt], ring])
    elem = []
    for k in range(n):
        a, b = 4 + k, 4 + (k + 1) % n
        elem += [[1, a, b, 2], [1, b, a, 3]]
    return coord, np.array(elem, np.int64)


def _merge(parts, dim):
    coords, elems, off = [], [], 0
    for c, e in parts:
        coords.append(c)
        elems.append(e + off)
        off += c.shape[0]
    return np.vstack(coords).reshape(-1, dim), np.vstack(elems)


@pytest.mark.parametrize("kind", ["t3", "t4"])
def test_long_rows(ctx, kind):
    """Rows longer than the lanes a row gets (8 for triangles, 16 for
    tetrahedra) and longer than twice that, which the assembly sums in place:
    fans of 5, 12 and 40 triangles, double cones of 6, 20 and 40 tetrahedra.
    One ring node of each part is fixed; matrix and right-hand side bitwise."""
    rng = np.random.default_rng(22)
    if kind == "t3":
        parts = [_fan2d(np.array([3.0 * i, 0.1 * i]), n, 1.0 + 0.1 * i)
                 for i, n in enumerate((5, 12, 40))]
        coord, elem = _merge(parts, 2)
        okind, coef = O.DC2DT3, np.exp(rng.standard_normal(elem.shape[0]))
    else:
        parts = [_double_cone(np.array([3.0 * i, 0.2 * i, -0.1 * i]), n, 1.0 + 0.1 * i)
                 for i, n in enumerate((6, 20, 40))]
        coord, elem = _merge(parts, 3)
        okind, coef = O.LAPLACE3DT4, np.exp(rng.standard_normal(elem.shape[0]))
    elem = _oriented(coord, elem)
    code = np.zeros((coord.shape[0], 1), np.int32)
    code[elem[::7, -1] - 1, 0] = 1
    m = O.Mesh(coord, code, elem, 1)
    P, a, b, dof = check_plan(ctx, m, okind, f=rng.standard_normal(m.nb_nodes), coef=coef,
                              bc=rng.standard_normal(m.nb_nodes))
    lens = np.diff(P.A.pattern()[0])
    lanes = 8 if kind == "t3" else 16
    assert lens.max() > 2 * lanes and ((lens > lanes) & (lens <= 2 * lanes)).any()


@pytest.mark.parametrize("mode", [0, 1])
def test_elas2dq4_cantilever(ctx, mode):
    m = O.mesh2d(12, 4, quad=True, nb_dof=2, codes=(1, 0, 0, 0), box=(0.0, 4.0, 0.0, 1.0))
    m.code[:, 1] = m.code[:, 0]  # clamp both DOFs on x = 0
    rng = np.random.default_rng(3)
    E = np.exp(0.5 * rng.standard_normal(m.nb_elements))
    P, a, b, dof = check_plan(ctx, m, O.ELAS2DQ4, coef=E, prm=[0.3, mode],
                              bc=np.zeros(2 * m.nb_nodes))
    # right-edge tractions, side order
    ny = 4
    nx = 12
    sides = np.array([[nx * (ny + 1) + j + 1, nx * (ny + 1) + j + 2] for j in range(ny)], np.int32)
    tr = rng.standard_normal((ny, 2))
    gb = dev(b)
    P.traction(sides, dev(tr), gb)
    ref = O.traction_rhs(m, sides.astype(np.int64), tr, dof, b)
    assert np.array_equal(host(gb), ref)


@pytest.mark.parametrize("mode", [0, 1])
def test_elas2dq4_body_force(ctx, mode):
    """Body force (fx, fy per node) at the 2x2 Gauss points, added before the
    Dirichlet lifting (Equa_Solid.h:184-190), with nonzero imposed
    displacements: right-hand side bitwise."""
    m = O.mesh2d(10, 6, quad=True, nb_dof=2, codes=(1, 0, 2, 0), box=(0.0, 3.0, 0.0, 1.0))
    m.code[:, 1] = m.code[:, 0]
    rng = np.random.default_rng(9)
    E = np.exp(0.5 * rng.standard_normal(m.nb_elements))
    bf = rng.standard_normal(2 * m.nb_nodes)
    bc = 1e-2 * rng.standard_normal(2 * m.nb_nodes)
    check_plan(ctx, m, O.ELAS2DQ4, f=bf, coef=E, prm=[0.3, mode], bc=bc)


def test_elas2dq4_tractions_on_two_edges(ctx):
    """Sides listed out of geometric order over the right and top edges: the
    corner node receives contributions from sides far apart in the list;
    every entry of b adds them in side order, bitwise the oracle's."""
    import ofeli_amd as oa
    nx, ny = 9, 5
    m = O.mesh2d(nx, ny, quad=True, nb_dof=2, codes=(1, 0, 0, 0), box=(0.0, 2.0, 0.0, 1.0))
    m.code[:, 1] = m.code[:, 0]
    rng = np.random.default_rng(10)
    E = np.exp(0.3 * rng.standard_normal(m.nb_elements))
    P, a, b, dof = check_plan(ctx, m, O.ELAS2DQ4, coef=E, prm=[0.3, 1])
    right = [[nx * (ny + 1) + j + 1, nx * (ny + 1) + j + 2] for j in range(ny)]
    top = [[i * (ny + 1) + ny + 1, (i + 1) * (ny + 1) + ny + 1] for i in range(nx)]
    sides = np.array(right[::-1] + top[::2] + top[1::2], np.int32)
    tr = rng.standard_normal((sides.shape[0], 2))
    gb = dev(b)
    P.traction(sides, dev(tr), gb)
    ref = O.traction_rhs(m, sides.astype(np.int64), tr, dof, b)
    assert np.array_equal(host(gb), ref)
    # a second call with fewer sides reuses the plan's buffers
    gb2 = dev(b)
    P.traction(sides[:3], dev(tr[:3]), gb2)
    assert np.array_equal(host(gb2), O.traction_rhs(m, sides[:3].astype(np.int64), tr[:3], dof, b))


def test_insert_bc(ctx):
    m = O.mesh2d(7, 5)
    import ofeli_amd as oa
    P = oa.AssemblyPlan(ctx, m.coord, m.elem.astype(np.int32), m.code)
    v = np.random.default_rng(0).standard_normal(P.nb_eq)
    bc = np.random.default_rng(1).standard_normal(m.nb_nodes)
    u = _t().zeros(m.nb_nodes, dtype=_t().float64, device="cuda")
    P.insert_bc(dev(v), dev(bc), u)
    assert np.array_equal(host(u), O.insert_bc(m.code, v, bc))


def test_negative_jacobian_rejected(ctx):
    import ofeli_amd as oa
    m = O.mesh2d(3, 3)
    elem = m.elem.copy()
    elem[4] = elem[4][[0, 2, 1]]  # flip one triangle
    P = oa.AssemblyPlan(ctx, m.coord, elem.astype(np.int32), m.code)
    with pytest.raises(oa.OfhError) as e:
        P.build(oa.LAPLACE2DT3, matrix=True)
    assert e.value.code == -4
