"""Write the golden fixtures of tests/golden/*.npz.

The fixtures are produced from textbook finite-element formulas evaluated
with plain numpy dense algebra, independently of the oracle and of the
GPU path:
  * P1 triangle / tetrahedron: K_e = |T| * k * G G^T with G the gradients of
    the barycentric coordinates, from inv([[1, x_i, y_i(, z_i)]]);
  * Q4 elasticity: K_e = sum_gp w det(J) B^T D B with bilinear shape functions
    at the 2x2 Gauss points, D from (lambda, G) of plane strain / stress;
  * global matrices by dense scatter-add, then restricted to the free DOFs
    (DOF numbering: free DOFs in node order, node-major, component-minor);
  * solutions by numpy.linalg.solve.
The reference's own tests hold no numbers (its demos only print errors), so
these fixtures pin the mathematics the reference implements.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def grid2d(nx, ny, quad, box=(0.0, 1.0, 0.0, 1.0)):
    x0, x1, y0, y1 = box
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    coord = np.array([[x, y] for x in xs for y in ys])  # x outer, y inner
    nid = lambda i, j: i * (ny + 1) + j
    elem = []
    for i in range(1, nx + 1):
        for j in range(1, ny + 1):
            a, b, c, d = nid(i - 1, j - 1), nid(i, j - 1), nid(i, j), nid(i - 1, j)
            if quad:
                elem.append([a, b, c, d])
            else:
                elem.append([a, b, c])
                elem.append([c, d, a])
    return coord, np.array(elem)


def grid3d_kuhn(n):
    xs = np.linspace(0.0, 1.0, n + 1)
    coord = np.array([[x, y, z] for z in xs for y in xs for x in xs])
    nid = lambda i, j, k: (k * (n + 1) + j) * (n + 1) + i
    elem = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                c = [nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                     nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1),
                     nid(i, j + 1, k + 1)]
                for t in [(0, 1, 2, 6), (0, 1, 6, 5), (0, 3, 6, 2), (0, 3, 7, 6), (0, 4, 5, 6),
                          (0, 4, 6, 7)]:
                    elem.append([c[q] for q in t])
    return coord, np.array(elem)


def p1_stiffness(xe, k=1.0):
    d = xe.shape[1]
    M = np.hstack([np.ones((d + 1, 1)), xe])
    G = np.linalg.inv(M)[1:, :].T  # (d+1) x d gradients
    meas = abs(np.linalg.det(M)) / (2.0 if d == 2 else 6.0)
    return k * meas * G @ G.T, meas


def q4_stiffness(xe, E, nu, plane_stress):
    lam = nu * E / ((1 + nu) * (1 - 2 * nu))
    G = 0.5 * E / (1 + nu)
    if plane_stress:
        lam = 2 * lam * G / (lam + 2 * G)
    D = np.array([[lam + 2 * G, lam, 0], [lam, lam + 2 * G, 0], [0, 0, G]])
    g = 1.0 / np.sqrt(3.0)
    K = np.zeros((8, 8))
    for s in (-g, g):
        for t in (-g, g):
            dN = 0.25 * np.array([[-(1 - t), (1 - t), (1 + t), -(1 + t)],
                                  [-(1 - s), -(1 + s), (1 + s), (1 - s)]])
            J = dN @ xe
            dX = np.linalg.solve(J, dN)
            B = np.zeros((3, 8))
            B[0, 0::2] = dX[0]
            B[1, 1::2] = dX[1]
            B[2, 0::2] = dX[1]
            B[2, 1::2] = dX[0]
            K += np.linalg.det(J) * B.T @ D @ B
    return K


def q4_deviator(xe, E, nu):
    """The 2*G*eps:eps part alone (plane strain G)."""
    G = 0.5 * E / (1 + nu)
    g = 1.0 / np.sqrt(3.0)
    D = np.array([[2 * G, 0, 0], [0, 2 * G, 0], [0, 0, G]])
    K = np.zeros((8, 8))
    for s in (-g, g):
        for t in (-g, g):
            dN = 0.25 * np.array([[-(1 - t), (1 - t), (1 + t), -(1 + t)],
                                  [-(1 - s), -(1 + s), (1 + s), (1 - s)]])
            J = dN @ xe
            dX = np.linalg.solve(J, dN)
            B = np.zeros((3, 8))
            B[0, 0::2] = dX[0]
            B[1, 1::2] = dX[1]
            B[2, 0::2] = dX[1]
            B[2, 1::2] = dX[0]
            K += np.linalg.det(J) * B.T @ D @ B
    return K


def assemble_dense(coord, elem, nb_dof, kefun, free):
    nn = coord.shape[0]
    K = np.zeros((nn * nb_dof, nn * nb_dof))
    for e, en in enumerate(elem):
        Ke = kefun(e, coord[en])
        idx = np.array([n * nb_dof + c for n in en for c in range(nb_dof)])
        K[np.ix_(idx, idx)] += Ke
    return K[np.ix_(free, free)]


def to_csr(K):
    rp = [0]
    cols = []
    vals = []
    for i in range(K.shape[0]):
        nz = np.nonzero(np.abs(K[i]) > 0)[0]
        cols.extend(nz)
        vals.extend(K[i, nz])
        rp.append(len(cols))
    return np.array(rp), np.array(cols), np.array(vals)


def pattern_from_elements(elem, nb_dof, dof):
    """Boolean pattern: DOFs sharing an element (including structural zeros)."""
    n = int(dof.max())
    P = np.zeros((n, n), bool)
    for en in elem:
        d = np.array([dof[v, c] for v in en for c in range(nb_dof)])
        d = d[d > 0] - 1
        P[np.ix_(d, d)] = True
    return P


def main():
    out = {}
    # 2D P1, 6x5, Dirichlet everywhere, k per element, f nodal
    coord, elem = grid2d(6, 5, False)
    rng = np.random.default_rng(7)
    k = np.exp(0.5 * rng.standard_normal(len(elem)))
    nn = coord.shape[0]
    bnd = (np.isclose(coord[:, 0], 0) | np.isclose(coord[:, 0], 1) | np.isclose(coord[:, 1], 0)
           | np.isclose(coord[:, 1], 1))
    free = np.nonzero(~bnd)[0]
    K = assemble_dense(coord, elem, 1, lambda e, xe: p1_stiffness(xe, k[e])[0], free)
    f = 1.0 + rng.random(nn)
    # lumped source: F_i = sum_e |T|/3 f_i
    F = np.zeros(nn)
    for en in elem:
        _, area = p1_stiffness(coord[en])
        F[en] += area / 3.0 * f[en]
    F = F[free]
    dof = np.zeros((nn, 1), int)
    dof[free, 0] = np.arange(1, len(free) + 1)
    Pt = pattern_from_elements(elem, 1, dof)
    rp, ci = [0], []
    for i in range(len(free)):
        nz = np.nonzero(Pt[i])[0]
        ci.extend(nz)
        rp.append(len(ci))
    out["p1tri_rp"] = np.array(rp)
    out["p1tri_col"] = np.array(ci)
    out["p1tri_K"] = K
    out["p1tri_F"] = F
    out["p1tri_k"] = k
    out["p1tri_f"] = f
    out["p1tri_x"] = np.linalg.solve(K, F)

    # 3D Kuhn P1, 3^3 cells, all boundary nodes Dirichlet 0, k = 1
    coord3, elem3 = grid3d_kuhn(3)
    bnd3 = np.any(np.isclose(coord3, 0) | np.isclose(coord3, 1), axis=1)
    free3 = np.nonzero(~bnd3)[0]
    K3 = assemble_dense(coord3, elem3, 1, lambda e, xe: p1_stiffness(xe)[0], free3)
    out["p1tet_K"] = K3
    vols = np.array([p1_stiffness(coord3[en])[1] for en in elem3])
    out["p1tet_vol"] = vols

    # Q4 cantilever 4x2 cells on [0,4]x[0,1], clamped x=0, E lognormal
    cq, eq = grid2d(4, 2, True, box=(0.0, 4.0, 0.0, 1.0))
    E = np.exp(0.5 * rng.standard_normal(len(eq)))
    clamp = np.isclose(cq[:, 0], 0.0)
    freeq = np.array([n * 2 + c for n in range(cq.shape[0]) for c in range(2) if not clamp[n]])
    for mode, stress in ((0, False), (1, True)):
        if mode == 1:
            Kq = assemble_dense(cq, eq, 2, lambda e, xe: q4_stiffness(xe, E[e], 0.3, True), freeq)
        else:
            # reference parity: plane strain, 2*Dev + Dil = textbook strain + Dev
            Kq = assemble_dense(cq, eq, 2, lambda e, xe: q4_stiffness(xe, E[e], 0.3, False)
                                + q4_deviator(xe, E[e], 0.3), freeq)
        out["q4_K_mode%d" % mode] = Kq
    out["q4_E"] = E
    np.savez_compressed(os.path.join(HERE, "fe_fixtures.npz"), **out)
    print("wrote", os.path.join(HERE, "fe_fixtures.npz"))


if __name__ == "__main__":
    main()
