"""DILU_PREC / ILU_PREC of the oracle (SpMatrix::DILUFactorize / DILUSolve,
SpMatrix_impl.h:834-900; the reference's ILU routines call the DILU ones)
pinned against closed forms and an independent dense restatement."""
import numpy as np
import pytest

from oracle import oracle as O


def csr(Ad):
    """1-based CSR of a dense matrix's nonzero pattern (sorted columns)."""
    n = Ad.shape[0]
    rp, ci, a = [0], [], []
    for i in range(n):
        for j in range(n):
            if Ad[i, j] != 0.0:
                ci.append(j + 1)
                a.append(Ad[i, j])
        rp.append(len(ci))
    return np.array(rp, np.int64), np.array(ci, np.int64), np.array(a)


def dense_dilu(Ad):
    """D-ILU written as rows pulling from earlier rows: D_l = a_ll - sum over
    i < l with both a_li and a_il in the pattern of a_li a_il / D_i; then
    M = (D + L) D^-1 (D + U)."""
    n = Ad.shape[0]
    D = np.zeros(n)
    for l in range(n):
        D[l] = Ad[l, l]
        for i in range(l):
            if Ad[l, i] != 0.0 and Ad[i, l] != 0.0:
                D[l] -= Ad[l, i] * Ad[i, l] / D[i]
    L = np.tril(Ad, -1)
    U = np.triu(Ad, 1)
    M = (np.diag(D) + L) @ np.diag(1.0 / D) @ (np.diag(D) + U)
    return D, M


def laplace2d(n, seed=0, unsym=False):
    rng = np.random.default_rng(seed)
    N = n * n
    Ad = np.zeros((N, N))
    for y in range(n):
        for x in range(n):
            i = y * n + x
            Ad[i, i] = 4.0 + rng.random()
            for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                if 0 <= x + dx < n and 0 <= y + dy < n:
                    Ad[i, (y + dy) * n + x + dx] = -1.0 - 0.5 * rng.random()
    if unsym:  # drop some upper entries: their lower partners find no transpose
        for i in range(0, N - 1, 3):
            Ad[i, i + 1] = 0.0
    return Ad


def test_tridiagonal_is_exact():
    """On a tridiagonal matrix D-ILU is the exact LDU factorisation: the
    preconditioner solve is A^-1 b and preconditioned CG takes one step."""
    rng = np.random.default_rng(1)
    n = 40
    Ad = np.diag(2.0 + rng.random(n)) + np.diag(-0.5 - 0.3 * rng.random(n - 1), 1)
    Ad = Ad + np.triu(Ad, 1).T
    rp, ci, a = csr(Ad)
    b = rng.standard_normal(n)
    piv = O.dilu_factorize(rp, ci, a)
    x = O.dilu_solve(rp, ci, a, piv, b)
    assert np.linalg.norm(x - np.linalg.solve(Ad, b)) <= 1e-12 * np.linalg.norm(x)
    for prec in (O.DILU, O.ILU):
        ret, xs, _ = O.cg(rp, ci, a, b, prec=prec, max_it=10, toler=1e-12)
        assert ret == 1


def test_diagonal_matrix():
    d = np.array([2.0, 4.0, 0.5, 8.0])
    rp, ci, a = csr(np.diag(d))
    piv = O.dilu_factorize(rp, ci, a)
    assert np.array_equal(piv, 1.0 / d)
    b = np.array([1.0, -2.0, 3.0, 0.25])
    assert np.array_equal(O.dilu_solve(rp, ci, a, piv, b), b / d)


@pytest.mark.parametrize("unsym", [False, True])
def test_matches_dense_restatement(unsym):
    Ad = laplace2d(7, seed=3, unsym=unsym)
    rp, ci, a = csr(Ad)
    D, M = dense_dilu(Ad)
    piv = O.dilu_factorize(rp, ci, a)
    assert np.allclose(piv, 1.0 / D, rtol=1e-14, atol=0)
    b = np.random.default_rng(4).standard_normal(Ad.shape[0])
    x = O.dilu_solve(rp, ci, a, piv, b)
    xd = np.linalg.solve(M, b)
    assert np.linalg.norm(x - xd) <= 1e-12 * np.linalg.norm(xd)


def test_zero_pivot_and_cg_gain():
    Ad = laplace2d(12, seed=5)
    Ad = 0.5 * (Ad + Ad.T)  # CG needs a symmetric matrix
    rp, ci, a = csr(Ad)
    b = np.ones(Ad.shape[0])
    r_diag, _, _ = O.cg(rp, ci, a, b, prec=O.DIAG, max_it=1000, toler=1e-10)
    r_dilu, x, _ = O.cg(rp, ci, a, b, prec=O.DILU, max_it=1000, toler=1e-10)
    assert 0 < r_dilu < r_diag
    assert np.linalg.norm(Ad @ x - b) <= 1e-9 * np.linalg.norm(b)
    a0 = a.copy()
    k = rp[3] + list(ci[rp[3]:rp[4]]).index(4)  # row 4's diagonal entry
    a0[k] = 0.0
    with pytest.raises(RuntimeError):
        O.dilu_factorize(rp, ci, a0)
