"""CGS_SOLVER and BICG_SOLVER of the oracle (CGS.h:91-171, BiCG.h:94-170),
the transposed product (SpMatrix::TMult, SpMatrix_impl.h:761-767) and
Prec::TransSolve (Prec_impl.h:259-296), pinned against dense linear algebra,
closed forms and properties a plausible mistake breaks."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O


def csr(Ad):
    S = sp.csr_matrix(Ad)
    S.sort_indices()
    return S.indptr.astype(np.int64), S.indices.astype(np.int64) + 1, S.data.copy()


def unsym(n, seed):
    rng = np.random.default_rng(seed)
    return np.eye(n) * 4 + rng.standard_normal((n, n)) * 0.5


def lap1d(n):
    return np.diag(2 * np.ones(n)) - np.diag(np.ones(n - 1), 1) - np.diag(np.ones(n - 1), -1)


def sparse_unsym(n, seed):
    """A diagonally dominant unsymmetric sparse matrix with an unsymmetric
    pattern (some entries have no transposed partner)."""
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=0.15, random_state=seed).toarray()
    A += np.diag(np.abs(A).sum(axis=1) + 1.0 + rng.random(n))
    return A


def dense_dilu(Ad):
    """M = (D+L) D^-1 (D+U) with D_l = a_ll - sum_{i<l} a_li a_il / D_i."""
    n = Ad.shape[0]
    D = np.diag(Ad).astype(float).copy()
    for l in range(n):
        for i in range(l):
            if Ad[l, i] != 0 and Ad[i, l] != 0:
                D[l] -= Ad[l, i] * Ad[i, l] / D[i]
    L, U = np.tril(Ad, -1), np.triu(Ad, 1)
    return D, (np.diag(D) + L) @ np.diag(1 / D) @ (np.diag(D) + U)


def test_spmv_t_is_the_transposed_product():
    A = sparse_unsym(40, 1)
    rp, ci, a = csr(A)
    x = np.random.default_rng(2).standard_normal(40)
    assert np.allclose(O.spmv_t(rp, ci, a, x), A.T @ x, rtol=1e-14, atol=1e-14)
    # small integers: every partial sum is exact, so the result is bit-exact
    Ai = np.round(A * 3)
    np.fill_diagonal(Ai, np.diag(Ai) + 1)
    rp, ci, a = csr(Ai)
    xi = np.arange(40.0) - 7
    assert np.array_equal(O.spmv_t(rp, ci, a, xi), Ai.T @ xi)


@pytest.mark.parametrize("seed", [3, 4])
def test_trans_solve_ilu_solves_with_m_transposed(seed):
    A = sparse_unsym(30, seed)
    rp, ci, a = csr(A)
    _, M = dense_dilu(A)
    b = np.random.default_rng(seed).standard_normal(30)
    x = O.prec_trans_apply(rp, ci, a, O.ILU, b)
    xd = np.linalg.solve(M.T, b)
    assert np.linalg.norm(x - xd) <= 1e-12 * np.linalg.norm(xd)
    # the forward (non-transposed) solve is a different operator here
    assert np.linalg.norm(O.dilu_solve(rp, ci, a, O.dilu_factorize(rp, ci, a), b) - xd) > \
        1e-6 * np.linalg.norm(xd)


def test_trans_solve_cases():
    """Symmetric A: M^T = M, so the ILU transposed solve equals DILUSolve;
    DIAG scales; IDENT, DILU and SSOR have no case and return b."""
    S = lap1d(15) + np.diag(np.linspace(0.1, 1.0, 15))
    rp, ci, a = csr(S)
    b = np.random.default_rng(5).standard_normal(15)
    piv = O.dilu_factorize(rp, ci, a)
    assert np.allclose(O.prec_trans_apply(rp, ci, a, O.ILU, b),
                       O.dilu_solve(rp, ci, a, piv, b), rtol=1e-13, atol=1e-15)
    assert np.array_equal(O.prec_trans_apply(rp, ci, a, O.DIAG, b), b * (1.0 / np.diag(S)))
    for p in (O.IDENT, O.DILU, O.SSOR):
        assert np.array_equal(O.prec_trans_apply(rp, ci, a, p, b), b)


@pytest.mark.parametrize("prec", [1, 2, 4])
def test_bicg_finite_termination(prec):
    """BiCG on an unsymmetric n x n system reaches the solution in about n
    iterations (exact arithmetic: at most n); using A instead of A^T for the
    shadow update loses this and does not converge in 200 iterations."""
    A = unsym(12, 0)
    rp, ci, a = csr(A)
    b = np.ones(12)
    ret, x, res = O.bicg(rp, ci, a, b, prec=prec, max_it=200, toler=1e-10)
    xs = np.linalg.solve(A, b)
    assert 0 < ret <= 14
    assert np.linalg.norm(x - xs) <= 1e-9 * np.linalg.norm(xs)
    assert res < 1e-10


@pytest.mark.parametrize("prec", [1, 2])
def test_bicg_equals_cg_on_spd(prec):
    """With A = A^T and a symmetric preconditioner the shadow sequence equals
    the primal one, so BiCG's iterates are CG's."""
    n = 60
    A = lap1d(n) + np.diag(np.linspace(0.01, 0.5, n))
    rp, ci, a = csr(A)
    b = np.random.default_rng(6).standard_normal(n)
    rb, xb, _ = O.bicg(rp, ci, a, b, prec=prec, max_it=500, toler=1e-10)
    rc, xc, _ = O.cg(rp, ci, a, b, prec=prec, max_it=500, toler=1e-10)
    assert abs(rb - rc) <= 1
    assert np.linalg.norm(xb - xc) <= 1e-8 * np.linalg.norm(xc)


def test_bicg_return_conventions():
    A = unsym(8, 1)
    rp, ci, a = csr(A)
    b = np.ones(8)
    xs = np.linalg.solve(A, b)
    assert O.bicg(rp, ci, a, b, x0=xs, max_it=10, toler=1e-8)[0] == 0
    assert O.bicg(rp, ci, a, b, max_it=3, toler=1e-30)[0] == -4  # -(max_it+1)
    assert O.bicg(rp, ci, a, b, max_it=0, toler=1e-30)[0] == -1


def test_cgs_one_by_one_closed_form():
    """1 x 1: y = b, v = a b, alpha = 1/a, w = b, x = b/a, r = 0 after one
    iteration (a = 2, b = 3: every step exact)."""
    rp, ci, a = np.array([0, 1], np.int64), np.array([1], np.int64), np.array([2.0])
    for prec in (O.IDENT, O.DIAG):
        ret, x, res = O.cgs(rp, ci, a, np.array([3.0]), prec=prec, max_it=10, toler=1e-12)
        assert ret == 1 and x[0] == 1.5 and res == 0.0


@pytest.mark.parametrize("prec", [1, 2, 5])
def test_cgs_residual_recurrence_matches_true_residual(prec):
    """x += alpha w and r -= alpha A w keep r = b - A x: the reported
    (recursive) residual equals the true one after a few iterations."""
    A = sparse_unsym(30, 7)
    rp, ci, a = csr(A)
    b = np.random.default_rng(8).standard_normal(30)
    ret, x, res = O.cgs(rp, ci, a, b, prec=prec, max_it=6, toler=1e-300)
    assert ret == -6  # CGS returns -max_it, not -(max_it+1)
    true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert abs(res - true) <= 1e-10 * true + 1e-15


def test_cgs_keeps_q_zero():
    """The reference's q is never assigned (CGS.h:127-135): its CGS does not
    converge on the 20-point 1D Laplacian in 200 iterations, where textbook
    CGS (q = u - alpha v) takes 10 -- a restatement that assigned q would
    converge here."""
    A = lap1d(20)
    rp, ci, a = csr(A)
    b = np.ones(20)
    ret, x, _ = O.cgs(rp, ci, a, b, prec=O.IDENT, max_it=200, toler=1e-10)
    xs = np.linalg.solve(A, b)
    assert ret == -200
    assert np.linalg.norm(x - xs) > 0.1 * np.linalg.norm(xs)
    assert O.cgs(rp, ci, a, b, x0=xs, max_it=5, toler=1e-8)[0] == 0
