"""Pins of the previous-solutions projection oracle (CPU only): the guess is
the Galerkin (A-norm-optimal) vector of the span of the stored solutions,
computed here independently by a dense solve of the normal equations."""
import numpy as np
import scipy.sparse as sp

from oracle import proj as PJ


def spd(n=60, seed=0):
    rng = np.random.default_rng(seed)
    B = sp.random(n, n, density=0.1, random_state=seed)
    A = (B @ B.T + sp.identity(n) * 2.0).tocsr()
    return A, rng


def test_guess_is_galerkin_over_stored_solutions():
    A, rng = spd()
    Ad = A.toarray()
    P = PJ.Projection(A, 8)
    xs = []
    for k in range(5):
        x = np.linalg.solve(Ad, rng.standard_normal(A.shape[0]))
        P.update(x)
        xs.append(x)
    W = np.array(xs).T
    b = rng.standard_normal(A.shape[0])
    want = W @ np.linalg.solve(W.T @ Ad @ W, W.T @ b)
    assert np.allclose(P.guess(b), want, rtol=0, atol=1e-10 * np.abs(want).max())
    # the basis is A-orthonormal
    G = P.V[:5] @ Ad @ P.V[:5].T
    assert np.allclose(G, np.eye(5), atol=1e-12)
    # a stored solution is reproduced from its right-hand side
    assert np.allclose(P.guess(Ad @ xs[2]), xs[2], atol=1e-10 * np.abs(xs[2]).max())


def test_guess_is_never_worse_than_the_last_solution_in_a_norm():
    A, rng = spd(seed=1)
    Ad = A.toarray()
    P = PJ.Projection(A, 8)
    x = np.zeros(A.shape[0])
    for k in range(12):
        b = np.sin(0.3 * k + np.arange(A.shape[0]) * 0.1) + 0.05 * rng.standard_normal(A.shape[0])
        xs = np.linalg.solve(Ad, b)
        if k > 0:
            g = P.guess(b)
            e_g, e_w = g - xs, x - xs
            assert e_g @ Ad @ e_g <= e_w @ Ad @ e_w * (1 + 1e-12)
        P.update(xs)
        x = xs


def test_full_basis_is_rebuilt_from_the_last_half_and_repeats_store_zero():
    A, rng = spd(seed=2)
    Ad = A.toarray()
    P = PJ.Projection(A, 8)
    xs = [np.linalg.solve(Ad, rng.standard_normal(A.shape[0])) for _ in range(11)]
    for x in xs:
        P.update(x)
    # update 9 found V full and rebuilt it from solutions 6..9; 10, 11 added
    assert P.count == 6 and P.head == 6
    W = np.array(xs[5:]).T
    b = rng.standard_normal(A.shape[0])
    want = W @ np.linalg.solve(W.T @ Ad @ W, W.T @ b)
    assert np.allclose(P.guess(b), want, atol=1e-10 * np.abs(want).max())
    G = P.V[:6] @ Ad @ P.V[:6].T
    assert np.allclose(G, np.eye(6), atol=1e-10)
    v_before = P.V.copy()
    P.update(xs[-1])  # nothing new: a zero direction fills slot 6
    assert not P.V[6].any() and np.array_equal(P.V[:6], v_before[:6]) and P.count == 7
