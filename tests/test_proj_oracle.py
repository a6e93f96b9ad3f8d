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


def test_full_basis_is_rebuilt_and_repeats_are_dropped():
    A, rng = spd(seed=2)
    Ad = A.toarray()
    P = PJ.Projection(A, 8)
    xs = [np.linalg.solve(Ad, rng.standard_normal(A.shape[0])) for _ in range(11)]
    for x in xs:
        assert P.update(x)
    # update 9 found V full and rebuilt it from solutions 6..8 and 9; 10, 11 added
    assert P.count == 6 and P.head == 6
    W = np.array(xs[5:]).T
    b = rng.standard_normal(A.shape[0])
    want = W @ np.linalg.solve(W.T @ Ad @ W, W.T @ b)
    assert np.allclose(P.guess(b), want, atol=1e-10 * np.abs(want).max())
    G = P.V[:6] @ Ad @ P.V[:6].T
    assert np.allclose(G, np.eye(6), atol=1e-10)
    v_before = P.V.copy()
    # nothing new: a solution in the span (or a combination) is dropped
    assert not P.update(xs[-1]) and not P.update(0.3 * xs[-2] - 2.0 * xs[-4])
    assert np.array_equal(P.V, v_before) and P.count == 6 and len(P.X) == 8


def test_minres_guess_is_least_squares_over_stored_solutions():
    A, rng = spd(seed=3)
    Ad = A.toarray()
    P = PJ.Projection(A, 8, minres=True)
    xs = []
    for k in range(6):
        x = np.linalg.solve(Ad, rng.standard_normal(A.shape[0]))
        P.update(x)
        xs.append(x)
    W = np.array(xs).T
    b = rng.standard_normal(A.shape[0])
    y, *_ = np.linalg.lstsq(Ad @ W, b, rcond=None)
    want = W @ y
    assert np.allclose(P.guess(b), want, atol=1e-10 * np.abs(want).max())
    AV = P.AV[:6]
    assert np.allclose(AV @ AV.T, np.eye(6), atol=1e-12)
    # its residual is no larger than that of the last solution
    assert np.linalg.norm(b - Ad @ P.guess(b)) <= np.linalg.norm(b - Ad @ xs[-1]) * (1 + 1e-12)
