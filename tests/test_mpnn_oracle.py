"""Pins of the CPU forward/backward of the learned initial guess (CPU only).

The network has no counterpart in the reference, so its oracle is pinned by
finite differences of its own loss (float64), by the exact bfloat16 rounding
of known values, and by properties the definition fixes: an untrained net
(zero decoder) returns the warm start with loss 1, and the guess is
equivariant under scaling of the right-hand side and the iterates.
"""
import numpy as np
import pytest

from oracle import mpnn as N
from oracle import oracle as O


def small_system(nx=5, seed=3, transient=True):
    m = O.mesh2d(nx, nx)
    dof, neq = O.number_equations(m.code)
    rp, ci = O.node_graph(m.elem, 1, dof, neq)
    rng = np.random.default_rng(seed)
    k = np.exp(0.5 * rng.standard_normal(m.elem.shape[0]))
    f = 1.0 + rng.random(m.coord.shape[0])
    a, b = O.assemble(O.DC2DT3, m, dof, neq, rp, ci, f=f, coef=k)
    if transient:  # add a lumped mass / dt so the system is a BE step
        a = a.copy()
        for i in range(neq):
            for p in range(rp[i], rp[i + 1]):
                if ci[p] - 1 == i:
                    a[p] += 0.5
    g = N.Graph(rp, ci - 1, a)
    xs = np.linalg.solve(O.to_dense(rp, ci, a), b)
    x_n = xs * (0.8 + 0.1 * rng.random(neq))
    x_p = x_n * 0.9
    return g, b, xs, x_n, x_p


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -7, -2.5, 0.0,
                  1.0 + 2.0 ** -8 + 2.0 ** -20, 3.0 * 2.0 ** -130])
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -7, -2.5, 0.0, 1.0 + 2.0 ** -7,
                     3.0 * 2.0 ** -130])
    assert np.array_equal(N.bf16(x), want)


def test_parameter_layout():
    L = 4
    o = N.offsets(L)
    assert N.n_params(L) == 8 * 64 + 64 + L * (64 * 128 + 64 + 64 + 128 * 64 + 64) + 65
    assert o["dec_b"] == N.n_params(L) - 1
    th = np.arange(N.n_params(L), dtype=np.float64)
    assert np.array_equal(N.pack(N.unpack(th, L), L), th)


def test_untrained_decoder_returns_warm_start():
    g, b, xs, x_n, x_p = small_system()
    th = N.init_params(2, seed=1)
    P = N.unpack(th, 2)
    P["dec_w"][:] = 0.0
    th = N.pack(dict(P, dec_b=0.0), 2)
    x0, cache = N.forward(g, th, 2, x_n, x_p, b)
    assert np.array_equal(x0, x_n)
    L, gd = N.loss(g, cache, xs)
    assert L == 1.0 and not gd.any()


@pytest.mark.parametrize("ext", [0, 1])
def test_loss_is_the_a_norm_error_ratio_after_the_line_search(ext):
    """With the exact solution as x*, L equals |x0 - x*|_A^2 / |x_b - x*|_A^2,
    and the line search never makes the base worse."""
    g, b, xs, x_n, x_p = small_system()
    A = g.A.toarray()
    for seed in range(3):
        th = N.init_params(2, seed=seed)
        x0, cache = N.forward(g, th, 2, x_n, x_p, b, ext=ext)
        L, _ = N.loss(g, cache, xs)
        xb = N.base(x_n, x_p, ext)
        e0, eb = x0 - xs, xb - xs
        ratio = (e0 @ A @ e0) / (eb @ A @ eb)
        assert abs(L - ratio) <= 1e-9 and 0 <= L <= 1


def test_scale_equivariance():
    g, b, xs, x_n, x_p = small_system()
    th = N.init_params(3, seed=2)
    x0, _ = N.forward(g, th, 3, x_n, x_p, b)
    x0c, _ = N.forward(g, th, 3, 8.0 * x_n, 8.0 * x_p, 8.0 * b)
    assert np.allclose(x0c, 8.0 * x0, rtol=1e-12, atol=0)


def test_features_definition():
    g, b, xs, x_n, x_p = small_system()
    F, s = N.features(g, x_n, x_p, b)
    A = g.A.toarray()
    r = b - A @ x_n
    z = r / np.diag(A)
    assert np.isclose(s, np.sqrt(np.mean(z * z)))
    assert np.allclose(F[:, 0], z / s, rtol=1e-6)
    assert np.allclose(F[:, 1], (r - A @ z) / np.diag(A) / s, rtol=1e-6, atol=1e-7)
    assert np.allclose(F[:, 2], (x_n - x_p) / s, rtol=1e-6)
    assert abs(F[:, 3].mean()) < 1e-6 and not F[:, 4:].any()


@pytest.mark.parametrize("layers,ext", [(1, 0), (2, 0), (2, 1)])
def test_gradient_matches_finite_differences(layers, ext):
    g, b, xs, x_n, x_p = small_system()
    th = N.init_params(layers, seed=5)
    L, grad, _ = N.loss_and_grad(g, th, layers, x_n, x_p, b, xs, ext=ext)
    assert 0 < L

    def f(t):
        _, cache = N.forward(g, t, layers, x_n, x_p, b, ext=ext)
        return N.loss(g, cache, xs)[0]

    rng = np.random.default_rng(0)
    h = 1e-6
    # directional derivatives along random directions
    for _ in range(4):
        v = rng.standard_normal(th.size)
        fd = (f(th + h * v) - f(th - h * v)) / (2 * h)
        assert abs(fd - grad @ v) <= 1e-6 * max(1.0, abs(fd)), (fd, grad @ v)
    # and single coordinates of every parameter group
    o = N.offsets(layers)
    idx = [o[k] + 3 for k in o if k != "dec_b"] + [o["dec_b"]]
    for i in idx:
        e = np.zeros_like(th)
        e[i] = h
        fd = (f(th + e) - f(th - e)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1e-2, abs(fd)), (i, fd, grad[i])


def test_bf16_forward_close_to_fp64():
    g, b, xs, x_n, x_p = small_system(nx=8)
    th = N.init_params(4, seed=7)
    x0, _ = N.forward(g, th, 4, x_n, x_p, b)
    x0b, _ = N.forward(g, th, 4, x_n, x_p, b, bf=True)
    d = x0 - x_n
    assert np.linalg.norm(x0b - x0) <= 0.05 * np.linalg.norm(d)
    assert not np.array_equal(x0b, x0)


@pytest.mark.parametrize("ext", [0, 1])
def test_residual_loss_is_the_residual_ratio_after_the_line_search(ext):
    """Minimum-residual objective: L equals |b - A x0|^2 / |b - A x_b|^2, and
    the step minimises |b - A (x_b + alpha d)| over alpha (checked against a
    dense scan of alpha)."""
    g, b, xs, x_n, x_p = small_system()
    A = g.A.toarray()
    for seed in range(3):
        th = N.init_params(2, seed=seed)
        x0, cache = N.forward(g, th, 2, x_n, x_p, b, ext=ext, obj=1)
        L, _ = N.loss_residual(g, cache)
        xb = N.base(x_n, x_p, ext)
        r0, rb = b - A @ x0, b - A @ xb
        assert abs(L - (r0 @ r0) / (rb @ rb)) <= 1e-9 and 0 <= L <= 1
        d = cache["d"]
        alpha = (x0 - xb) @ d / (d @ d)
        scan = [np.linalg.norm(b - A @ (xb + a * d)) for a in alpha * np.linspace(0.9, 1.1, 21)]
        assert np.linalg.norm(r0) <= min(scan) * (1 + 1e-12)


@pytest.mark.parametrize("layers,ext", [(1, 0), (2, 1)])
def test_residual_gradient_matches_finite_differences(layers, ext):
    g, b, xs, x_n, x_p = small_system()
    th = N.init_params(layers, seed=6)
    L, grad, _ = N.loss_and_grad(g, th, layers, x_n, x_p, b, None, ext=ext, obj=1)
    assert 0 < L < 1

    def f(t):
        _, cache = N.forward(g, t, layers, x_n, x_p, b, ext=ext, obj=1)
        return N.loss_residual(g, cache)[0]

    rng = np.random.default_rng(1)
    h = 1e-6
    for _ in range(4):
        v = rng.standard_normal(th.size)
        fd = (f(th + h * v) - f(th - h * v)) / (2 * h)
        assert abs(fd - grad @ v) <= 1e-6 * max(1.0, abs(fd)), (fd, grad @ v)
    o = N.offsets(layers)
    for i in [o[k] + 3 for k in o if k != "dec_b"] + [o["dec_b"]]:
        e = np.zeros_like(th)
        e[i] = h
        fd = (f(th + e) - f(th - e)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1e-2, abs(fd)), (i, fd, grad[i])
