"""Learned initial guess (ofh_mpnn_*) against the CPU oracle oracle/mpnn.py.

The oracle runs float64 with the operands of the dense products rounded to
bfloat16 exactly as the GPU feeds its matrix cores, so the GPU differs only
by FP32 accumulation order (and the rare ReLU or rounding flips it causes):
tolerances below are relative L2 norms, stated per check.
"""
import numpy as np
import pytest

from oracle import mpnn as N

pytestmark = pytest.mark.gpu


def _t():
    import torch
    return torch


def be_system(ctx, nx, seed=0, dt=1e-3):
    """Backward-Euler heat matrix (DC2DT3 + lumped capacity / dt) on an
    nx x nx grid, with iterates x_n, x_p near the solution x*."""
    import ofeli_amd as oa
    torch = _t()
    coord, code, elem = oa.api.mesh2d(nx, nx)
    rng = np.random.default_rng(seed)
    P = oa.AssemblyPlan(ctx, coord, elem, code)
    k = torch.as_tensor(np.exp(0.5 * rng.standard_normal(elem.shape[0])), device="cuda")
    xy = coord
    f = torch.as_tensor(np.exp(-((xy[:, 0] - 0.4) ** 2 + (xy[:, 1] - 0.6) ** 2) / 0.02), device="cuda")
    u = torch.as_tensor(0.1 * np.sin(np.pi * xy[:, 0]) * np.sin(np.pi * xy[:, 1]), device="cuda")
    b = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
    P.build(oa.DC2DT3_BE, coef=k, f=f, u_prev=u, prm=(dt, 1.0, 1.0), matrix=True, b=b)
    xs = torch.zeros_like(b)
    info = oa.cg(P.A, b, xs, max_it=5000, toler=1e-13)
    assert info.nb_it > 0
    pert = torch.as_tensor(1.0 + 0.2 * rng.standard_normal(P.nb_eq), device="cuda")
    x_n = xs * pert
    x_p = x_n * 0.95
    rp, col = P.A.pattern()
    g = N.Graph(rp, col, P.A.values())
    return P, g, b, xs, x_n, x_p


def host(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_untrained_network_returns_warm_start(ctx):
    import ofeli_amd as oa
    P, g, b, xs, x_n, x_p = be_system(ctx, 12)
    M = oa.LearnedGuess(P.A, layers=4, seed=1)
    assert M.nparams == N.n_params(4)
    x0 = _t().empty_like(b)
    M.guess(x_n, x_p, b, x0)
    assert _t().equal(x0, x_n)
    L, _ = M.loss_grad(x_n, x_p, b, xs, x0, want_grad=False)
    assert abs(L - 1.0) < 1e-12


@pytest.mark.parametrize("nx,layers", [(12, 4), (41, 2), (150, 4)])
def test_forward_matches_oracle(ctx, nx, layers):
    import ofeli_amd as oa
    P, g, b, xs, x_n, x_p = be_system(ctx, nx, seed=nx)
    M = oa.LearnedGuess(P.A, layers=layers)
    th = N.init_params(layers, seed=nx).astype(np.float32)
    M.set_params(th)
    assert np.array_equal(M.params(), th)
    x0 = _t().empty_like(b)
    M.guess(x_n, x_p, b, x0)
    xo, _ = N.forward(g, th.astype(np.float64), layers, host(x_n), host(x_p), host(b), bf=True)
    d_gpu = host(x0) - host(x_n)
    d_orc = xo - host(x_n)
    assert np.linalg.norm(d_orc) > 0
    assert rel(d_gpu, d_orc) < 5e-3, rel(d_gpu, d_orc)


@pytest.mark.parametrize("nx,layers", [(12, 2), (41, 4)])
def test_loss_and_gradient_match_oracle(ctx, nx, layers):
    import ofeli_amd as oa
    P, g, b, xs, x_n, x_p = be_system(ctx, nx, seed=7 + nx)
    M = oa.LearnedGuess(P.A, layers=layers)
    th = N.init_params(layers, seed=3).astype(np.float32)
    M.set_params(th)
    x0 = _t().empty_like(b)
    L, grad = M.loss_grad(x_n, x_p, b, xs, x0)
    Lo, go, _ = N.loss_and_grad(g, th.astype(np.float64), layers, host(x_n), host(x_p), host(b),
                                host(xs), bf=True)
    assert abs(L - Lo) <= 2e-3 * abs(Lo), (L, Lo)
    o = N.offsets(layers)
    keys = sorted(o, key=o.get) + ["end"]
    bounds = [o[k] for k in keys[:-1]] + [N.n_params(layers)]
    for k, a, e in zip(keys[:-1], bounds[:-1], bounds[1:]):
        err = rel(grad[a:e], go[a:e])
        assert err < 2e-2, (k, err)
    assert rel(grad, go) < 1e-2


def test_graph_replay_equals_direct_launch(ctx):
    import ofeli_amd as oa
    torch = _t()
    P, g, b, xs, x_n, x_p = be_system(ctx, 30, seed=4)
    M = oa.LearnedGuess(P.A, layers=4)
    M.set_params(N.init_params(4, seed=9).astype(np.float32))
    M.use_graph(False)
    ref = torch.empty_like(b)
    M.guess(x_n, x_p, b, ref)
    M.use_graph(True)
    for rep in range(4):
        x0 = torch.full_like(b, float("nan"))
        M.guess(x_n, x_p, b, x0)
        torch.cuda.synchronize()
        assert torch.equal(x0, ref), rep
        y = torch.zeros_like(b)
        oa.spmv(P.A, x0, y)  # eager work between replays
        torch.cuda.synchronize()


def test_training_lowers_the_loss(ctx):
    import ofeli_amd as oa
    torch = _t()
    P, g, b, xs, x_n, x_p = be_system(ctx, 24, seed=5)
    M = oa.LearnedGuess(P.A, layers=4, seed=2)
    x0 = torch.empty_like(b)
    losses = [M.train_step(x_n, x_p, b, xs, x0, lr=2e-3) for _ in range(60)]
    assert abs(losses[0] - 1.0) < 1e-12  # the untrained net returns the warm start
    assert min(losses[-5:]) < 0.9 * losses[0], losses[::10]
    L, _ = M.loss_grad(x_n, x_p, b, xs, x0, want_grad=False)
    assert L < 0.9


def test_rejects_unsymmetric_matrix(ctx):
    import ofeli_amd as oa
    rp = np.array([0, 2, 4])
    col = np.array([0, 1, 0, 1])
    A = oa.CsrMatrix(ctx, 2, rp, col, np.array([2.0, -1.0, -0.5, 2.0]))
    with pytest.raises(oa.OfhError):
        oa.LearnedGuess(A)
