"""Learned initial guess (ofh_mpnn_*) against the CPU oracle oracle/mpnn.py.

The oracle runs float64 with the operands of the dense products rounded to
bfloat16 exactly as the GPU feeds its matrix cores, so the GPU differs only
by FP32 accumulation order (and the rare ReLU or rounding flips it causes):
tolerances below are relative L2 norms, stated per check.
"""
import numpy as np
import pytest

from oracle import mpnn as N
from tests.gpu_systems import be_system

pytestmark = pytest.mark.gpu


def _t():
    import torch
    return torch


def host(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("ext", [0, 1])
def test_guess_is_the_line_search_from_its_base(ctx, ext):
    """A zero decoder returns the base exactly (loss 1); a random network's
    guess is never worse than the base in the A-norm."""
    import ofeli_amd as oa
    torch = _t()
    P, g, b, xs, x_n, x_p = be_system(ctx, 12)
    M = oa.LearnedGuess(P.A, layers=4, seed=1, extrapolate=ext)
    assert M.nparams == N.n_params(4)
    xb = x_n + (x_n - x_p) if ext else x_n
    x0 = torch.empty_like(b)
    th = M.params()
    o = N.offsets(4)
    th0 = th.copy()
    th0[o["dec_w"]:] = 0.0
    M.set_params(th0)
    M.guess(x_n, x_p, b, x0)
    assert torch.equal(x0, xb)
    L, _ = M.loss_grad(x_n, x_p, b, xs, x0, want_grad=False)
    assert L == 1.0
    M.set_params(th)
    M.guess(x_n, x_p, b, x0)
    y = torch.empty_like(b)

    def anorm2(e):
        oa.spmv(P.A, e, y)
        return float(torch.dot(e, y))

    assert anorm2(x0 - xs) <= anorm2(xb - xs) * (1 + 1e-12)
    L, _ = M.loss_grad(x_n, x_p, b, xs, x0, want_grad=False)
    assert 0.0 <= L <= 1.0
    assert abs(L - anorm2(x0 - xs) / anorm2(xb - xs)) < 1e-6
    # residual objective: never a larger residual than the base's, and the
    # loss is the squared residual ratio
    M = oa.LearnedGuess(P.A, layers=4, seed=1, extrapolate=ext, objective="residual")

    def res2(x):
        oa.spmv(P.A, x, y)
        return float(torch.sum((b - y) ** 2))

    M.guess(x_n, x_p, b, x0)
    assert res2(x0) <= res2(xb) * (1 + 1e-12)
    L, _ = M.loss_grad(x_n, x_p, b, None, x0, want_grad=False)
    assert abs(L - res2(x0) / res2(xb)) < 1e-6


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


@pytest.mark.parametrize("nx,layers,ext,obj", [(12, 2, 0, 0), (41, 4, 0, 0), (33, 3, 1, 0),
                                                (12, 2, 0, 1), (41, 4, 1, 1)])
def test_loss_and_gradient_match_oracle(ctx, nx, layers, ext, obj):
    import ofeli_amd as oa
    P, g, b, xs, x_n, x_p = be_system(ctx, nx, seed=7 + nx)
    M = oa.LearnedGuess(P.A, layers=layers, extrapolate=ext,
                        objective="residual" if obj else "energy")
    th = N.init_params(layers, seed=3).astype(np.float32)
    M.set_params(th)
    x0 = _t().empty_like(b)
    L, grad = M.loss_grad(x_n, x_p, b, None if obj else xs, x0)
    Lo, go, xo = N.loss_and_grad(g, th.astype(np.float64), layers, host(x_n), host(x_p), host(b),
                                 host(xs), bf=True, ext=ext, obj=obj)
    xb = N.base(host(x_n), host(x_p), ext)
    assert rel(host(x0) - xb, xo - xb) < 5e-3
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
    assert 0.0 < losses[0] <= 1.0
    assert min(losses[-5:]) < 0.9 * losses[0], losses[::10]
    L, _ = M.loss_grad(x_n, x_p, b, xs, x0, want_grad=False)
    assert L < 0.9 * losses[0]


def test_rejects_unsymmetric_matrix(ctx):
    import ofeli_amd as oa
    rp = np.array([0, 2, 4])
    col = np.array([0, 1, 0, 1])
    A = oa.CsrMatrix(ctx, 2, rp, col, np.array([2.0, -1.0, -0.5, 2.0]))
    with pytest.raises(oa.OfhError):
        oa.LearnedGuess(A)


def test_trained_network_beats_its_base_on_held_out_steps(ctx):
    """What training buys: on a 64^2 backward-Euler run with a moving
    source, a network trained on the first 40 steps' solutions is applied to
    the next 20 (never seen).  Its start vector lowers the A-norm error of its
    base (the extrapolation 2u^n - u^(n-1)) on every held-out step and by at
    least 10 % of the error energy on average; an untrained network of the
    same shape gains far less.  The CG iterations from both starts are
    reported, not asserted (an A-norm gain need not shorten CG's tail)."""
    import ofeli_amd as oa
    from ofeli_amd import workloads as W
    from ofeli_amd.transient import HeatBE
    torch = _t()
    d = W.config3(64)
    S = HeatBE(ctx, d["coord"], d["elem"], d["code"], d["coef"], d["dt"])
    xy = torch.as_tensor(d["coord"], device="cuda")
    F = [W.source_field(xy, d["path"], (s + 1) * d["dt"] * 8) for s in range(60)]
    data = []
    for s in range(60):
        S.rhs(F[s])
        S.start("warm")
        x_n, x_p, b = S.x.clone(), S.xp.clone(), S.b.clone()
        xs = torch.zeros_like(b)
        info = oa.cg(S.A, b, xs, max_it=20000, toler=1e-12)
        assert info.nb_it > 0
        data.append((x_n, x_p, b, xs))
        S.x0.copy_(xs)
        S.advance()
    net = oa.LearnedGuess(S.A, layers=4, seed=0, extrapolate=True)
    fresh = oa.LearnedGuess(S.A, layers=4, seed=0, extrapolate=True)
    netr = oa.LearnedGuess(S.A, layers=4, seed=0, extrapolate=True, objective="residual")
    x0 = torch.empty_like(S.b)
    rng = np.random.default_rng(0)
    steps = 8 * 40
    k = 0
    for ep in range(8):
        for i in rng.permutation(40):
            lr = 2e-3 * 0.5 * (1 + np.cos(np.pi * k / steps))
            net.train_step(*data[i], x0, lr=lr)
            x_n, x_p, b, _ = data[i]
            netr.train_step(x_n, x_p, b, None, x0, lr=lr)
            k += 1
    y = torch.empty_like(S.b)

    def energy(x, xs):
        e = x - xs
        oa.spmv(S.A, e, y)
        return float(torch.dot(e, y))

    def res2(x, b):
        oa.spmv(S.A, x, y)
        return float(torch.sum((b - y) ** 2))

    ratios, ratios_fresh, its, rres = [], [], [], []
    for x_n, x_p, b, xs in data[40:]:
        base = 2.0 * x_n - x_p
        net.guess(x_n, x_p, b, x0)
        r = energy(x0, xs) / energy(base, xs)
        fresh.guess(x_n, x_p, b, x0)
        ratios_fresh.append(energy(x0, xs) / energy(base, xs))
        ratios.append(r)
        net.guess(x_n, x_p, b, x0)
        xb = base.clone()
        xr = torch.empty_like(x0)
        netr.guess(x_n, x_p, b, xr)
        rres.append(res2(xr, b) / res2(base, b))
        its.append((oa.cg(S.A, b, xb, max_it=20000, toler=1e-8).iterations,
                    oa.cg(S.A, b, x0, max_it=20000, toler=1e-8).iterations,
                    oa.cg(S.A, b, xr, max_it=20000, toler=1e-8).iterations))
    print("held-out error energy ratio learned/base: mean %.3f max %.3f; untrained %.3f; "
          "residual-trained residual ratio: mean %.3f max %.3f; "
          "CG iterations base / energy-trained / residual-trained: %s"
          % (np.mean(ratios), max(ratios), np.mean(ratios_fresh), np.mean(rres), max(rres),
             np.mean(its, axis=0)))
    assert max(rres) <= 1.0 + 1e-9
    assert max(ratios) <= 1.0 + 1e-9
    assert np.mean(ratios) < 0.9
    assert np.mean(ratios) < np.mean(ratios_fresh) - 0.05


def _kuhn_family(ctx, n, seeds):
    """A 3D Kuhn P1 Poisson plan and k-fields of several seeds (same mesh)."""
    import ofeli_amd as oa
    from ofeli_amd import workloads as W
    torch = _t()
    coord, code, elem = oa.api.mesh3d_kuhn(n, n, n)
    P = oa.AssemblyPlan(ctx, coord, elem, code)
    pts = W.centroids(coord, elem)
    ks = [torch.as_tensor(np.exp(W.fourier_field(W.SplitMix64(s), pts, corr=0.2)), device="cuda")
          for s in seeds]
    f = torch.ones(coord.shape[0], dtype=torch.float64, device="cuda")
    return P, ks, f


def test_refresh_equals_a_network_built_on_the_new_values(ctx):
    """ofh_mpnn_refresh re-reads the graph weights from new values: the
    guess then equals bitwise that of a network with the same parameters
    created on the new values."""
    import ofeli_amd as oa
    torch = _t()
    P, ks, f = _kuhn_family(ctx, 8, (11, 12))
    b = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
    zero = torch.zeros_like(b)
    P.build(oa.LAPLACE3DT4, coef=ks[0], f=f, matrix=True, b=b)
    M1 = oa.LearnedGuess(P.A, layers=2, seed=3)
    P.build(oa.LAPLACE3DT4, coef=ks[1], f=f, matrix=True, b=b)
    M1.refresh()
    M2 = oa.LearnedGuess(P.A, layers=2, seed=3)
    x1, x2 = torch.empty_like(b), torch.empty_like(b)
    M1.guess(zero, zero, b, x1)
    M2.guess(zero, zero, b, x2)
    torch.cuda.synchronize()
    assert torch.equal(x1, x2)


def test_learned_start_on_a_held_out_coefficient_field(ctx):
    """Steady solves with new coefficient fields: networks trained on the
    solutions for 8 k-fields of a 14^3 Kuhn mesh (base x = 0, line search
    along their direction) start the solve of a 9th, unseen k-field."""
    import ofeli_amd as oa
    torch = _t()
    P, ks, f = _kuhn_family(ctx, 14, range(21, 30))
    n = P.nb_eq
    zero = torch.zeros(n, dtype=torch.float64, device="cuda")
    data = []
    for k in ks:
        b = torch.zeros(n, dtype=torch.float64, device="cuda")
        P.build(oa.LAPLACE3DT4, coef=k, f=f, matrix=True, b=b)
        xs = torch.zeros_like(b)
        assert oa.cg(P.A, b, xs, max_it=5000, toler=1e-11).nb_it > 0
        data.append((k, b, xs))
    out = {}
    for obj in ("energy", "residual"):
        M = oa.LearnedGuess(P.A, layers=4, seed=0, objective=obj)
        x0 = torch.empty_like(zero)
        rng = np.random.default_rng(0)
        for ep in range(30):
            for i in rng.permutation(8):
                k, b, xs = data[i]
                P.build(oa.LAPLACE3DT4, coef=k, f=f, matrix=True, b=b)
                M.refresh()
                M.train_step(zero, zero, b, xs, x0, lr=2e-3, want_loss=False)
        k, b, xs = data[8]
        P.build(oa.LAPLACE3DT4, coef=k, f=f, matrix=True, b=b)
        M.refresh()
        loss, _ = M.loss_grad(zero, zero, b, xs, x0, want_grad=False)
        x_l = torch.empty_like(b)
        M.guess(zero, zero, b, x_l)
        r = torch.empty_like(b)
        oa.spmv(P.A, x_l, r)
        rel0 = float(torch.linalg.norm(b - r) / torch.linalg.norm(b))
        i_l = oa.cg(P.A, b, x_l, max_it=5000, toler=1e-8)
        out[obj] = (loss, rel0, i_l.nb_it)
        M.close()
    x_z = torch.zeros_like(b)
    i_z = oa.cg(P.A, b, x_z, max_it=5000, toler=1e-8)
    print("held-out k-field: zero start %d it; " % i_z.nb_it + "; ".join(
        "%s-trained: loss %.3f, start relres %.3f, %d it" % (k, *v) for k, v in out.items()))
    assert out["energy"][0] < 0.5 and out["residual"][1] < 1.0
