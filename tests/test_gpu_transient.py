"""The device-resident backward-Euler driver (ofeli_amd.transient.HeatBE)
against the oracle's time stepping, as the reference's TimeStepping runs it
over DC2DT3 (TimeStepping.cpp:174-212, 444-456; Equa_Therm.h:162-218):
per step (C/dt + K) u^{n+1} = C/dt u^n + F^{n+1}, CG+DIAG_PREC warm-started
from u^n (the persistent solution vector, TimeStepping.cpp:207).

Per step, from the GPU's own u^n: the right-hand side is bitwise the
oracle's and the GPU solution is within 1e-10 relative L2 of the oracle's
solve from the same start; over 6 free-running steps the GPU's u stays
within 1e-8 of the oracle's.  The matrix, assembled once, is bitwise the
oracle's.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

NX, DT, STEPS = 24, 2e-3, 6


def source(coord, t):
    cx, cy = 0.5 + 0.3 * np.sin(9.0 * t), 0.5 + 0.3 * np.cos(7.0 * t)
    return 50.0 * np.exp(-((coord[:, 0] - cx) ** 2 + (coord[:, 1] - cy) ** 2) / 0.01)


def test_heat_be_steps_match_oracle(ctx):
    import torch
    import ofeli_amd as oa
    from ofeli_amd.transient import HeatBE
    coord, code, elem = oa.api.mesh2d(NX, NX)
    rng = np.random.default_rng(21)
    k = np.exp(0.75 * rng.standard_normal(elem.shape[0]))
    H = HeatBE(ctx, coord, elem, code, k, DT)
    m = O.mesh2d(NX, NX)
    dof, neq = O.number_equations(m.code)
    rp, ci = O.node_graph(m.elem, 1, dof, neq)
    prm = [DT, 1.0, 1.0]
    a, _ = O.assemble(O.DC2DT3_BE, m, dof, neq, rp, ci, f=np.zeros(m.nb_nodes), coef=k, prm=prm,
                      u_prev=np.zeros(m.nb_nodes))
    assert np.array_equal(H.A.values(), a)
    u_free = np.zeros(m.nb_nodes)  # oracle's free-running nodal solution
    free = m.code[:, 0] == 0
    for s in range(STEPS):
        t = (s + 1) * DT
        f = source(coord, t)
        u_gpu = H.u.cpu().numpy()
        # one step of the driver
        H.rhs(torch.as_tensor(f, device="cuda"))
        _, bo = O.assemble(O.DC2DT3_BE, m, dof, neq, rp, ci, f=f, coef=k, prm=prm, u_prev=u_gpu,
                           matrix=False)
        assert np.array_equal(H.b.cpu().numpy(), bo), s
        H.start("warm")
        assert np.array_equal(H.x0.cpu().numpy(), u_gpu[free])
        info = H.solve(toler=1e-8)
        ret, xo, _ = O.cg(rp, ci, a, bo, x0=u_gpu[free], max_it=5000, toler=1e-8)
        assert ret > 0 and info.nb_it > 0 and abs(info.nb_it - ret) <= 1, (s, info.nb_it, ret)
        x = H.x0.cpu().numpy()
        assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo), s
        H.advance("warm")
        assert np.array_equal(H.u.cpu().numpy(), O.insert_bc(m.code, x))
        # the oracle's own free-running trajectory
        _, bf = O.assemble(O.DC2DT3_BE, m, dof, neq, rp, ci, f=f, coef=k, prm=prm, u_prev=u_free,
                           matrix=False)
        r2, xf, _ = O.cg(rp, ci, a, bf, x0=u_free[free], max_it=5000, toler=1e-8)
        assert r2 > 0
        u_free = O.insert_bc(m.code, xf)
        e = np.linalg.norm(H.u.cpu().numpy() - u_free) / np.linalg.norm(u_free)
        assert e <= 1e-8, (s, e)
    assert H.step_count == STEPS


def test_projection_judges_the_network(ctx):
    """HeatBE(judge_every=k): every k-th step computes both the projected
    start and the network's correction of it and keeps the network only if
    its start's residual is below judge_keep times the projection's.  An
    untrained network does not pass: after each judged step the projection
    alone starts the solves, and every solve meets the stop test."""
    import torch
    import ofeli_amd as oa
    from ofeli_amd.transient import HeatBE
    coord, code, elem = oa.api.mesh2d(NX, NX)
    k = np.exp(0.5 * np.random.default_rng(4).standard_normal(elem.shape[0]))
    H = HeatBE(ctx, coord, elem, code, k, DT, judge_every=4, judge_keep=0.5)
    proj = oa.ProjectionGuess(H.A, m=8, minres=True)
    net = oa.LearnedGuess(H.A, layers=2, seed=5)
    for s in range(12):
        info = H.step(torch.as_tensor(source(coord, (s + 1) * DT), device="cuda"),
                      guess=(proj, net), toler=1e-8)
        assert info.nb_it >= 0 and info.relres <= 1e-8
    assert H.judged == 3
    assert H.net_steps <= H.judged
