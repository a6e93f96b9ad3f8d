"""Small GPU-assembled systems shared by the GPU tests."""
import numpy as np

from oracle import mpnn as N


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
