"""Start vectors from previous solutions (ofh_proj_*) against oracle/proj.py."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import proj as PJ
from tests.gpu_systems import be_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("minres", [False, True])
def test_projection_matches_oracle_through_a_rebuild(ctx, minres):
    import torch
    import ofeli_amd as oa
    P, g, b, xs, x_n, x_p = be_system(ctx, 20, seed=11)
    A = P.A
    rp, col = A.pattern()
    Ah = sp.csr_matrix((A.values(), col, rp), shape=(A.n, A.n))
    rng = np.random.default_rng(0)
    G = oa.ProjectionGuess(A, m=8, minres=minres)
    O = PJ.Projection(Ah, 8, minres=minres)
    x0 = torch.empty_like(b)
    counts = []
    for k in range(13):
        bk = torch.as_tensor(rng.standard_normal(A.n), device="cuda")
        G.guess(bk, x0)
        want = O.guess(bk.cpu().numpy())
        got = x0.cpu().numpy()
        if k == 0:
            assert not got.any()
        else:
            assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want), k
        x = torch.zeros_like(b)
        info = oa.cg(A, bk, x, max_it=5000, toler=1e-12)
        assert info.nb_it > 0
        G.update(x)
        O.update(x.cpu().numpy())
        assert G.size() == O.count
        counts.append(O.count)
    assert counts[7] == 8 and counts[8] == 4  # the ninth update rebuilt the basis
    # a combination of stored solutions is dropped: nothing changes
    xa = torch.zeros_like(b)
    G.guess(bk, x0)
    y = x0.clone()
    G.update(1.5 * y)
    assert G.size() == O.count and O.update(1.5 * y.cpu().numpy()) is False


def test_projection_start_is_no_worse_than_warm_in_a_norm(ctx):
    import torch
    import ofeli_amd as oa
    P, g, b, xs, x_n, x_p = be_system(ctx, 24, seed=12)
    A = P.A
    G = oa.ProjectionGuess(A, m=16)
    x0 = torch.empty_like(b)
    y = torch.empty_like(b)
    prev = torch.zeros_like(b)
    base = b.clone()
    for k in range(20):
        bk = base * (1.0 + 0.3 * np.sin(0.4 * k)) + 0.01 * k * torch.roll(base, k)
        x = torch.zeros_like(b)
        oa.cg(A, bk, x, max_it=5000, toler=1e-12)
        if k > 0:
            G.guess(bk, x0)
            e_g, e_w = x0 - x, prev - x
            oa.spmv(A, e_g, y)
            eg = float(torch.dot(e_g, y))
            oa.spmv(A, e_w, y)
            ew = float(torch.dot(e_w, y))
            assert eg <= ew * (1 + 1e-8) + 1e-30, (k, eg, ew)
        G.update(x)
        prev = x
