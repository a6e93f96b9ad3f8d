"""Solver chunks replayed from captured graphs give bitwise the results of
direct launches, across repeated solves with eager work in between, for the
two-kernel (L2-resident) and three-kernel CG forms and both mixed forms."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def poisson(ctx, nx):
    import torch
    import ofeli_amd as oa
    coord, code, elem = oa.api.mesh2d(nx, nx)
    rng = np.random.default_rng(nx)
    P = oa.AssemblyPlan(ctx, coord, elem, code)
    b = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
    P.build(oa.DC2DT3, coef=torch.as_tensor(np.exp(rng.standard_normal(elem.shape[0])), device="cuda"),
            f=torch.as_tensor(1.0 + rng.random(coord.shape[0]), device="cuda"), matrix=True, b=b)
    return P, b


@pytest.mark.parametrize("nx", [100, 660])  # 9,801 rows (two kernels), 434,281 (three)
@pytest.mark.parametrize("mixed", [False, True])
def test_graph_chunks_equal_direct_launches(ctx, nx, mixed):
    import torch
    import ofeli_amd as oa
    P, b = poisson(ctx, nx)
    fn = oa.cg_mixed if mixed else oa.cg
    ctx.use_graphs(False)
    x_ref = torch.zeros_like(b)
    i_ref = fn(P.A, b, x_ref, max_it=20000, toler=1e-9)
    ctx.use_graphs(True)
    try:
        for rep in range(3):
            x = torch.zeros_like(b)
            info = fn(P.A, b, x, max_it=20000, toler=1e-9)
            torch.cuda.synchronize()
            assert torch.equal(x, x_ref), rep
            assert info.iterations == i_ref.iterations and info.nb_it == i_ref.nb_it
            y = torch.zeros_like(b)
            oa.spmv(P.A, x, y)  # eager work between replays
            torch.cuda.synchronize()
        # FP64 CG stops on the recursive residual (CG.h:144); mixed on the true one
        assert i_ref.nb_it > 0 and info.true_relres <= (1e-9 if mixed else 1.2e-9)
    finally:
        ctx.use_graphs(False)


@pytest.mark.parametrize("mixed", [False, True])
def test_recreated_matrix_never_replays_a_stale_graph(ctx, mixed):
    """A matrix destroyed and re-created with the same size (the allocator may
    hand back the same addresses, the SELL copy is rebuilt) must not replay a
    graph captured for the old one: with graphs on, the new solve equals
    direct launches bitwise."""
    import torch
    import ofeli_amd as oa
    fn = oa.cg_mixed if mixed else oa.cg
    ctx.use_graphs(True)
    try:
        P, b = poisson(ctx, 100)
        x = torch.zeros_like(b)
        fn(P.A, b, x, max_it=20000, toler=1e-9)
        P.close()
        del P, b
        torch.cuda.synchronize()
        coord, code, elem = oa.api.mesh2d(100, 100)
        rng = np.random.default_rng(12345)
        P2 = oa.AssemblyPlan(ctx, coord, elem, code)
        b2 = torch.zeros(P2.nb_eq, dtype=torch.float64, device="cuda")
        P2.build(oa.DC2DT3, coef=torch.as_tensor(np.exp(2.0 * rng.standard_normal(elem.shape[0])),
                                                 device="cuda"),
                 f=torch.as_tensor(rng.random(coord.shape[0]), device="cuda"), matrix=True, b=b2)
        x_g = torch.zeros_like(b2)
        i_g = fn(P2.A, b2, x_g, max_it=20000, toler=1e-9)
        ctx.use_graphs(False)
        x_d = torch.zeros_like(b2)
        i_d = fn(P2.A, b2, x_d, max_it=20000, toler=1e-9)
        torch.cuda.synchronize()
        assert torch.equal(x_g, x_d)
        assert i_g.iterations == i_d.iterations and i_g.nb_it == i_d.nb_it > 0
    finally:
        ctx.use_graphs(False)
