"""Batched persistent CG (ofh_batch_*) against per-system oracle solves."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _t():
    import torch
    return torch


def small_batch(ctx, sizes=(5, 9, 12, 7, 16, 6), seed=0):
    import ofeli_amd as oa
    rng = np.random.default_rng(seed)
    meshes, ks, fs = [], [], []
    for n in sizes:
        meshes.append(oa.api.mesh2d(n, n))
        ks.append(np.exp(0.75 * rng.standard_normal(2 * n * n)))
        fs.append(1.0 + rng.random((n + 1) ** 2))
    coord, code, elem, sys_rows = oa.concat_meshes(meshes)
    P = oa.AssemblyPlan(ctx, coord, elem, code)
    b = _t().zeros(P.nb_eq, dtype=_t().float64, device="cuda")
    P.build(oa.DC2DT3, coef=_t().as_tensor(np.concatenate(ks), device="cuda"),
            f=_t().as_tensor(np.concatenate(fs), device="cuda"), matrix=True, b=b)
    return P, b, sys_rows, sizes, ks, fs


def oracle_system(n, k, f):
    m = O.mesh2d(n, n)
    dof, neq = O.number_equations(m.code)
    rp, ci = O.node_graph(m.elem, 1, dof, neq)
    a, b = O.assemble(O.DC2DT3, m, dof, neq, rp, ci, f=f, coef=k)
    return rp, ci, a, b


def test_batch_blocks_equal_per_system_assembly(ctx):
    P, b, sys_rows, sizes, ks, fs = small_batch(ctx)
    rp, col = P.A.pattern()
    val = P.A.values()
    bh = b.cpu().numpy()
    for s, n in enumerate(sizes):
        orp, oci, oa_, ob = oracle_system(n, ks[s], fs[s])
        r0, r1 = sys_rows[s], sys_rows[s + 1]
        k0 = rp[r0]
        assert np.array_equal(rp[r0:r1 + 1] - k0, orp)
        assert np.array_equal(col[rp[r0]:rp[r1]] - r0, oci - 1)
        assert np.array_equal(val[rp[r0]:rp[r1]], oa_)
        assert np.array_equal(bh[r0:r1], ob)


@pytest.mark.parametrize("block,tpr", [(256, 1), (512, 4), (1024, 8), (1024, 0)])
def test_batch_fp64_matches_oracle(ctx, block, tpr):
    import ofeli_amd as oa
    P, b, sys_rows, sizes, ks, fs = small_batch(ctx)
    B = oa.Batch(P.A, sys_rows)
    B.set_config(block, tpr)
    x = _t().zeros(P.nb_eq, dtype=_t().float64, device="cuda")
    it, st, rr, info = B.cg(b, x, max_it=1000, toler=1e-10)
    xh = x.cpu().numpy()
    for s, n in enumerate(sizes):
        orp, oci, oa_, ob = oracle_system(n, ks[s], fs[s])
        ret, xo, _ = O.cg(orp, oci, oa_, ob, max_it=1000, toler=1e-10)
        assert ret > 0 and abs(st[s] - ret) <= 1 and it[s] == st[s]
        r0, r1 = sys_rows[s], sys_rows[s + 1]
        assert np.linalg.norm(xh[r0:r1] - xo) / np.linalg.norm(xo) < 1e-9
    assert info.nb_it > 0


def test_batch_mixed_fp64_verified(ctx):
    import ofeli_amd as oa
    P, b, sys_rows, sizes, ks, fs = small_batch(ctx, sizes=(8, 20, 33, 12, 25))
    B = oa.Batch(P.A, sys_rows)
    x = _t().zeros(P.nb_eq, dtype=_t().float64, device="cuda")
    it, st, rr, info = B.cg(b, x, mixed=True, max_it=2000, toler=1e-8)
    assert (st > 0).all() and (rr <= 1e-8).all()
    xh = x.cpu().numpy()
    for s, n in enumerate(sizes):
        orp, oci, oa_, ob = oracle_system(n, ks[s], fs[s])
        ret, xo, _ = O.cg(orp, oci, oa_, ob, max_it=2000, toler=1e-12)
        r0, r1 = sys_rows[s], sys_rows[s + 1]
        xs = xh[r0:r1]
        true = np.linalg.norm(ob - O.spmv(orp, oci, oa_, xs)) / np.linalg.norm(ob)
        assert true <= 1e-8
        # error against the oracle's rtol-1e-12 solution, at the FP64 level
        assert np.linalg.norm(xs - xo) / np.linalg.norm(xo) <= 1e-7


def test_batch_autotune_and_not_block_diagonal(ctx):
    import ofeli_amd as oa
    P, b, sys_rows, sizes, ks, fs = small_batch(ctx)
    B = oa.Batch(P.A, sys_rows)
    x = _t().zeros(P.nb_eq, dtype=_t().float64, device="cuda")
    blk, tpr, ms = B.autotune(b, x, max_it=1000, toler=1e-10)
    assert (blk, tpr) == (1024, 0) or (blk in (256, 512, 1024) and tpr in (1, 2, 4, 8))
    assert ms > 0
    assert B.config() == (blk, tpr)
    assert float(x.abs().max()) == 0.0  # the guesses are restored
    bad = sys_rows.copy()
    bad[1] += 1  # a row whose columns cross into the previous system
    with pytest.raises(oa.OfhError):
        oa.Batch(P.A, bad)


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_graph_replays_equal_eager(ctx, mixed):
    """The captured launch replayed several times, with eager work and
    synchronisations in between, gives bitwise the eager launch's result."""
    import ofeli_amd as oa
    torch = _t()
    P, b, sys_rows, sizes, ks, fs = small_batch(ctx, sizes=(6, 14, 9, 21))
    B = oa.Batch(P.A, sys_rows)
    B.use_graph(False)
    x_eager = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
    it_e, st_e, rr_e, _ = B.cg(b, x_eager, mixed=mixed, max_it=1500, toler=1e-9)
    B.use_graph(True)
    x = torch.zeros_like(x_eager)
    for rep in range(4):
        x.zero_()
        it, st, rr, _ = B.cg(b, x, mixed=mixed, max_it=1500, toler=1e-9)
        torch.cuda.synchronize()
        assert torch.equal(x, x_eager), rep
        assert np.array_equal(it, it_e) and np.array_equal(st, st_e) and np.array_equal(rr, rr_e)
        y = torch.zeros_like(x)
        oa.spmv(P.A, x, y)  # eager work between replays
        torch.cuda.synchronize()
    # a second right-hand side recaptures, the first is still reproduced after
    b2 = 2.0 * b
    x2 = torch.zeros_like(x)
    B.cg(b2, x2, mixed=mixed, max_it=1500, toler=1e-9)
    x.zero_()
    B.cg(b, x, mixed=mixed, max_it=1500, toler=1e-9)
    assert torch.equal(x, x_eager)
    assert float(torch.linalg.norm(x2 - 2.0 * x_eager) / torch.linalg.norm(x2)) < 1e-7


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_resident_equals_streamed(ctx, mixed):
    """The resident form (direction in LDS, r and x in registers / global)
    sums every row and every dot product in the streamed 1024-thread form's
    order, so both give bitwise the same iterates.  Systems span one to
    twelve row slots per thread (an 111 x 111 cell square has 110^2 = 12,100
    equations, the limit is 12,288)."""
    import ofeli_amd as oa
    torch = _t()
    sizes = (111, 70, 33, 5, 12, 45)
    P, b, sys_rows, _, ks, fs = small_batch(ctx, sizes=sizes, seed=3)
    B = oa.Batch(P.A, sys_rows)
    B.use_graph(False)
    out = {}
    for tpr in (1, 0):
        B.set_config(1024, tpr)
        x = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
        it, st, rr, info = B.cg(b, x, mixed=mixed, max_it=4000, toler=1e-9)
        out[tpr] = (x, it, st, rr)
    x1, it1, st1, rr1 = out[1]
    x0, it0, st0, rr0 = out[0]
    assert (st0 > 0).all() and (rr0 <= 1e-9).all()
    assert np.array_equal(it0, it1) and np.array_equal(st0, st1) and np.array_equal(rr0, rr1)
    assert torch.equal(x0, x1)
    # true residual per system
    y = torch.zeros_like(x0)
    oa.spmv(P.A, x0, y)
    r = (b - y).cpu().numpy()
    bh = b.cpu().numpy()
    for s in range(len(sizes)):
        r0, r1 = sys_rows[s], sys_rows[s + 1]
        assert np.linalg.norm(r[r0:r1]) / np.linalg.norm(bh[r0:r1]) <= (1.2e-9 if not mixed else 1e-9)


def test_batch_autotune_cache(ctx):
    """The autotuner's choice is kept per batch shape: a new batch of the
    same systems takes it without timing (same result, reported time equal
    to the first tuning's), a batch of another shape is timed anew."""
    import time
    import ofeli_amd as oa
    torch = _t()
    P, b, sys_rows, sizes, ks, fs = small_batch(ctx, sizes=(14, 9, 22), seed=7)
    x = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
    B1 = oa.Batch(P.A, sys_rows)
    r1 = B1.autotune(b, x, max_it=1000, toler=1e-9)
    B2 = oa.Batch(P.A, sys_rows)
    t0 = time.perf_counter()
    r2 = B2.autotune(b, x, max_it=1000, toler=1e-9)
    t_cached = time.perf_counter() - t0
    assert r2 == r1 and B2.config() == r1[:2]
    assert t_cached < 0.05, t_cached
    # another solver variant is another key: timed anew, still a valid shape
    blk, tpr, ms = B2.autotune(b, x, mixed=True, max_it=1000, toler=1e-9)
    assert (blk, tpr) == (1024, 0) or (blk in (256, 512, 1024) and tpr in (1, 2, 4, 8))
    assert float(x.abs().max()) == 0.0


def test_batch_resident_size_limit(ctx):
    """Systems above 12,288 rows cannot use the resident form: set_config
    refuses it and the autotuner does not offer it."""
    import ofeli_amd as oa
    torch = _t()
    P, b, sys_rows, sizes, ks, fs = small_batch(ctx, sizes=(112, 6))  # 111^2 = 12,321 rows
    B = oa.Batch(P.A, sys_rows)
    with pytest.raises(oa.OfhError):
        B.set_config(1024, 0)
    x = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
    blk, tpr, ms = B.autotune(b, x, max_it=3000, toler=1e-8)
    assert tpr in (1, 2, 4, 8)


def test_batch_resident_follows_new_values(ctx):
    """The resident form reads an ELL copy of the matrix; re-assembling the
    matrix with new coefficients after a solve refreshes that copy, so the
    next resident solve again equals the streamed one bitwise."""
    import ofeli_amd as oa
    torch = _t()
    sizes = (20, 9, 31)
    P, b, sys_rows, _, ks, fs = small_batch(ctx, sizes=sizes, seed=5)
    B = oa.Batch(P.A, sys_rows)
    B.set_config(1024, 0)
    x = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
    B.cg(b, x, max_it=2000, toler=1e-9)
    rng = np.random.default_rng(11)
    k2 = np.concatenate([np.exp(0.5 * rng.standard_normal(2 * n * n)) for n in sizes])
    f2 = np.concatenate([1.0 + rng.random((n + 1) ** 2) for n in sizes])
    b.zero_()
    P.build(oa.DC2DT3, coef=torch.as_tensor(k2, device="cuda"), f=torch.as_tensor(f2, device="cuda"),
            matrix=True, b=b)
    out = {}
    for tpr in (0, 1):
        B.set_config(1024, tpr)
        for mixed in (False, True):
            x = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
            it, st, rr, _ = B.cg(b, x, mixed=mixed, max_it=2000, toler=1e-9)
            out[(tpr, mixed)] = (x, it)
    for mixed in (False, True):
        assert torch.equal(out[(0, mixed)][0], out[(1, mixed)][0])
        assert np.array_equal(out[(0, mixed)][1], out[(1, mixed)][1])
