"""GPU parity of the C-ABI against the CPU oracle: SpMV, Prec, SSOR, CG,
BiCGStab, GMRes and the mixed-precision CG.

Integer/index results and every quantity the reference computes by
sequential sums (SpMV row sums, SSOR sweeps) must be bitwise equal; iterative
solutions are compared with the tolerance stated in each test.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def dev(a):
    torch = _torch()
    return torch.as_tensor(np.ascontiguousarray(a, np.float64), device="cuda")


def host(t):
    _torch().cuda.synchronize()
    return t.detach().cpu().numpy()


def dc2d_system(n=48, seed=3):
    """Oracle-assembled DC2DT3 system with lognormal k and a bump source."""
    m = O.mesh2d(n, n)
    dof, neq = O.number_equations(m.code)
    rp, ci = O.node_graph(m.elem, 1, dof, neq)
    rng = np.random.default_rng(seed)
    k = np.exp(0.75 * rng.standard_normal(m.nb_elements))
    f = 1.0 + rng.random(m.nb_nodes)
    a, b = O.assemble(O.DC2DT3, m, dof, neq, rp, ci, f=f, coef=k)
    return rp, ci, a, b


def gpu_csr(ctx, rp, ci, a):
    import ofeli_amd as oa
    return oa.CsrMatrix(ctx, len(rp) - 1, rp.astype(np.int32), (ci - 1).astype(np.int32), a)


def rel(x, y):
    return np.linalg.norm(x - y) / np.linalg.norm(y)


def test_spmv_bitexact(ctx):
    import ofeli_amd as oa
    rp, ci, a, b = dc2d_system()
    A = gpu_csr(ctx, rp, ci, a)
    x = np.random.default_rng(1).standard_normal(A.n)
    y = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
    oa.spmv(A, dev(x), y)
    assert np.array_equal(host(y), O.spmv(rp, ci, a, x))


@pytest.mark.parametrize("shape", [(28, 28, 28), (400, 20)])
def test_spmv_sell_bitexact_and_value_changes(ctx, shape):
    """y = A x from the SELL-64 copy (ragged last chunk: 21,952 and 8,000 rows
    are not multiples of 64 x 4) bitwise the oracle's row sums, also after the
    values change (the copy is gathered again) and with non-finite x entries
    (padding adds -0.0, never 0 * inf); FP32 bitwise the sequential float sums."""
    import ofeli_amd as oa
    torch = _torch()
    rp, ci, a = grid_system(shape, 31)
    A = gpu_csr(ctx, rp, ci, a)
    lens = np.diff(rp)
    pad = np.zeros(-(-A.n // 64) * 64, np.int64)
    pad[:A.n] = lens
    stored = int(pad.reshape(-1, 64).max(axis=1).sum() * 64)
    assert A.sell_info() == {"in_use": True, "stored": stored, "col_bytes": 2}
    rng = np.random.default_rng(32)
    x = rng.standard_normal(A.n)
    y = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    oa.spmv(A, dev(x), y)
    assert np.array_equal(host(y), O.spmv(rp, ci, a, x))
    a2 = a * (1.0 + 0.5 * np.cos(np.arange(len(a))))
    A.set_values(a2)
    oa.spmv(A, dev(x), y)
    assert np.array_equal(host(y), O.spmv(rp, ci, a2, x))
    xs = x.copy()
    xs[[5, 77]] = [np.inf, np.nan]
    oa.spmv(A, dev(xs), y)
    with np.errstate(invalid="ignore"):
        assert np.array_equal(host(y), O.spmv(rp, ci, a2, xs), equal_nan=True)
    x32 = x.astype(np.float32)
    y32 = torch.zeros(A.n, dtype=torch.float32, device="cuda")
    oa.spmv_f32(A, torch.as_tensor(x32, device="cuda"), y32)
    a32 = a2.astype(np.float32)
    ref = np.zeros(A.n, np.float32)
    for i in range(A.n):  # sequential float32 row sums
        s = np.float32(0.0)
        for k in range(rp[i], rp[i + 1]):
            s = np.float32(s + np.float32(a32[k] * x32[ci[k] - 1]))
        ref[i] = s
    assert np.array_equal(y32.cpu().numpy(), ref)


def seq_f32(rp, ci, a32, x32):
    """Sequential float32 row sums (the CSR row order, one rounding per step)."""
    ref = np.zeros(len(rp) - 1, np.float32)
    with np.errstate(invalid="ignore", over="ignore"):
        for i in range(len(rp) - 1):
            s = np.float32(0.0)
            for k in range(rp[i], rp[i + 1]):
                s = np.float32(s + np.float32(a32[k] * x32[ci[k] - 1]))
            ref[i] = s
    return ref


def tridiag_system(n=5000, seed=7):
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                rows.append(i)
                cols.append(j)
                vals.append(2.5 if j == i else -rng.uniform(0.5, 1.0))
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, np.asarray(rows) + 1, 1)
    return np.cumsum(rp), np.asarray(cols, np.int64) + 1, np.asarray(vals)


@pytest.mark.parametrize("case", ["grid3d", "tridiag"])
def test_spmv_f32_diagonal_copy_bitexact(ctx, case):
    """ofh_spmv_f32 from the diagonal copy (7 diagonals of a 3D grid: 8 slots
    per row; a tridiagonal matrix: 4) bitwise the sequential float32 CSR row
    sums, and ofh_spmv_f64 from the FP64 copy (the 3D grid; the tridiagonal
    matrix keeps SELL-64, whose 3 entries per row take fewer bytes) bitwise
    the oracle's, through changes of the values (the copies are gathered
    again) and with non-finite x entries: a row's absent diagonals are masked
    out, so an inf or nan only reaches the rows that really couple to it.
    The 3D grid's values are symmetric bit for bit, so the products read the
    symmetric form (4 slots per row, lower values from their mirrors); values
    scaled entry by entry are not, and the full copy returns; values doubled
    are symmetric again.  The random tridiagonal matrix is not symmetric."""
    import ofeli_amd as oa
    torch = _torch()
    rp, ci, a = grid_system((28, 28, 28), 31) if case == "grid3d" else tridiag_system()
    A = gpu_csr(ctx, rp, ci, a)
    want = {"grid3d": (True, 7, 8, True), "tridiag": (False, 3, 4, False)}[case]
    assert A.dia_info() == {"fp32": True, "fp64": want[0], "diagonals": want[1],
                            "row_slots": want[2], "symmetric": want[3]}
    x64 = np.random.default_rng(34).standard_normal(A.n)
    y64 = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    rng = np.random.default_rng(33)
    x32 = rng.standard_normal(A.n).astype(np.float32)
    y32 = torch.zeros(A.n, dtype=torch.float32, device="cuda")
    unsym = a * (1.0 + 0.5 * np.cos(np.arange(len(a))))
    for vals, sym in ((a, want[3]), (unsym, False), (2.0 * a, want[3]), (unsym, False)):
        if vals is not a:
            A.set_values(vals)
        assert A.dia_info()["symmetric"] is sym
        oa.spmv_f32(A, torch.as_tensor(x32, device="cuda"), y32)
        assert np.array_equal(y32.cpu().numpy(), seq_f32(rp, ci, vals.astype(np.float32), x32))
        oa.spmv(A, dev(x64), y64)
        assert np.array_equal(host(y64), O.spmv(rp, ci, vals, x64))
        if vals is a:
            continue
        xs = x32.copy()
        xs[[0, 5, A.n - 1, A.n // 2]] = [np.inf, np.nan, -np.inf, np.nan]
        oa.spmv_f32(A, torch.as_tensor(xs, device="cuda"), y32)
        ref = seq_f32(rp, ci, vals.astype(np.float32), xs)
        got = y32.cpu().numpy()
        assert np.isfinite(ref).sum() > A.n - 40
        assert np.array_equal(got, ref, equal_nan=True)
        xs64 = x64.copy()
        xs64[[0, 5, A.n - 1, A.n // 2]] = [np.inf, np.nan, -np.inf, np.nan]
        oa.spmv(A, dev(xs64), y64)
        with np.errstate(invalid="ignore"):
            assert np.array_equal(host(y64), O.spmv(rp, ci, vals, xs64), equal_nan=True)


def test_mixed_cg_diag_scaled_and_unscaled_inner_iterations(ctx):
    """Mixed CG with DIAG_PREC on a matrix with a diagonal copy runs its FP32
    inner iteration on D^-1/2 A D^-1/2 (positive pivots); on -A (negative
    pivots) it keeps the Jacobi-preconditioned inner iteration on A.  Both
    reach the FP64 true-residual tolerance and the same solution, and the
    scaled form needs no more inner iterations than FP64 CG needs by much."""
    import ofeli_amd as oa
    torch = _torch()
    rp, ci, a = grid_system((28, 28, 28), 31)
    xs = np.cos(np.arange(len(rp) - 1) * 0.01)
    b = O.spmv(rp, ci, a, xs)
    out = {}
    for sign in (1.0, -1.0):
        A = gpu_csr(ctx, rp, ci, sign * a)
        assert A.dia_info()["fp32"]
        x = torch.zeros(A.n, dtype=torch.float64, device="cuda")
        info = oa.cg_mixed(A, dev(sign * b), x, prec=oa.DIAG_PREC, max_it=5000, toler=1e-10)
        assert info.nb_it > 0 and info.fallback == 0, info.as_dict()
        assert info.true_relres <= 1e-10
        assert rel(host(x), xs) <= 1e-8
        out[sign] = info.iterations
    x = torch.zeros(len(xs), dtype=torch.float64, device="cuda")
    it64 = oa.cg(gpu_csr(ctx, rp, ci, a), dev(b), x, prec=oa.DIAG_PREC, max_it=5000,
                 toler=1e-10).iterations
    print("inner iterations: scaled %d, unscaled %d, FP64 CG %d" % (out[1.0], out[-1.0], it64))
    assert out[1.0] <= 1.1 * it64 + 5


def _drop_entry(rp, ci, a, k):
    """The CSR arrays without stored entry k."""
    row = int(np.searchsorted(rp, k, side="right")) - 1
    rp2 = rp.copy()
    rp2[row + 1:] -= 1
    return rp2, np.delete(ci, k), np.delete(a, k)


def _entry(rp, ci, i, j):
    """Index of stored entry (i, j), 0-based row and column."""
    k = rp[i] + int(np.searchsorted(ci[rp[i]:rp[i + 1]], j + 1))
    assert ci[k] == j + 1
    return k


def test_symmetric_form_needs_mirrored_bits(ctx):
    """The products read the symmetric form of the diagonal copy only where
    every entry's mirror is stored with the same 64 bits: the 3D grid system
    does; one value moved by one ulp, a mirrored pair stored as +0.0 / -0.0,
    or one off-diagonal entry dropped without its mirror (the diagonals still
    pair up) do not; a mirrored pair dropped together does.  In every case
    ofh_spmv_f64 and ofh_spmv_f32 stay the CSR row sums bit for bit."""
    import ofeli_amd as oa
    torch = _torch()
    rp, ci, a = grid_system((12, 12, 12), 31)
    n = len(rp) - 1
    i, j = 500, 501  # a coupled pair of rows
    kij, kji = _entry(rp, ci, i, j), _entry(rp, ci, j, i)
    ulp = a.copy()
    ulp[kij] = np.nextafter(ulp[kij], np.inf)
    zeros = a.copy()
    zeros[kij], zeros[kji] = 0.0, -0.0
    one = _drop_entry(rp, ci, a, kij)
    both = _drop_entry(*_drop_entry(rp, ci, a, max(kij, kji)), min(kij, kji))
    cases = [((rp, ci, a), True), ((rp, ci, ulp), False), ((rp, ci, zeros), False),
             (one, False), (both, True)]
    x64 = np.random.default_rng(41).standard_normal(n)
    x32 = x64.astype(np.float32)
    for (r, c, v), sym in cases:
        A = gpu_csr(ctx, r, c, v)
        info = A.dia_info()
        assert info["fp32"] and info["diagonals"] == 7 and info["symmetric"] is sym
        y64 = torch.zeros(n, dtype=torch.float64, device="cuda")
        oa.spmv(A, dev(x64), y64)
        assert np.array_equal(host(y64), O.spmv(r, c, v, x64))
        y32 = torch.zeros(n, dtype=torch.float32, device="cuda")
        oa.spmv_f32(A, torch.as_tensor(x32, device="cuda"), y32)
        assert np.array_equal(y32.cpu().numpy(), seq_f32(r, c, v.astype(np.float32), x32))
        A.close()


def test_diagonal_copy_not_used_where_it_does_not_pay(ctx):
    """A 2D 5-point grid (5 diagonals: 8 floats per row against ~30 bytes of
    SELL-64 entries) and a matrix with far couplings (5 diagonals, but mostly
    3 entries per row) keep the SELL-64 copies in both precisions."""
    rp, ci, a = grid_system((400, 20), 31)
    info = gpu_csr(ctx, rp, ci, a).dia_info()
    assert info["fp32"] is False and info["fp64"] is False
    rp, ci, a = far_coupled_system()
    info = gpu_csr(ctx, rp, ci, a).dia_info()
    assert info["fp32"] is False and info["fp64"] is False


def far_coupled_system(n=40000, far=35000):
    """SPD system: a 1D Laplacian plus couplings i <-> i + far.  The far
    columns lie more than 32,767 rows from their 64-row chunk, so the SELL-64
    copy keeps 32-bit columns only."""
    rows, cols, vals = [], [], []
    for i in range(n):
        ent = {i: 4.0 + (i % 7) * 0.25}
        if i > 0:
            ent[i - 1] = -1.0
        if i + 1 < n:
            ent[i + 1] = -1.0
        if i + far < n:
            ent[i + far] = -0.5
        if i - far >= 0:
            ent[i - far] = -0.5
        for j in sorted(ent):
            rows.append(i)
            cols.append(j)
            vals.append(ent[j])
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, np.asarray(rows) + 1, 1)
    return np.cumsum(rp), np.asarray(cols, np.int64) + 1, np.asarray(vals)


def test_sell_with_32bit_columns_only(ctx):
    """The SELL-64 products with columns too far for 16-bit offsets: y = A x
    bitwise the oracle's row sums (FP64 and sequential FP32), and CG, mixed CG
    and BiCGStab (their products read the same copy) solve to the tolerance."""
    import ofeli_amd as oa
    torch = _torch()
    rp, ci, a = far_coupled_system()
    A = gpu_csr(ctx, rp, ci, a)
    assert A.sell_info() == {"in_use": True, "stored": 130112, "col_bytes": 4}
    rng = np.random.default_rng(41)
    x = rng.standard_normal(A.n)
    y = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    oa.spmv(A, dev(x), y)
    assert np.array_equal(host(y), O.spmv(rp, ci, a, x))
    x32, a32 = x.astype(np.float32), a.astype(np.float32)
    y32 = torch.zeros(A.n, dtype=torch.float32, device="cuda")
    oa.spmv_f32(A, torch.as_tensor(x32, device="cuda"), y32)
    ref = np.zeros(A.n, np.float32)
    for i in range(A.n):  # sequential float32 row sums
        s = np.float32(0.0)
        for k in range(rp[i], rp[i + 1]):
            s = np.float32(s + np.float32(a32[k] * x32[ci[k] - 1]))
        ref[i] = s
    assert np.array_equal(y32.cpu().numpy(), ref)
    b = O.spmv(rp, ci, a, np.ones(A.n))  # solution 1
    for fn in (oa.cg, oa.cg_mixed, oa.bicgstab):
        xs = torch.zeros(A.n, dtype=torch.float64, device="cuda")
        info = fn(A, dev(b), xs, prec=oa.DIAG_PREC, max_it=2000, toler=1e-10)
        assert info.true_relres <= 2e-10, fn  # recursive-residual stop tests
        assert rel(host(xs), np.ones(A.n)) < 1e-8, fn


def test_spmv_column_compression_is_exact(ctx):
    """16-bit block-relative columns and 32-bit columns give the same bits."""
    import ofeli_amd as oa
    rp, ci, a, b = dc2d_system(60)
    A = gpu_csr(ctx, rp, ci, a)
    x = np.random.default_rng(9).standard_normal(A.n)
    ref = O.spmv(rp, ci, a, x)
    for enable in (False, True, False):
        A.compress_columns(enable)
        y = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
        oa.spmv(A, dev(x), y)
        assert np.array_equal(host(y), ref)
        xs = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
        info = oa.cg(A, dev(b), xs, max_it=3000, toler=1e-10)
        assert info.nb_it > 0


def test_spmv_add_bitexact(ctx):
    """MultAdd(x,y) and MultAdd(a,x,y) keep the reference's term order."""
    import ofeli_amd as oa
    rp, ci, a, b = dc2d_system(20)
    A = gpu_csr(ctx, rp, ci, a)
    rng = np.random.default_rng(2)
    x, y0 = rng.standard_normal(A.n), rng.standard_normal(A.n)
    y = dev(y0)
    oa.spmv_add(A, 1.0, dev(x), y)
    ref = y0.copy()
    for i in range(A.n):
        for k in range(rp[i], rp[i + 1]):
            ref[i] += a[k] * x[ci[k] - 1]
    assert np.array_equal(host(y), ref)
    y = dev(y0)
    oa.spmv_add(A, -0.3, dev(x), y)
    ref = y0.copy()
    for i in range(A.n):
        for k in range(rp[i], rp[i + 1]):
            ref[i] += -0.3 * a[k] * x[ci[k] - 1]
    assert np.array_equal(host(y), ref)


def test_spmv_long_rows_bitexact(ctx):
    """Rows longer than the LDS tile take the chunked path; sums stay sequential."""
    import ofeli_amd as oa
    n = 9000
    rng = np.random.default_rng(5)
    rows = []
    for i in range(n):
        if i in (3, 4500):
            cols = np.arange(n)  # two dense rows
        else:
            cols = np.unique(np.concatenate([[i], rng.integers(0, n, 6)]))
        rows.append(cols)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate(rows).astype(np.int64) + 1
    a = rng.standard_normal(rp[-1])
    A = gpu_csr(ctx, rp, ci, a)
    x = rng.standard_normal(n)
    y = _torch().zeros(n, dtype=_torch().float64, device="cuda")
    oa.spmv(A, dev(x), y)
    assert np.array_equal(host(y), O.spmv(rp, ci, a, x))


def test_prec_diag_and_ssor_bitexact(ctx):
    import ofeli_amd as oa
    rp, ci, a, b = dc2d_system(32)
    A = gpu_csr(ctx, rp, ci, a)
    r = np.random.default_rng(4).standard_normal(A.n)
    z = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
    oa.prec_apply(A, oa.DIAG_PREC, dev(r), z)
    diag = a[[np.searchsorted(ci[rp[i]:rp[i + 1]], i + 1) + rp[i] for i in range(A.n)]]
    assert np.array_equal(host(z), r * (1.0 / diag))
    oa.ssor_solve(A, dev(r), z)
    assert np.array_equal(host(z), O.ssor_solve(rp, ci, a, r))


def drop_upper(rp, ci, a, every=3):
    """The same matrix without the first upper entry of every `every`-th row:
    a structurally unsymmetric pattern (its lower partner has no transpose)."""
    keep = np.ones(len(ci), bool)
    for i in range(0, len(rp) - 1, every):
        up = [k for k in range(rp[i], rp[i + 1]) if ci[k] > i + 1]
        if up:
            keep[up[0]] = False
    nrp = np.concatenate([[0], np.cumsum([keep[rp[i]:rp[i + 1]].sum() for i in range(len(rp) - 1)])])
    return nrp.astype(rp.dtype), ci[keep], a[keep]


@pytest.mark.parametrize("unsym", [False, True])
def test_prec_dilu_bitexact(ctx, unsym):
    """DILU_PREC / ILU_PREC (the reference's ILU is its DILU): the GPU's
    level-scheduled factorization and sweeps give bitwise the oracle's
    SpMatrix::DILUSolve; new values refresh the pivots."""
    import ofeli_amd as oa
    rp, ci, a, b = dc2d_system(32)
    if unsym:
        rp, ci, a = drop_upper(rp, ci, a)
    A = gpu_csr(ctx, rp, ci, a)
    r = np.random.default_rng(5).standard_normal(A.n)
    z = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
    z2 = _torch().zeros_like(z)
    piv = O.dilu_factorize(rp, ci, a)
    oa.prec_apply(A, oa.DILU_PREC, dev(r), z)
    assert np.array_equal(host(z), O.dilu_solve(rp, ci, a, piv, r))
    oa.prec_apply(A, oa.ILU_PREC, dev(r), z2)
    assert _torch().equal(z, z2)
    a2 = a * (1.0 + 0.1 * np.cos(np.arange(len(a))))
    A.set_values(a2)
    oa.prec_apply(A, oa.DILU_PREC, dev(r), z)
    assert np.array_equal(host(z), O.dilu_solve(rp, ci, a2, O.dilu_factorize(rp, ci, a2), r))


def grid_system(shape, seed):
    """Laplacian of a structured grid (5- or 7-point) with random positive
    off-diagonal weights, rows sorted, 1-based columns: wide levels for a 3D
    grid, hundreds of one-wave levels for a long thin 2D strip."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    mats = [sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], (m, m)) for m in shape]
    L = mats[0]
    for M in mats[1:]:
        L = sp.kronsum(L, M)
    L = sp.csr_matrix(L)
    U = sp.triu(L, 1)
    U.data *= rng.uniform(0.5, 1.5, U.nnz)
    S = (U + U.T).tocsr()
    A = (S + sp.diags(-np.asarray(S.sum(axis=1)).ravel() + 0.1)).tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int64), A.indices.astype(np.int64) + 1, A.data.copy()


@pytest.mark.parametrize("shape", [(28, 28, 28), (400, 20)])
@pytest.mark.parametrize("unsym", [False, True])
def test_sweep_precs_bitexact_large(ctx, shape, unsym):
    """SSOR_PREC and DILU_PREC sweeps bitwise the oracle's on 21,952 rows of a
    3D grid (82 levels up to ~9 waves wide) and on a 400 x 20 strip (419
    levels of one wave each: the persistent workgroups take many chunks)."""
    import ofeli_amd as oa
    rp, ci, a = grid_system(shape, 11)
    if unsym:
        rp, ci, a = drop_upper(rp, ci, a)
    A = gpu_csr(ctx, rp, ci, a)
    r = np.random.default_rng(12).standard_normal(A.n)
    z = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
    oa.ssor_solve(A, dev(r), z)
    assert np.array_equal(host(z), O.ssor_solve(rp, ci, a, r))
    oa.prec_apply(A, oa.DILU_PREC, dev(r), z)
    assert np.array_equal(host(z), O.dilu_solve(rp, ci, a, O.dilu_factorize(rp, ci, a), r))
    for _ in range(3):  # repeated sweeps reuse the ticket counters
        oa.prec_apply(A, oa.DILU_PREC, dev(r), z)
    assert np.array_equal(host(z), O.dilu_solve(rp, ci, a, O.dilu_factorize(rp, ci, a), r))


def hub_system(seed):
    """A 60 x 40 grid Laplacian plus hub rows coupled to 30-45 random other
    rows (symmetric pattern, diagonally dominant): rows whose lower and upper
    parts are longer than the entries a sweep loads with its chunk."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    rp, ci, a = grid_system((60, 40), seed)
    n = len(rp) - 1
    A = sp.csr_matrix((a, ci - 1, rp), shape=(n, n)).tolil()
    for h in (7, 500, 1500, 2399):
        for j in rng.choice(n, size=int(rng.integers(30, 46)), replace=False):
            if j != h:
                w = -rng.uniform(0.01, 0.1)
                A[h, j] = w
                A[j, h] = w
    A = A.tocsr()
    A.setdiag(0.0)
    A.eliminate_zeros()
    A = (A + sp.diags(-np.asarray(A.sum(axis=1)).ravel() + 0.1)).tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int64), A.indices.astype(np.int64) + 1, A.data.copy()


@pytest.mark.parametrize("case", ["chain", "hubs"])
def test_sweep_precs_chains_and_long_rows(ctx, case):
    """A tridiagonal chain of 3,000 one-row levels (chunks of many levels, each
    row reading the previous one through LDS), and hub rows with 30-45
    entries on each side of the diagonal (entries loaded in batches after
    the first ones): SSOR and DILU sweeps bitwise the oracle's."""
    import ofeli_amd as oa
    rp, ci, a = grid_system((3000,), 13) if case == "chain" else hub_system(14)
    A = gpu_csr(ctx, rp, ci, a)
    r = np.random.default_rng(15).standard_normal(A.n)
    z = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
    oa.ssor_solve(A, dev(r), z)
    assert np.array_equal(host(z), O.ssor_solve(rp, ci, a, r))
    oa.prec_apply(A, oa.DILU_PREC, dev(r), z)
    assert np.array_equal(host(z), O.dilu_solve(rp, ci, a, O.dilu_factorize(rp, ci, a), r))


@pytest.mark.parametrize("prec", [1, 2, 3, 4, 5])
def test_cg_matches_oracle(ctx, prec):
    """CG: same return convention, iteration count within 1, and the GPU and
    oracle solutions agree to relative L2 <= 1e-9 at rtol 1e-10."""
    import ofeli_amd as oa
    rp, ci, a, b = dc2d_system(40)
    A = gpu_csr(ctx, rp, ci, a)
    ret, xo, reso = O.cg(rp, ci, a, b, prec=prec, max_it=5000, toler=1e-10)
    x = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
    info = oa.cg(A, dev(b), x, prec=prec, max_it=5000, toler=1e-10)
    assert ret > 0 and info.nb_it > 0
    assert abs(info.nb_it - ret) <= 1
    assert rel(host(x), xo) <= 1e-9
    assert info.true_relres <= 2e-10


def test_cg_deterministic_and_warm_start(ctx):
    import ofeli_amd as oa
    rp, ci, a, b = dc2d_system(40)
    A = gpu_csr(ctx, rp, ci, a)
    xs = []
    for _ in range(2):
        x = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
        oa.cg(A, dev(b), x, max_it=5000, toler=1e-9)
        xs.append(host(x))
    assert np.array_equal(xs[0], xs[1])
    # warm start from the solution: already converged -> 0 (CG.h:116-123)
    info = oa.cg(A, dev(b), dev(xs[0]), max_it=5000, toler=1e-6)
    assert info.nb_it == 0


def test_cg_not_converged_convention(ctx):
    import ofeli_amd as oa
    rp, ci, a, b = dc2d_system(40)
    A = gpu_csr(ctx, rp, ci, a)
    ret, _, _ = O.cg(rp, ci, a, b, max_it=7, toler=1e-12)
    x = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
    info = oa.cg(A, dev(b), x, max_it=7, toler=1e-12)
    assert ret == -8 and info.nb_it == -8 and info.iterations == 7


@pytest.mark.parametrize("prec", [1, 2, 3, 5])
def test_bicgstab_matches_oracle(ctx, prec):
    """BiCGStab's residual history is erratic, so rounding differences in the
    tree-summed dot products move the step that first crosses the tolerance by
    a few iterations.  The algorithm itself is pinned by the early iterates
    (equal to ~1e-12), the final solutions by relative L2 <= 1e-8."""
    import ofeli_amd as oa
    rp, ci, a, b = dc2d_system(32)
    A = gpu_csr(ctx, rp, ci, a)
    for k in (1, 2, 5):
        _, xo, _ = O.bicgstab(rp, ci, a, b, prec=prec, max_it=k, toler=1e-14)
        x = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
        info = oa.bicgstab(A, dev(b), x, prec=prec, max_it=k, toler=1e-14)
        assert info.nb_it == -(k + 1)
        assert rel(host(x), xo) <= 1e-11
    ret, xo, _ = O.bicgstab(rp, ci, a, b, prec=prec, max_it=3000, toler=1e-10)
    x = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
    info = oa.bicgstab(A, dev(b), x, prec=prec, max_it=3000, toler=1e-10)
    assert ret > 0 and info.nb_it > 0
    assert abs(info.nb_it - ret) <= max(2, ret // 8)
    assert rel(host(x), xo) <= 1e-8
    assert info.true_relres <= 1e-9


@pytest.mark.parametrize("prec", [1, 2, 3, 5])
def test_gmres_matches_oracle(ctx, prec):
    import ofeli_amd as oa
    rp, ci, a, b = dc2d_system(24)
    A = gpu_csr(ctx, rp, ci, a)
    ret, xo, _ = O.gmres(rp, ci, a, b, 20, prec=prec, max_it=3000, toler=1e-10)
    x = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
    info = oa.gmres(A, dev(b), x, 20, prec=prec, max_it=3000, toler=1e-10)
    assert ret > 0 and info.nb_it > 0
    assert abs(info.nb_it - ret) <= 2
    assert rel(host(x), xo) <= 1e-8


@pytest.mark.parametrize("prec", ["DIAG_PREC", "IDENT_PREC", "SSOR_PREC", "DILU_PREC"])
def test_mixed_cg_fp64_verified(ctx, prec):
    """Mixed precision: stops on the true FP64 residual; its error against
    the FP64 oracle solution stays at the solver-accuracy level (<= 1e-7)."""
    import ofeli_amd as oa
    rp, ci, a, b = dc2d_system(64)
    A = gpu_csr(ctx, rp, ci, a)
    ret, xo, _ = O.cg(rp, ci, a, b, max_it=10000, toler=1e-12)
    x = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
    info = oa.cg_mixed(A, dev(b), x, prec=getattr(oa, prec), max_it=10000, toler=1e-8)
    assert info.nb_it > 0 and info.fallback == 0
    assert info.true_relres <= 1e-8
    # the inner FP32 iteration keeps the preconditioner's effect: about the
    # FP64 iteration count of the same preconditioner (oracle, rtol 1e-8)
    r8, _, _ = O.cg(rp, ci, a, b, prec=getattr(O, prec.split("_")[0]), max_it=10000, toler=1e-8)
    assert r8 > 0 and info.iterations <= 1.3 * r8 + 10, (info.iterations, r8)
    # error against the oracle's rtol-1e-12 solution: FP64 CG at rtol 1e-8
    # sits at the same level, so 1e-7 leaves room only for rounding
    e = rel(host(x), xo)
    print("mixed %s: err vs oracle %.2e" % (prec, e))
    assert e <= 1e-7


def sym_tridiag_system(n=6000, seed=11):
    """Symmetric positive definite tridiagonal system (3 diagonals: the FP32
    diagonal copy with 4 slots per row)."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    off = -rng.uniform(0.5, 1.0, n - 1)
    d = 2.05 + rng.uniform(0.0, 0.1, n)
    M = sp.diags([off, d, off], [-1, 0, 1], format="csr")
    M.sort_indices()
    b = rng.standard_normal(n)
    return M.indptr.astype(np.int64), M.indices.astype(np.int64) + 1, M.data, b


@pytest.mark.parametrize("system", ["dc2d", "tridiag", "tiny"])
@pytest.mark.parametrize("prec", ["DIAG_PREC", "IDENT_PREC"])
def test_mixed_cg_one_and_two_kernel_inner_forms(ctx, prec, system):
    """ofh_ctx_mixed_form: the single-reduction inner iteration (one kernel per
    inner iteration, z = r on the FP32 diagonal copy; 8 and 4 slots per row
    here; "tiny": 63 rows, less than one workgroup's rows) and the two-kernel
    form both stop on the true FP64 residual, at the oracle's rtol-1e-12
    solution to 1e-7 with about the FP64 iteration count of the same
    preconditioner.  The per-kind kernel timings show which form
    ran.  The max_it convention (-(max_it+1), max_it inner iterations) holds
    for both."""
    import ofeli_amd as oa
    rp, ci, a, b = (dc2d_system(64) if system == "dc2d" else
                    sym_tridiag_system(63) if system == "tiny" else sym_tridiag_system())
    A = gpu_csr(ctx, rp, ci, a)
    assert A.dia_info()["fp32"] and A.dia_info()["row_slots"] == (8 if system == "dc2d" else 4)
    _, xo, _ = O.cg(rp, ci, a, b, max_it=20000, toler=1e-12)
    r8, _, _ = O.cg(rp, ci, a, b, prec=getattr(O, prec.split("_")[0]), max_it=20000, toler=1e-8)
    assert r8 > 0
    try:
        for form in (1, 2):
            ctx.mixed_form(form)
            ctx.profile(True)
            x = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
            info = oa.cg_mixed(A, dev(b), x, prec=getattr(oa, prec), max_it=20000, toler=1e-8)
            n_one = len(ctx.profile_read(oa.PROF_MX_ITER))
            n_two = len(ctx.profile_read(oa.PROF_MX_UPD))
            ctx.profile(False)
            assert (n_one > 0 and n_two == 0) if form == 2 else (n_one == 0 and n_two > 0)
            assert info.nb_it > 0 and info.fallback == 0, info.as_dict()
            assert info.true_relres <= 1e-8
            assert info.iterations <= 1.3 * r8 + 10, (info.iterations, r8)
            e = rel(host(x), xo)
            print("%s %s form %d: %d inner it (FP64 %d), %d replacements, err vs oracle %.2e"
                  % (system, prec, form, info.iterations, r8, info.outer, e))
            assert e <= 1e-7
            x = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
            info = oa.cg_mixed(A, dev(b), x, prec=getattr(oa, prec), max_it=7, toler=1e-12)
            assert info.nb_it == -8 and info.iterations == 7
    finally:
        ctx.profile(False)
        ctx.mixed_form(0)


def test_mixed_cg_one_kernel_form_needs_symmetric_values(ctx):
    """The one-kernel inner iteration forms (p, Ap) through (p_old, A r) =
    (A p_old, r), so it runs only where the values are symmetric bit for bit:
    one value moved by one ulp makes even a requested one-kernel form take the
    two-kernel iteration (no k_mx_cg1 launch), which still meets the FP64
    tolerance."""
    import ofeli_amd as oa
    rp, ci, a, b = sym_tridiag_system()
    a = a.copy()
    a[5] = np.nextafter(a[5], -np.inf)
    A = gpu_csr(ctx, rp, ci, a)
    assert A.dia_info()["fp32"] and not A.dia_info()["symmetric"]
    try:
        ctx.mixed_form(2)
        ctx.profile(True)
        x = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
        info = oa.cg_mixed(A, dev(b), x, prec=oa.DIAG_PREC, max_it=20000, toler=1e-8)
        assert len(ctx.profile_read(oa.PROF_MX_ITER)) == 0
        assert len(ctx.profile_read(oa.PROF_MX_UPD)) > 0
    finally:
        ctx.profile(False)
        ctx.mixed_form(0)
    assert info.nb_it > 0 and info.fallback == 0 and info.true_relres <= 1e-8


def test_mixed_cg_many_partials(ctx):
    """Mixed CG on 2,000,000 rows: more update partials than the product
    kernel prefetches (it folds them after the product instead); the solve
    still ends on the true FP64 residual with the known solution."""
    import scipy.sparse as sp
    import ofeli_amd as oa
    n = 2_000_000
    d = 4.0 + (np.arange(n) % 7) * 0.25
    M = sp.diags([-np.ones(n - 1), d, -np.ones(n - 1)], [-1, 0, 1], format="csr")
    M.sort_indices()
    rp, ci, a = M.indptr.astype(np.int64), M.indices.astype(np.int64) + 1, M.data
    A = gpu_csr(ctx, rp, ci, a)
    b = M @ np.ones(n)
    x = _torch().zeros(n, dtype=_torch().float64, device="cuda")
    info = oa.cg_mixed(A, dev(b), x, max_it=1000, toler=1e-10)
    assert info.nb_it > 0 and info.fallback == 0
    assert info.true_relres <= 1e-10
    assert rel(host(x), np.ones(n)) < 1e-9


def test_zero_pivot_error(ctx):
    import ofeli_amd as oa
    rp, ci, a, b = dc2d_system(8)
    a = a.copy()
    d = np.searchsorted(ci[rp[5]:rp[6]], 6) + rp[5]
    a[d] = 0.0
    A = gpu_csr(ctx, rp, ci, a)
    x = _torch().zeros(A.n, dtype=_torch().float64, device="cuda")
    with pytest.raises(oa.OfhError) as e:
        oa.cg(A, dev(b), x)
    assert e.value.code == -3
    with pytest.raises(oa.OfhError) as e:
        oa.cg(A, dev(b), x, prec=oa.DILU_PREC)
    assert e.value.code == -3


@pytest.mark.parametrize("n", [1, 1000, 262_145, 3_000_001])
def test_blas1_entry_points(ctx, n):
    """ofh_dot_f64 / ofh_norm2_f64 (Vect_impl.h:2218-2225, 660-668) are
    fixed-order trees, not the reference's sequential sum: within 1e-13
    relative of the oracle's sequential sum, bitwise equal over repeated
    calls; ofh_axpy_f64 (Vect += a*x) is one rounded multiply-add per entry
    without contraction: bitwise y + a*x."""
    import ofeli_amd as oa
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    y = rng.standard_normal(n) + 0.5
    dx, dy = dev(x), dev(y)
    d = oa.dot(ctx, dx, dy)
    do = O.lib().orc_dot(n, x, y)
    assert abs(d - do) <= 1e-13 * np.sum(np.abs(x * y))
    assert oa.dot(ctx, dx, dy) == d
    nr = oa.norm2(ctx, dx)
    assert abs(nr - O.lib().orc_norm2(n, x)) <= 1e-13 * nr
    a = 0.3125 - 1e-3
    oa.axpy(ctx, a, dx, dy)
    assert np.array_equal(host(dy), y + a * x)


def test_blas1_edge_sizes(ctx):
    import ofeli_amd as oa
    torch = _torch()
    z = torch.zeros(0, dtype=torch.float64, device="cuda")
    assert oa.dot(ctx, z, z) == 0.0 and oa.norm2(ctx, z) == 0.0
    oa.axpy(ctx, 2.0, z, z)
