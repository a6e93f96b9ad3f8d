"""The 3D variable-coefficient Poisson path end to end on the GPU, on the
product's own generators: Kuhn P1 tetrahedra (DC3DT4 Diffusion with a
per-element k = exp(g), g the 64-mode Gaussian field of correlation 0.1,
f = 1, Dirichlet 0), assembled on the GPU and solved by FP64 CG+DIAG_PREC
(CG.h:99-157) and by the mixed FP32/FP64 CG, against the oracle.

Bars: pattern and values bitwise; FP64 CG iterations within 1 of the
oracle's and solution within 1e-9 relative L2 of the oracle's at the same
tolerance (the two follow the same iteration; dot products are summed in a
different order); mixed CG true FP64 residual <= 1e-8 and error against the
oracle's rtol-1e-12 solution <= 1e-7 relative L2 (FP64 CG at rtol 1e-8 is
itself 7e-10 / 9e-10 from it on these two meshes).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [(20, 20, 20), (23, 17, 13)]


def problem(dims, seed_offset=1):
    from ofeli_amd import api, workloads as W
    coord, code, elem = api.mesh3d_kuhn(*dims)
    rng = W.SplitMix64(W.SEED_BASE + seed_offset)
    k = np.exp(W.fourier_field(rng, W.centroids(coord, elem)))
    return coord, code, elem, k


@pytest.fixture(scope="module", params=CASES, ids=lambda d: "x".join(map(str, d)))
def system(request, ctx):
    import torch
    import ofeli_amd as oa
    dims = request.param
    coord, code, elem, k = problem(dims)
    print("poisson3d case", dims, flush=True)
    m = O.mesh3d_kuhn(*dims)
    assert np.array_equal(coord, m.coord) and np.array_equal(code, m.code)
    assert np.array_equal(elem.astype(np.int64), m.elem)
    dof, neq = O.number_equations(m.code)
    rp, ci = O.node_graph(m.elem, 1, dof, neq)
    a, b = O.assemble(O.LAPLACE3DT4, m, dof, neq, rp, ci, f=np.ones(m.nb_nodes), coef=k)
    P = oa.AssemblyPlan(ctx, coord, elem, code)
    bg = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
    P.build(oa.LAPLACE3DT4, coef=torch.as_tensor(k, device="cuda"),
            f=torch.ones(coord.shape[0], dtype=torch.float64, device="cuda"), matrix=True, b=bg)
    ret12, x12, _ = O.cg(rp, ci, a, b, max_it=20000, toler=1e-12)
    assert ret12 > 0
    yield dict(P=P, b=bg, rp=rp, ci=ci, a=a, bo=b, x12=x12, dims=dims)
    P.close()


def rel(x, y):
    return np.linalg.norm(x - y) / np.linalg.norm(y)


def test_pattern_values_and_rhs_bitwise(system):
    P = system["P"]
    grp, gcol = P.A.pattern()
    assert np.array_equal(grp.astype(np.int64), system["rp"])
    assert np.array_equal(gcol.astype(np.int64) + 1, system["ci"])
    assert np.array_equal(P.A.values(), system["a"])
    assert np.array_equal(system["b"].cpu().numpy(), system["bo"])


def test_products_stream_the_sell_copy_with_16bit_columns(system):
    info = system["P"].A.sell_info()
    assert info["in_use"] and info["col_bytes"] == 2
    assert system["P"].A.nnz <= info["stored"] <= 1.25 * system["P"].A.nnz


def test_products_read_the_diagonal_copies(system):
    """The Kuhn numbering of the free nodes is a regular grid: every entry lies
    on one of 15 diagonals, so ofh_spmv_f32 and the mixed CG's inner product
    read the FP32 diagonal copy and ofh_spmv_f64 and FP64 CG the FP64 one (16
    slots per row), bitwise the CSR row sums.  The assembled values are
    symmetric bit for bit (the element matrices are, and every entry sums
    its elements in ascending order), so the products read the symmetric
    form (8 slots per row)."""
    import torch
    import ofeli_amd as oa
    A = system["P"].A
    assert A.dia_info() == {"fp32": True, "fp64": True, "diagonals": 15, "row_slots": 16,
                            "symmetric": True}
    x64 = np.random.default_rng(6).standard_normal(A.n)
    y64 = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    oa.spmv(A, torch.as_tensor(x64, device="cuda"), y64)
    assert np.array_equal(y64.cpu().numpy(), O.spmv(system["rp"], system["ci"], system["a"], x64))
    rng = np.random.default_rng(5)
    x32 = rng.standard_normal(A.n).astype(np.float32)
    y32 = torch.zeros(A.n, dtype=torch.float32, device="cuda")
    oa.spmv_f32(A, torch.as_tensor(x32, device="cuda"), y32)
    rp, ci, a32 = system["rp"], system["ci"], system["a"].astype(np.float32)
    ref = np.zeros(A.n, np.float32)
    prod = a32 * x32[ci - 1]
    for j in range(int(np.diff(rp).max())):  # column j of every row, in order
        rows = np.nonzero(np.diff(rp) > j)[0]
        ref[rows] = ref[rows] + prod[rp[rows] + j]
    assert np.array_equal(y32.cpu().numpy(), ref)


def test_fp64_cg_diag_matches_oracle(system):
    import torch
    import ofeli_amd as oa
    P = system["P"]
    ret, xo, _ = O.cg(system["rp"], system["ci"], system["a"], system["bo"], max_it=20000,
                      toler=1e-8)
    x = torch.zeros_like(system["b"])
    info = oa.cg(P.A, system["b"], x, prec=oa.DIAG_PREC, max_it=20000, toler=1e-8)
    assert ret > 0 and info.nb_it > 0
    assert abs(info.nb_it - ret) <= 1, (info.nb_it, ret)
    e = rel(x.cpu().numpy(), xo)
    print("fp64 cg: it %d (oracle %d), rel err vs oracle %.2e" % (info.nb_it, ret, e))
    assert e <= 1e-9
    assert info.true_relres <= 1.01e-8
    assert rel(x.cpu().numpy(), system["x12"]) <= 1e-7


def test_mixed_cg_matches_oracle(system):
    import torch
    import ofeli_amd as oa
    P = system["P"]
    x = torch.zeros_like(system["b"])
    info = oa.cg_mixed(P.A, system["b"], x, prec=oa.DIAG_PREC, max_it=20000, toler=1e-8)
    assert info.nb_it > 0 and info.fallback == 0, info.as_dict()
    assert info.true_relres <= 1e-8
    e = rel(x.cpu().numpy(), system["x12"])
    print("mixed cg: %d inner it, %d replacements, true relres %.2e, err vs oracle %.2e"
          % (info.iterations, info.outer, info.true_relres, e))
    assert e <= 1e-7


@pytest.mark.parametrize("form", [1, 2])
def test_mixed_cg_inner_forms(system, ctx, form):
    """Both forms of the inner iteration on the 15-diagonal copy (16 slots per
    row): two kernels, and the single-reduction one kernel.  Same bars as the
    default solve."""
    import torch
    import ofeli_amd as oa
    P = system["P"]
    try:
        ctx.mixed_form(form)
        x = torch.zeros_like(system["b"])
        info = oa.cg_mixed(P.A, system["b"], x, prec=oa.DIAG_PREC, max_it=20000, toler=1e-8)
    finally:
        ctx.mixed_form(0)
    assert info.nb_it > 0 and info.fallback == 0, info.as_dict()
    assert info.true_relres <= 1e-8
    e = rel(x.cpu().numpy(), system["x12"])
    print("mixed cg form %d: %d inner it, %d replacements, err vs oracle %.2e"
          % (form, info.iterations, info.outer, e))
    assert e <= 1e-7


def test_mixed_cg_ident_prec(system):
    import torch
    import ofeli_amd as oa
    P = system["P"]
    x = torch.zeros_like(system["b"])
    info = oa.cg_mixed(P.A, system["b"], x, prec=oa.IDENT_PREC, max_it=20000, toler=1e-8)
    assert info.nb_it > 0 and info.true_relres <= 1e-8
    assert rel(x.cpu().numpy(), system["x12"]) <= 1e-7
