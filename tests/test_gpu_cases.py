"""Load cases of one cantilever (Elas2DQ4 plane stress): right-hand sides
from the GPU traction assembly, solved one after another with the start
vector projected on the previous solutions, against per-case oracle solves
of the oracle-assembled system."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_cantilever_load_cases_match_oracle(ctx):
    import torch
    import ofeli_amd as oa
    from ofeli_amd import workloads as W
    d = W.config4(32, 8, ncases=12)
    P = oa.AssemblyPlan(ctx, d["coord"], d["elem"], d["code"])
    A = P.A
    b0 = torch.zeros(A.n, dtype=torch.float64, device="cuda")
    E = torch.as_tensor(d["coef"], device="cuda")
    P.build(oa.ELAS2DQ4, coef=E, prm=d["prm"], matrix=True, b=b0)
    # oracle system
    m = O.Mesh(d["coord"], d["code"], d["elem"].astype(np.int64), 2)
    dof, neq = O.number_equations(d["code"])
    rp, ci = O.node_graph(m.elem, 2, dof, neq)
    a, _ = O.assemble(O.ELAS2DQ4, m, dof, neq, rp, ci, coef=d["coef"], prm=list(d["prm"]))
    assert np.array_equal(A.values(), a)
    basis_gpu, basis_orc = [], []
    for tr in d["basis"]:
        bk = torch.zeros_like(b0)
        P.traction(d["sides"], torch.as_tensor(tr, device="cuda"), bk)
        bo = O.traction_rhs(m, d["sides"].astype(np.int64), tr.ravel(), dof, np.zeros(neq))
        assert np.array_equal(bk.cpu().numpy(), bo)
        basis_gpu.append(bk)
        basis_orc.append(bo)
    G = oa.ProjectionGuess(A, m=16, minres=True)
    x0 = torch.empty_like(b0)
    its, errs = [], []
    for c in range(d["alpha"].shape[0]):
        al = d["alpha"][c]
        b = sum(float(al[k]) * basis_gpu[k] for k in range(8))
        bo = sum(al[k] * basis_orc[k] for k in range(8))
        info = oa.solve_with_projection(A, b, x0, G, toler=1e-8, tighten=1e-3)
        its.append(info.iterations if info.nb_it != 0 else 0)
        assert info.true_relres <= 1e-8
        ret, xo, _ = O.cg(rp, ci, a, bo, max_it=20000, toler=1e-12)
        assert ret > 0
        e = np.linalg.norm(x0.cpu().numpy() - xo) / np.linalg.norm(xo)
        errs.append(e)
        assert e <= 1e-7, (c, e)
    print("load cases: max err vs oracle %.2e" % max(errs))
    # the first 8 cases span every case: the later ones need no iterations
    assert min(its[:8]) > 0 and max(its[8:]) == 0 and G.size() == 8, its


def test_mixed_cg_converges_on_an_ill_conditioned_system(ctx):
    """k jumps by 1e8 across the domain: the FP32 inner iteration cannot
    resolve the slow modes; the mixed solver must still return a solution
    meeting the FP64 stop test (through its FP64 fallback if needed)."""
    import torch
    import ofeli_amd as oa
    coord, code, elem = oa.api.mesh2d(96, 96)
    cx = coord[elem - 1].mean(axis=1)[:, 0]
    k = np.where(cx < 0.5, 1e-4, 1e4)
    P = oa.AssemblyPlan(ctx, coord, elem, code)
    b = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
    P.build(oa.DC2DT3, coef=torch.as_tensor(k, device="cuda"),
            f=torch.ones(coord.shape[0], dtype=torch.float64, device="cuda"), matrix=True, b=b)
    x64 = torch.zeros_like(b)
    i64 = oa.cg(P.A, b, x64, max_it=100000, toler=1e-8)
    xt = torch.zeros_like(b)  # the judge: FP64 CG far below the tolerance
    it_ = oa.cg(P.A, b, xt, max_it=200000, toler=1e-13)
    x = torch.zeros_like(b)
    info = oa.cg_mixed(P.A, b, x, max_it=100000, toler=1e-8)
    assert i64.nb_it > 0 and it_.nb_it > 0
    assert info.nb_it > 0 and info.true_relres <= 1e-8, info.as_dict()
    assert info.iterations < 6 * i64.iterations, (info.iterations, i64.iterations)
    nt = torch.linalg.norm(xt)
    e64 = float(torch.linalg.norm(x64 - xt) / nt)
    emx = float(torch.linalg.norm(x - xt) / nt)
    print("ill-conditioned: err FP64@1e-8 %.2e, mixed %.2e (fallback %d)"
          % (e64, emx, info.fallback))
    # the mixed solution is as accurate as the FP64 solve of the same tolerance
    assert emx <= max(1e-7, 10 * e64)


@pytest.mark.parametrize("prec", ["DIAG_PREC", "SSOR_PREC"])
def test_mixed_cg_on_a_small_cantilever(ctx, prec):
    """Mixed CG on a Q4 plane-stress cantilever (48 x 12 quads): the FP32
    inner iteration (q carried by recurrence, recomputed as A p after each
    FP64 replacement) reaches the FP64 stop test without the FP64 fallback,
    and its solution is within 1e-7 of the oracle's rtol-1e-12 solution.
    (From 64 x 16 quads on, with DIAG_PREC, FP32 no longer resolves the
    slowest modes and the solver finishes in FP64 CG: tested by the
    ill-conditioned case above.)"""
    import torch
    import ofeli_amd as oa
    from ofeli_amd import workloads as W
    d = W.config4(48, 12, ncases=1)
    P = oa.AssemblyPlan(ctx, d["coord"], d["elem"], d["code"])
    b = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
    P.build(oa.ELAS2DQ4, coef=torch.as_tensor(d["coef"], device="cuda"), prm=d["prm"], matrix=True,
            b=b)
    tr = sum(float(d["alpha"][0][k]) * torch.as_tensor(d["basis"][k], device="cuda").reshape(-1, 2)
             for k in range(8))
    P.traction(d["sides"], tr.contiguous(), b)
    rp, col = P.A.pattern()
    a = P.A.values()
    ret, xo, _ = O.cg(rp.astype(np.int64), col.astype(np.int64) + 1, a, b.cpu().numpy(),
                      prec=O.DIAG, max_it=100000, toler=1e-12)
    assert ret > 0
    x = torch.zeros_like(b)
    info = oa.cg_mixed(P.A, b, x, prec=getattr(oa, prec), max_it=100000, toler=1e-8)
    e = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
    print("cantilever mixed %s: %d inner it, %d repl, fallback %d, err %.2e"
          % (prec, info.iterations, info.outer, info.fallback, e))
    assert info.nb_it > 0 and info.true_relres <= 1e-8
    assert info.fallback == 0
    assert e <= 1e-7
