"""ofh_cg_noalloc, the CG entry point that allocates nothing, through the
C-ABI against ofh_cg (the entry point it mirrors) on the same systems: two
calls in a row, the second on a smaller system (the scratch array keeps the
first call's contents), IDENT and DIAG, plus its size limit and its
zero-pivot error.  The two entry points run the same iteration; their dot
products are folded over different partitions (ofh_cg reads the SELL-64
copy), so iterates agree to rounding: iteration counts within 1, solutions
within 1e-10 relative L2 of each other, both meeting the stop test."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def system(ctx, nx, seed):
    import torch
    import ofeli_amd as oa
    coord, code, elem = oa.api.mesh2d(nx, nx)
    rng = np.random.default_rng(seed)
    P = oa.AssemblyPlan(ctx, coord, elem, code)
    b = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
    P.build(oa.DC2DT3, coef=torch.as_tensor(np.exp(0.75 * rng.standard_normal(elem.shape[0])),
                                            device="cuda"),
            f=torch.as_tensor(1.0 + rng.random(coord.shape[0]), device="cuda"), matrix=True, b=b)
    return P, b


def noalloc(A, b, x, prec, max_it=20000, toler=1e-9):
    import ofeli_amd as oa
    info = oa.SolveInfo()
    oa.check(oa.lib().ofh_cg_noalloc(A.h, prec, C.c_void_p(b.data_ptr()),
                                     C.c_void_p(x.data_ptr()), max_it, toler, C.byref(info)))
    return info


def rel(a, b):
    return float((a - b).norm() / b.norm())


@pytest.mark.parametrize("prec_name", ["DIAG_PREC", "IDENT_PREC"])
def test_noalloc_cg_equals_cg_on_two_calls(ctx, prec_name):
    import torch
    import ofeli_amd as oa
    prec = getattr(oa, prec_name)
    systems = [system(ctx, 300, 1), system(ctx, 64, 2)]  # 89,401 then 3,969 rows
    xs = [torch.zeros_like(b) for _, b in systems]
    # the HIP runtime may set up its own launch resources at a process's
    # first launches of these kernels; after one call, calls take no memory
    torch.cuda.synchronize()
    f_first = torch.cuda.mem_get_info()[0]
    noalloc(systems[1][0].A, systems[1][1], torch.zeros_like(systems[1][1]), prec)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    print("device memory taken by the first call (runtime set-up): %d B" % (f_first - free0))
    infos = [noalloc(P.A, b, x, prec) for (P, b), x in zip(systems, xs)]
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0  # no device memory taken
    for (P, b), x, info in zip(systems, xs, infos):
        xr = torch.zeros_like(b)
        ir = oa.cg(P.A, b, xr, prec=prec, max_it=20000, toler=1e-9)
        print("%s n=%d: noalloc %d it (relres %.2e), cg %d it, rel diff %.2e"
              % (prec_name, P.nb_eq, info.nb_it, info.true_relres, ir.nb_it, rel(x, xr)))
        assert info.nb_it > 0 and ir.nb_it > 0
        assert abs(info.nb_it - ir.nb_it) <= 1
        assert info.relres <= 1e-9 and info.true_relres <= 1.2e-9
        assert rel(x, xr) <= 1e-10
    # a repeat of the first system after the smaller one gives the same result
    x2 = torch.zeros_like(systems[0][1])
    info2 = noalloc(systems[0][0].A, systems[0][1], x2, prec)
    torch.cuda.synchronize()
    assert info2.nb_it == infos[0].nb_it and torch.equal(x2, xs[0])
    for P, _ in systems:
        P.close()


def test_noalloc_limits_and_errors(ctx):
    import torch
    import ofeli_amd as oa
    # beyond the scratch maximum of 262,144 rows: a clear error, no launch
    n = 262_145
    rp = np.arange(n + 1, dtype=np.int32)
    A = oa.CsrMatrix(ctx, n, rp, np.arange(n, dtype=np.int32), np.full(n, 2.0))
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    with pytest.raises(oa.OfhError) as e:
        noalloc(A, b, x, oa.DIAG_PREC)
    assert e.value.code == -1 and "262144" in str(e.value)
    A.close()
    # unsupported preconditioner
    P, b = system(ctx, 16, 3)
    x = torch.zeros_like(b)
    with pytest.raises(oa.OfhError) as e:
        noalloc(P.A, b, x, oa.SSOR_PREC)
    assert e.value.code == -7
    # a zero diagonal entry: the reference's Prec::inv_diag error
    a = P.A.values().copy()
    rp, col = P.A.pattern()
    d = rp[5] + int(np.searchsorted(col[rp[5]:rp[6]], 5))
    a[d] = 0.0
    P.A.set_values(a)
    with pytest.raises(oa.OfhError) as e:
        noalloc(P.A, b, x, oa.DIAG_PREC)
    assert e.value.code == -3
    # IDENT needs no pivots; max_it = 0 reports no convergence from zero
    info = noalloc(P.A, b, x, oa.IDENT_PREC, max_it=0)
    assert info.nb_it == -1
    P.close()
