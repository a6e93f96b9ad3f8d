"""Benchmark of the sparse linear-solve path on one MI355X.

Workload: the 3D variable-coefficient Poisson problem (unit cube, 96^3 cells,
6 Kuhn tets per cell: 857,375 free DOFs, 12,645,163 nonzeros; k_e lognormal
from a 64-mode Gaussian field; f = 1; Dirichlet 0), assembled on the GPU and
solved to rtol 1e-8 with CG + DIAG_PREC.  One step = one solve from x0 = 0.
`value` is the wall time per solve of the mixed-precision CG (FP32 inner,
FP64 residual replacement, FP64-verified); the plain FP64 CG time, the
speedup and the SpMV roofline are reported beside it.

`--config 2` benches the batched workload instead: 512 uneven 2D steady
heat systems (DC2DT3, n x n cells with n uniform in [16, 96], 1,826,484 rows
in all) solved in one persistent launch (FP64 and mixed), against a loop of
per-system FP64 CG solves.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 96] [--config 1|2]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time to rtol 1e-8 per solve, mixed+learned vs plain FP64 CG; SpMV HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=96)
    ap.add_argument("--toler", type=float, default=1e-8)
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--cold-reps", type=int, default=20,
                    help="SpMV launches timed after a 1 GiB cache-evicting write each")
    ap.add_argument("--replica", type=int, default=128,
                    help="config 1: n of the cube replica (n^3 cells) for the cold SpMV "
                         "roofline beyond the Infinity Cache; 0 skips it")
    ap.add_argument("--cpu-reps", type=int, default=5, help="config 1: full oracle solves timed")
    ap.add_argument("--learned-seeds", type=int, default=16,
                    help="config 1: k-fields the learned start is trained on (0 skips the leg)")
    ap.add_argument("--learned-epochs", type=int, default=8)
    ap.add_argument("--learned-layers", type=int, default=4)
    ap.add_argument("--learned-objective", default="residual", choices=("energy", "residual"))
    ap.add_argument("--cpu-iters", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep-precs", action="store_true",
                    help="config 1: also time CG with the sweep preconditioners (DILU, SSOR)")
    ap.add_argument("--krylov", action="store_true",
                    help="config 1: also time BiCG, BiCGStab and CGS (DIAG_PREC)")
    ap.add_argument("--mesh-io", action="store_true",
                    help="config 1: also time the .m mesh writer/reader on a 1024^2 P1 mesh and "
                         "the GPU plan + assembly of the mesh read back")
    ap.add_argument("--config", type=int, default=1, choices=(1, 2, 3, 4))
    ap.add_argument("--nx4", type=int, default=2048, help="config 4 cells along the beam")
    ap.add_argument("--ny4", type=int, default=256, help="config 4 cells across the beam")
    ap.add_argument("--cases", type=int, default=256, help="config 4 load cases")
    ap.add_argument("--base-cases", type=int, default=8,
                    help="config 4 cases solved by the per-case baselines")
    ap.add_argument("--train-cases", type=int, default=32)
    ap.add_argument("--ssor-cases", type=int, default=1,
                    help="config 4 cases solved by FP64 CG+SSOR_PREC from zero (about 70 s each)")
    ap.add_argument("--ssor-proj-cases", type=int, default=8,
                    help="config 4 cases after the first ones solved by FP64 CG+SSOR_PREC (and "
                         "CG+DIAG) from the projected start")
    ap.add_argument("--n3", type=int, default=512, help="config 3 grid (n x n cells)")
    ap.add_argument("--t-steps", type=int, default=2000, help="config 3 time steps")
    ap.add_argument("--train-steps", type=int, default=200, help="config 3 steps used for training")
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--proj-m", type=int, default=16, help="config 3 projection basis size")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the sharding design is written up beside the code); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.config == 2:
        return bench_batched(args)
    if args.config == 3:
        return bench_transient(args)
    if args.config == 4:
        return bench_cantilever(args)

    return bench_poisson3d(args)


# ---------------------------------------------------------------- config 1
def b_spmv64(n, nnz):
    """Algorithmic bytes of y = A x, FP64 CSR with int32 indices (B_spmv64):
    values and columns once, row pointers, x once, y written."""
    return 12 * nnz + 4 * (n + 1) + 16 * n


def b_spmv32(n, nnz):
    """B_spmv32: the same with FP32 values and vectors."""
    return 8 * nnz + 4 * (n + 1) + 8 * n


def kernel_bytes(n, nnz):
    """Algorithmic bytes per launch of each kernel of the CG iterations
    (FP64: B_spmv64 + 88 n per iteration; mixed: B_spmv32 + 36 n)."""
    return {
        # mixed, two kernels: the product of z (the FP32 CSR SpMV bytes; it
        # also reads p_old, q_old and writes p_new beside q, which B_cg32_iter
        # counts in the iteration's 44n vector term), and the update
        # on the scaled system D^-1/2 A D^-1/2 (reads x, p, r, q; writes x, r)
        "k_mx_dirspmv": b_spmv32(n, nnz),
        "k_mx_upd": 24 * n,
        # FP64, three kernels: q = A p, update (reads x, p, r, q, 1/a_ii;
        # writes x, r), direction (reads r, 1/a_ii, p; writes p)
        "k_cg_spmv": b_spmv64(n, nnz),
        "k_cg_upd": 56 * n,
        "k_cg_dir": 32 * n,
    }


def working(ms):
    """Launch durations of the launches that did work: the host enqueues
    launches past the solver's stop points, which return at once."""
    ms = np.asarray(ms)
    if ms.size == 0:
        return ms
    return ms[ms >= 0.5 * np.median(ms)]


def dia_row_bytes(dia, value_bytes):
    """Matrix bytes per row of the diagonal copy the products read: row_slots
    values (the last slot the mask), or the symmetric form's row_slots / 2
    values and a 16-bit mask."""
    if dia["symmetric"]:
        return value_bytes * dia["row_slots"] // 2 + 2
    return value_bytes * dia["row_slots"]


def dia_kernel(dia, typ):
    """Name of the standalone diagonal-copy SpMV kernel as rocprofv3 lists it
    (compacted) for the copy in use."""
    return "k_spmv_dia<%d,%s,%s>" % (dia["row_slots"], typ, "true" if dia["symmetric"] else "false")


def dia_layout(dia, what):
    if dia["symmetric"]:
        return ("%s: symmetric form of the diagonal copy, %d value slots per row (diagonal and "
                "upper values; a lower value read from its mirror) + a 16-bit mask, chunks of 64 "
                "rows stored slot-major" % (what, dia["row_slots"] // 2))
    return ("%s: diagonal copy, %d slots per row, chunks of 64 rows stored slot-major"
            % (what, dia["row_slots"]))


def pmc_traffic(kernel, n):
    """HBM bytes per launch of `kernel` on an n-row system from the committed
    rocprofv3 --pmc summaries (profiles/*.md lines "traffic <kernel> n=<rows>
    FETCH_KiB=<f> WRITE_KiB=<w> fetch_scale=<a> write_scale=<b>"), or None."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path, encoding="utf-8"):
            m = re.match(r"traffic (\S+) n=(\d+) FETCH_KiB=([\d.]+) WRITE_KiB=([\d.]+) "
                         r"fetch_scale=([\d.]+) write_scale=([\d.]+)", line.strip())
            if m and m.group(1) == kernel and int(m.group(2)) == n:
                f, w, fs, ws = map(float, m.groups()[2:])
                return {"bytes": round((f * fs + w * ws) * 1024), "source": os.path.relpath(path, ROOT)}
    return None


def spmv_timing(oa, ctx, A, reps_warm, reps_cold, fp32=False):
    """Standalone SpMV durations from dispatch-attached events: `reps_warm`
    back-to-back launches (matrix and vectors stay in the 256 MB Infinity
    Cache when they fit), and `reps_cold` launches each preceded by a 1 GiB
    streaming write and a 1 GiB read of another buffer, which evict the
    matrix from every cache and leave clean lines behind (so the timed
    launch does not also pay the write-back of the flush)."""
    import torch
    x = torch.rand(A.n, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    if fp32:
        x, y = x.float(), y.float()
    fn = oa.spmv_f32 if fp32 else oa.spmv
    kind = oa.PROF_SPMV32 if fp32 else oa.PROF_SPMV64
    for _ in range(3):
        fn(A, x, y)
    ctx.profile(True)
    for _ in range(reps_warm):
        fn(A, x, y)
    warm = ctx.profile_read(kind)
    cold = np.zeros(0)
    if reps_cold:
        flush = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        clean = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        ctx.profile(True)
        for k in range(reps_cold):
            flush.fill_(float(k))
            clean.max()
            fn(A, x, y)
        cold = ctx.profile_read(kind)
        del flush, clean
    ctx.profile(False)
    torch.cuda.synchronize()
    return warm, cold


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model}


def cpu_baseline_poisson(args, d, A, b, vals_gpu, it64):
    """The oracle's CG+DIAG (the reference's algorithm in its operation
    order) on the same system, one host thread pinned to one core: full
    solves, median of `cpu_reps` (-Ofast, as the reference builds), one
    strict -O2 -ffp-contract=off solve, and the element-loop assembly of the
    values and right-hand side (the pattern is the GPU's, checked equal)."""
    from oracle import oracle as O
    rp, cl = A.pattern()
    rp64 = rp.astype(np.int64)
    ci64 = cl.astype(np.int64) + 1
    bh = b.cpu().numpy()
    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    try:
        O.lib(fast=True)
        O.lib(fast=False)
        ts, rets = [], []
        for _ in range(args.cpu_reps):
            t = time.perf_counter()
            ret, xo, res = O.cg(rp64, ci64, vals_gpu, bh, prec=O.DIAG, max_it=20000,
                                toler=args.toler, fast=True)
            ts.append(time.perf_counter() - t)
            rets.append(ret)
            log("[bench] cpu -Ofast solve %.2f s (%d it)" % (ts[-1], ret))
        t = time.perf_counter()
        ret_s, xs_, _ = O.cg(rp64, ci64, vals_gpu, bh, prec=O.DIAG, max_it=20000,
                             toler=args.toler, fast=False)
        strict_s = time.perf_counter() - t
        log("[bench] cpu strict -O2 solve %.2f s (%d it)" % (strict_s, ret_s))
        m = O.Mesh(d["coord"], d["code"], d["elem"].astype(np.int64), 1)
        dof, neq = O.number_equations(m.code)
        t = time.perf_counter()
        a, bo = O.assemble(O.LAPLACE3DT4, m, dof, neq, rp64, ci64, f=d["f"], coef=d["coef"])
        asm_s = time.perf_counter() - t
        log("[bench] cpu assembly (values + RHS) %.2f s" % asm_s)
    finally:
        os.sched_setaffinity(0, old)
    med = float(np.median(ts))
    return {"value": round(med * 1e3, 1), "unit": "ms per FP64 CG+DIAG solve to rtol %g" % args.toler,
            "cores": 1, "kind": "port",
            "sample": "%d full solves of this system (%s iterations) by the oracle built -Ofast, "
                      "one thread pinned to core %d; median" % (len(ts), sorted(set(rets)), core),
            "solves_s": [round(t_, 3) for t_ in ts], "strict_O2_solve_s": round(strict_s, 3),
            "strict_O2_iterations": ret_s, "assembly_values_rhs_s": round(asm_s, 3),
            "assembly_bitwise_equal_gpu": bool(np.array_equal(a, vals_gpu) and np.array_equal(bo, bh)),
            "gpu_fp64_iterations": it64, "host": cpu_info()}


def learned_leg(args, oa, W, ctx, d, P, b, ms64_zero, ms32_zero, it64_zero):
    """Learned start vectors for steady solves with a new coefficient field:
    the network is trained on `learned_seeds` other k-fields of the same mesh
    (k = exp(g), g the 64-mode Gaussian field of another seed; solutions by
    FP64 CG to rtol 1e-10), then guesses the start of this configuration's
    own solve (a k-field it never saw).  Its base is x = 0: the start is the
    exact A-norm line search along the predicted direction.  Reported beside
    the zero start: CG iterations and time per solve, FP64 and mixed, and
    the inference time."""
    import torch
    A = P.A
    n = A.n
    f = torch.as_tensor(d["f"], device="cuda")
    pts = torch.as_tensor(W.centroids(d["coord"], d["elem"]), device="cuda")
    zero = torch.zeros(n, dtype=torch.float64, device="cuda")
    samples = []
    t0 = time.perf_counter()
    for sd in range(args.learned_seeds):
        rng = W.SplitMix64(W.SEED_BASE + 100 + sd)
        k = torch.exp(W.fourier_field_device(rng, pts))
        bs = torch.zeros(n, dtype=torch.float64, device="cuda")
        P.build(oa.LAPLACE3DT4, coef=k, f=f, matrix=True, b=bs)
        xs = torch.zeros(n, dtype=torch.float64, device="cuda")
        oa.cg(A, bs, xs, prec=oa.DIAG_PREC, max_it=20000, toler=1e-10)
        samples.append((k, bs, xs))
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    net = oa.LearnedGuess(A, layers=args.learned_layers, seed=0, objective=args.learned_objective)
    x0buf = torch.empty(n, dtype=torch.float64, device="cuda")
    rng = np.random.default_rng(0)
    hist = []
    t0 = time.perf_counter()
    for ep in range(args.learned_epochs):
        tot = 0.0
        for i in rng.permutation(len(samples)):
            k, bs, xs = samples[i]
            P.build(oa.LAPLACE3DT4, coef=k, f=f, matrix=True, b=bs)
            net.refresh()
            tot += net.train_step(zero, zero, bs, xs, x0buf, lr=args.lr)
        hist.append(round(tot / len(samples), 5))
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log("[bench] learned start: %d k-fields (%.1fs), loss per epoch %s (%.1fs)"
        % (len(samples), gen_s, hist, train_s))
    # this configuration's own system again
    P.build(oa.LAPLACE3DT4, coef=torch.as_tensor(d["coef"], device="cuda"), f=f, matrix=True, b=b)
    net.refresh()
    x0 = torch.zeros(n, dtype=torch.float64, device="cuda")
    net.guess(zero, zero, b, x0)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(5):
        net.guess(zero, zero, b, x0)
    ev1.record()
    torch.cuda.synchronize()
    infer_ms = ev0.elapsed_time(ev1) / 5
    res = {}
    for name, fn in (("fp64", oa.cg), ("mixed", oa.cg_mixed)):
        ts, its = [], []
        for rep in range(3):
            x = torch.zeros(n, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            t = time.perf_counter()
            net.guess(zero, zero, b, x)
            info = fn(A, b, x, prec=oa.DIAG_PREC, max_it=20000, toler=args.toler)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
            its.append(info.iterations)
        res[name] = {"ms_per_solve": round(1e3 * float(np.median(ts)), 3), "iterations": its[-1],
                     "true_relres": info.true_relres}
    r0 = torch.zeros(n, dtype=torch.float64, device="cuda")
    oa.spmv(A, x0, r0)
    relres0 = float(torch.linalg.norm(b - r0) / torch.linalg.norm(b))
    # what any start vector can gain here: FP64 CG from the exact solution
    # perturbed by a smooth relative error eps (a 16-mode field, correlation 0.3)
    xstar = torch.zeros(n, dtype=torch.float64, device="cuda")
    oa.cg(A, b, xstar, prec=oa.DIAG_PREC, max_it=40000, toler=1e-13)
    rows = torch.as_tensor(d["coord"][d["code"][:, 0] == 0], device="cuda")
    g = W.fourier_field_device(W.SplitMix64(7), rows, modes=16, corr=0.3)
    sens = {}
    for eps in (1e-1, 1e-2, 1e-3, 1e-4):
        x = xstar * (1.0 + eps * g)
        info = oa.cg(A, b, x, prec=oa.DIAG_PREC, max_it=20000, toler=args.toler)
        sens["%g" % eps] = info.iterations
    log("[bench] FP64 CG iterations from the exact solution with a smooth relative error eps: %s "
        "(zero start: %d)" % (sens, it64_zero))
    log("[bench] learned start on the held-out field: relres0 %.3f, FP64 %d it %.2f ms (zero: %d it %.2f ms), "
        "mixed %d it %.2f ms (zero %.2f ms), inference %.3f ms"
        % (relres0, res["fp64"]["iterations"], res["fp64"]["ms_per_solve"], it64_zero, ms64_zero,
           res["mixed"]["iterations"], res["mixed"]["ms_per_solve"], ms32_zero, infer_ms))
    return {"train_fields": len(samples), "layers": args.learned_layers, "epochs": args.learned_epochs,
            "train_loss_per_epoch": hist, "train_seconds": round(train_s, 2),
            "inference_ms": round(infer_ms, 3), "start_relres": relres0,
            "fp64_learned": res["fp64"], "mixed_learned": res["mixed"],
            "fp64_zero_ms": round(ms64_zero, 3), "fp64_zero_iterations": it64_zero,
            "mixed_zero_ms": round(ms32_zero, 3),
            "speedup_mixed_learned_vs_fp64_zero": round(ms64_zero / res["mixed"]["ms_per_solve"], 3),
            "objective": args.learned_objective,
            "fp64_iterations_from_perturbed_exact_solution": sens}


def bench_poisson3d(args):
    """3D variable-k Poisson (config 1): GPU assembly, the SpMV roofline warm
    and cold (96^3 and a 128^3 replica larger than the Infinity Cache), FP64
    CG and mixed CG per solve, per-kernel durations of one profiled solve of
    each, and the oracle on one host core."""
    import torch
    import ofeli_amd as oa
    from ofeli_amd import workloads as W

    t0 = time.time()
    d = W.config1(args.n)
    log("[bench] config1 n=%d generated in %.1fs" % (args.n, time.time() - t0))
    ctx = oa.Context(0)
    t0 = time.time()
    P = oa.AssemblyPlan(ctx, d["coord"], d["elem"], d["code"])
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    A = P.A
    n, nnz = A.n, A.nnz
    log("[bench] plan: n=%d nnz=%d setup %.2fs" % (n, nnz, setup_s))
    coef = torch.as_tensor(d["coef"], device="cuda")
    f = torch.as_tensor(d["f"], device="cuda")
    b = torch.zeros(n, dtype=torch.float64, device="cuda")
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    P.build(oa.LAPLACE3DT4, coef=coef, f=f, matrix=True, b=b)  # warm
    ev0.record()
    P.build(oa.LAPLACE3DT4, coef=coef, f=f, matrix=True, b=b)
    ev1.record()
    torch.cuda.synchronize()
    asm_ms = ev0.elapsed_time(ev1)
    log("[bench] assembly %.3f ms" % asm_ms)

    # standalone SpMV, warm and cold
    sell = A.sell_info()
    dia = A.dia_info()
    warm64, cold64 = spmv_timing(oa, ctx, A, args.spmv_reps, args.cold_reps)
    warm32, _ = spmv_timing(oa, ctx, A, args.spmv_reps, 0, fp32=True)
    B64, B32 = b_spmv64(n, nnz), b_spmv32(n, nnz)
    spmv_ms, spmv_cold_ms, spmv32_ms = warm64.mean(), cold64.mean(), warm32.mean()
    log("[bench] spmv64 warm %.4f ms (%.0f GB/s)  cold %.4f ms (%.0f GB/s)  spmv32 warm %.4f ms"
        % (spmv_ms, B64 / spmv_ms / 1e6, spmv_cold_ms, B64 / spmv_cold_ms / 1e6, spmv32_ms))
    # streaming reference: a device-to-device copy of 1 GiB
    big = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    big2 = torch.empty_like(big)
    big2.copy_(big)
    ev0.record()
    for _ in range(5):
        big2.copy_(big)
    ev1.record()
    torch.cuda.synchronize()
    copy_gbs = 5 * 2 * big.numel() * 8 / (ev0.elapsed_time(ev1) * 1e-3) / 1e9
    del big, big2
    log("[bench] device copy %.0f GB/s (read+write)" % copy_gbs)

    def solve(fn):
        xs = torch.zeros(n, dtype=torch.float64, device="cuda")
        info = fn(A, b, xs, prec=oa.DIAG_PREC, max_it=20000, toler=args.toler)
        return info, xs

    for _ in range(args.warmup):
        solve(oa.cg)
        solve(oa.cg_mixed)
    t64, t32 = [], []
    torch.cuda.synchronize()
    for k in range(args.steps):
        t = time.perf_counter()
        i64, x64 = solve(oa.cg)
        torch.cuda.synchronize()
        t64.append(time.perf_counter() - t)
        t = time.perf_counter()
        i32, x32 = solve(oa.cg_mixed)
        torch.cuda.synchronize()
        t32.append(time.perf_counter() - t)
        log("[bench] step %d: fp64 %.2f ms (%d it, true %.2e) | mixed %.2f ms (%d it, %d repl, true %.2e)"
            % (k, t64[-1] * 1e3, i64.iterations, i64.true_relres, t32[-1] * 1e3, i32.iterations,
               i32.outer, i32.true_relres))
    err = float(torch.linalg.norm(x32 - x64) / torch.linalg.norm(x64))
    ms64 = 1e3 * float(np.median(t64))
    ms32 = 1e3 * float(np.median(t32))
    it64 = i64.iterations

    # per-kernel durations: one profiled solve of each (after the timed region)
    kb = kernel_bytes(n, nnz)
    kernels = {}
    ctx.profile(True)
    solve(oa.cg_mixed)
    solve(oa.cg)
    for name, kind in (("k_mx_dirspmv", oa.PROF_MX_DIRSPMV), ("k_mx_upd", oa.PROF_MX_UPD),
                       ("k_cg_spmv", oa.PROF_CG_SPMV), ("k_cg_upd", oa.PROF_CG_UPD),
                       ("k_cg_dir", oa.PROF_CG_DIR)):
        ms = ctx.profile_read(kind)
        w = working(ms)
        if w.size == 0:
            continue
        avg = float(w.mean())
        kernels[name] = {"launches": int(ms.size), "working_launches": int(w.size),
                         "avg_us": round(avg * 1e3, 2), "bytes_per_launch": kb[name],
                         "GBs": round(kb[name] / avg / 1e6, 1),
                         "frac": round(kb[name] / avg / 1e6 / HBM_PEAK_GBS, 4)}
        log("[bench] %-13s %4d/%4d launches  %.2f us  %.0f GB/s" % (
            name, w.size, ms.size, avg * 1e3, kb[name] / avg / 1e6))
    ctx.profile(False)
    # share of the mixed step's kernel time per kernel
    mx_total = sum(kernels[k]["avg_us"] * kernels[k]["working_launches"]
                   for k in ("k_mx_dirspmv", "k_mx_upd") if k in kernels)

    learned = None
    if args.learned_seeds > 0:
        learned = learned_leg(args, oa, W, ctx, d, P, b, ms64, ms32, it64)

    # replica larger than the 256 MB Infinity Cache
    replica = None
    if args.replica:
        del P, A, b, x32, x64, coef, f
        torch.cuda.empty_cache()
        t0 = time.time()
        dr = W.config1(args.replica)
        Pr = oa.AssemblyPlan(ctx, dr["coord"], dr["elem"], dr["code"])
        br = torch.zeros(Pr.nb_eq, dtype=torch.float64, device="cuda")
        Pr.build(oa.LAPLACE3DT4, coef=torch.as_tensor(dr["coef"], device="cuda"),
                 f=torch.as_tensor(dr["f"], device="cuda"), matrix=True, b=br)
        torch.cuda.synchronize()
        nr, nnzr = Pr.A.n, Pr.A.nnz
        dia_r = Pr.A.dia_info()
        wr, cr = spmv_timing(oa, ctx, Pr.A, args.spmv_reps, args.cold_reps)
        Br = b_spmv64(nr, nnzr)
        own_r = (dia_row_bytes(dia_r, 8) + 16) * nr if dia_r["fp64"] else None
        replica = {"rows": nr, "nnz": nnzr, "bytes_per_launch": Br,
                   "kernel": dia_kernel(dia_r, "double") if dia_r["fp64"]
                             else "k_spmv_sell<double,2>",
                   "warm_ms": round(float(wr.mean()), 4), "cold_ms": round(float(cr.mean()), 4),
                   "warm_GBs": round(Br / wr.mean() / 1e6, 1), "cold_GBs": round(Br / cr.mean() / 1e6, 1),
                   "own_bytes_per_launch": own_r,
                   "own_bytes_definition": "matrix copy + x read 8n + y written 8n",
                   "own_cold_frac": None if own_r is None
                                    else round(own_r / cr.mean() / 1e6 / HBM_PEAK_GBS, 4)}
        log("[bench] replica %d^3: n=%d nnz=%d spmv64 warm %.4f ms (%.0f GB/s) cold %.4f ms (%.0f GB/s) (%.0fs)"
            % (args.replica, nr, nnzr, wr.mean(), Br / wr.mean() / 1e6, cr.mean(),
               Br / cr.mean() / 1e6, time.time() - t0))
        Pr.close()
        del Pr, br, dr

    cpu = None
    if not args.no_cpu:
        # the 96^3 system again (the GPU copies were released for the replica)
        Pc = oa.AssemblyPlan(ctx, d["coord"], d["elem"], d["code"])
        bc_ = torch.zeros(Pc.nb_eq, dtype=torch.float64, device="cuda")
        Pc.build(oa.LAPLACE3DT4, coef=torch.as_tensor(d["coef"], device="cuda"),
                 f=torch.as_tensor(d["f"], device="cuda"), matrix=True, b=bc_)
        cpu = cpu_baseline_poisson(args, d, Pc.A, bc_, Pc.A.values(), it64)
        Pc.close()

    dom = kernels.get("k_mx_dirspmv")
    traffic = pmc_traffic("k_mx_dirspmv", n)
    out = {
        "metric": METRIC, "value": round(ms32, 3), "unit": "ms per solve (mixed FP32/FP64 CG, rtol 1e-8)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms32, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "3D variable-k P1 Poisson, Kuhn tets %d^3 cells" % args.n,
                   "rows": n, "nnz": nnz, "toler": args.toler, "precond": "DIAG_PREC"},
        "fp64_cg_ms": round(ms64, 3), "fp64_iterations": it64, "mixed_iterations": i32.iterations,
        "mixed_replacements": i32.outer, "speedup_mixed_vs_fp64": round(ms64 / ms32, 3),
        "mixed_err_vs_fp64": err, "fp64_true_relres": i64.true_relres,
        "mixed_true_relres": i32.true_relres, "assembly_ms": round(asm_ms, 3),
        "plan_setup_s": round(setup_s, 3), "cg64_us_per_iter": round(ms64 * 1e3 / max(it64, 1), 2),
        "device_copy_GBs": round(copy_gbs, 1),
        "kernels": kernels,
        "mixed_kernel_share": {k: round(kernels[k]["avg_us"] * kernels[k]["working_launches"] / mx_total, 3)
                               for k in ("k_mx_dirspmv", "k_mx_upd") if k in kernels and mx_total},
        "roofline": None if dom is None else {
            "bound": "hbm", "achieved": dom["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": dom["frac"], "traffic": None if traffic is None else traffic["bytes"],
            "traffic_source": None if traffic is None else traffic["source"],
            "kernel": "k_mx_dirspmv_dia (mixed CG inner product, 2 rows per lane; q by "
                      "recurrence, direction, (p,q) partials) -- "
                      + dia_layout(dia, "FP32 D^-1/2 A D^-1/2"),
            "bytes_per_launch": dom["bytes_per_launch"], "avg_ms": round(dom["avg_us"] / 1e3, 5),
            "bytes_definition": "B_spmv32 = 8 nnz + 4(n+1) + 8n (the FP32 CSR SpMV)",
            "own_bytes_per_launch": (dia_row_bytes(dia, 4) + 20) * n,
            "own_bytes_definition": "matrix copy %d n + z 4n + p_old, q_old read, p_new, q "
                                    "written 16n" % dia_row_bytes(dia, 4),
            "own_frac": round((dia_row_bytes(dia, 4) + 20) * n / (dom["avg_us"] * 1e3)
                              / HBM_PEAK_GBS, 4)},
        "fp32_copy": {"diagonal_copy": dia, "inner_operator": "D^-1/2 A D^-1/2 (DIAG_PREC as a "
                      "symmetric scaling of the FP32 inner iteration)"},
        "mixed_iteration": None if dom is None or "k_mx_upd" not in kernels else {
            "us": round(dom["avg_us"] + kernels["k_mx_upd"]["avg_us"], 2),
            "bytes_per_iteration": b_spmv32(n, nnz) + 44 * n,
            "bytes_definition": "B_cg32_iter = B_spmv32 + 44n (the FP32 CSR inner CG iteration)",
            "GBs": round((b_spmv32(n, nnz) + 44 * n)
                         / ((dom["avg_us"] + kernels["k_mx_upd"]["avg_us"]) * 1e3), 1),
            "frac": round((b_spmv32(n, nnz) + 44 * n)
                          / ((dom["avg_us"] + kernels["k_mx_upd"]["avg_us"]) * 1e3)
                          / HBM_PEAK_GBS, 4)},
        "roofline_spmv64": {
            "kernel": (dia_kernel(dia, "double") + " (" + dia_layout(dia, "FP64") + ")")
                      if dia["fp64"] else
                      "k_spmv_sell<double,2> (SELL-64, %d-byte columns, %d stored entries)"
                      % (sell["col_bytes"], sell["stored"]),
            "bytes_per_launch": B64, "bytes_definition": "B_spmv64 = 12 nnz + 4(n+1) + 16n",
            "warm_ms": round(float(spmv_ms), 5), "warm_GBs": round(B64 / spmv_ms / 1e6, 1),
            "warm_frac": round(B64 / spmv_ms / 1e6 / HBM_PEAK_GBS, 4),
            "cold_ms": round(float(spmv_cold_ms), 5), "cold_GBs": round(B64 / spmv_cold_ms / 1e6, 1),
            "cold_frac": round(B64 / spmv_cold_ms / 1e6 / HBM_PEAK_GBS, 4),
            "cold_protocol": "each launch after a 1 GiB streaming write and a 1 GiB read",
            "own_bytes_per_launch": (dia_row_bytes(dia, 8) + 16) * n if dia["fp64"] else None,
            "own_bytes_definition": "matrix copy + x read 8n + y written 8n",
            "own_cold_frac": round((dia_row_bytes(dia, 8) + 16) * n / spmv_cold_ms / 1e6
                                   / HBM_PEAK_GBS, 4) if dia["fp64"] else None,
            "traffic": (pmc_traffic(dia_kernel(dia, "double") if dia["fp64"]
                                    else "k_spmv_sell<double,2>", n) or {}).get("bytes"),
            "spmv32_warm_ms": round(float(spmv32_ms), 5),
            "spmv32_warm_GBs": round(B32 / spmv32_ms / 1e6, 1),
            "spmv32_kernel": dia_kernel(dia, "float") if dia["fp32"]
                             else "k_spmv_sell<float,4>"},
        "roofline_cold": None if replica is None else dict(
            replica, workload="%d^3 Kuhn replica" % args.replica,
            bound="hbm", peak=HBM_PEAK_GBS, unit="GB/s",
            frac=round(replica["cold_GBs"] / HBM_PEAK_GBS, 4),
            bytes_definition="B_spmv64 = 12 nnz + 4(n+1) + 16n",
            cold_protocol="each launch after a 1 GiB streaming write and a 1 GiB read"),
        "cpu_baseline": cpu,
        "learned_start": learned,
    }
    if args.sweep_precs or args.krylov or args.mesh_io:
        Pe = oa.AssemblyPlan(ctx, d["coord"], d["elem"], d["code"])
        be = torch.zeros(Pe.nb_eq, dtype=torch.float64, device="cuda")
        Pe.build(oa.LAPLACE3DT4, coef=torch.as_tensor(d["coef"], device="cuda"),
                 f=torch.as_tensor(d["f"], device="cuda"), matrix=True, b=be)
        vals = Pe.A.values() if not args.no_cpu else None
        if args.sweep_precs:
            out["sweep_preconditioners"] = sweep_precs(args, oa, Pe.A, be, vals)
        if args.krylov:
            out["other_krylov"] = other_krylov(args, oa, Pe.A, be, ms64, it64, vals)
        if args.mesh_io:
            out["mesh_io"] = mesh_io_leg(args, oa, ctx)
    print(json.dumps(out), flush=True)


def mesh_io_leg(args, oa, ctx, n=1024):
    """OFELI .m files: Mesh::put and Mesh::get (C++ writer / reader) on a
    structured n x n P1 mesh, then the GPU plan (graph + numbering) and a
    Laplace2DT3 assembly of the mesh read back; the oracle's Python reader
    on the same file as the CPU baseline."""
    import tempfile
    import torch
    from ofeli_amd.api import MeshFile
    coord, code, elem = oa.api.mesh2d(n, n)
    m = MeshFile(2, 1, coord, code, elem, np.ones(len(elem), np.int32), 3,
                 np.zeros((0, 2), np.int32), np.zeros((0, 1), np.int32), 2)
    d = tempfile.mkdtemp()
    path = os.path.join(d, "mesh.m")
    t = time.perf_counter()
    oa.write_mesh(path, m)
    tw = time.perf_counter() - t
    size = os.path.getsize(path)
    t = time.perf_counter()
    r = oa.read_mesh(path)
    tr = time.perf_counter() - t
    torch.cuda.synchronize()
    t = time.perf_counter()
    P = oa.AssemblyPlan(ctx, r.coord, r.elem, r.code)
    b = torch.zeros(P.nb_eq, dtype=torch.float64, device="cuda")
    P.build(oa.LAPLACE2DT3, f=torch.ones(len(r.coord), dtype=torch.float64, device="cuda"),
            matrix=True, b=b)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t
    res = {"nodes": len(r.coord), "elements": len(r.elem), "file_MB": round(size / 1e6, 1),
           "write_s": round(tw, 3), "read_s": round(tr, 3),
           "read_MB_per_s": round(size / 1e6 / tr, 1), "write_MB_per_s": round(size / 1e6 / tw, 1),
           "plan_and_assembly_s": round(tp, 3), "rows": P.nb_eq}
    log("[bench] mesh io: %.1f MB, write %.3f s, read %.3f s (%.0f MB/s), plan+assembly %.3f s"
        % (size / 1e6, tw, tr, size / 1e6 / tr, tp))
    if not args.no_cpu:
        from oracle import mesh_io as MIO
        t = time.perf_counter()
        MIO.read(path)
        to = time.perf_counter() - t
        res["cpu_oracle_read_s"] = round(to, 2)
        log("[bench] oracle (Python) read %.2f s" % to)
    os.remove(path)
    os.rmdir(d)
    return res


def other_krylov(args, oa, A, b, cg_ms, cg_it, vals):
    """BiCG (with its A^T product), BiCGStab, CGS and GMRes(50) on the same
    system with DIAG_PREC.  The reference's CGS does not converge here (its q
    stays 0), so it runs a fixed 200 iterations and is reported per iteration."""
    import torch
    n = A.n
    res = {"cg": {"solve_ms": round(cg_ms, 2), "iterations": cg_it,
                  "us_per_iteration": round(cg_ms * 1e3 / cg_it, 1)}}
    for name, fn, max_it, toler in (("bicg", oa.bicg, 20000, args.toler),
                                    ("bicgstab", oa.bicgstab, 20000, args.toler),
                                    ("cgs_200_iterations", oa.cgs, 200, 1e-300),
                                    ("gmres_50", lambda A_, b_, x_, prec, max_it, toler:
                                     oa.gmres(A_, b_, x_, 50, prec=prec, max_it=max_it,
                                              toler=toler), 2000, args.toler)):
        ts = []
        for rep in range(3):
            x = torch.zeros(n, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            t = time.perf_counter()
            info = fn(A, b, x, prec=oa.DIAG_PREC, max_it=max_it, toler=toler)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        ms = 1e3 * min(ts[1:])
        res[name] = {"solve_ms": round(ms, 2), "iterations": info.iterations, "nb_it": info.nb_it,
                     "true_relres": info.true_relres,
                     "us_per_iteration": round(ms * 1e3 / max(info.iterations, 1), 1)}
        log("[bench] %s: %.2f ms, %d it (%.1f us/it), true %.2e"
            % (name, ms, info.iterations, ms * 1e3 / max(info.iterations, 1), info.true_relres))
    if vals is not None:
        from oracle import oracle as O
        rp, cl = A.pattern()
        bh = b.cpu().numpy()
        t = time.perf_counter()
        O.bicg(rp.astype(np.int64), cl.astype(np.int64) + 1, vals, bh, prec=O.DIAG, max_it=20,
               toler=1e-300, fast=True)
        per_it = (time.perf_counter() - t) / 20
        res["cpu_bicg_ms_per_iteration"] = round(per_it * 1e3, 2)
        log("[bench] cpu BiCG+DIAG %.1f ms/it" % (per_it * 1e3))
    return res


def sweep_precs(args, oa, A, b, vals):
    """CG with the preconditioners applied by level-scheduled triangular
    sweeps (DILU_PREC = ILU_PREC, SSOR_PREC) on the same system: the first
    application (level sets, factorisation), one application, one solve,
    and the oracle's CG iteration with the same preconditioner on one core."""
    import torch
    n = A.n
    res = {}
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for name, pc in (("dilu", oa.DILU_PREC), ("ssor", oa.SSOR_PREC)):
        z = torch.zeros(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t = time.perf_counter()
        oa.prec_apply(A, pc, b, z)
        torch.cuda.synchronize()
        first = time.perf_counter() - t
        ev0.record()
        for _ in range(3):
            oa.prec_apply(A, pc, b, z)
        ev1.record()
        torch.cuda.synchronize()
        apply_ms = ev0.elapsed_time(ev1) / 3
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t = time.perf_counter()
        info = oa.cg(A, b, x, prec=pc, max_it=20000, toler=args.toler)
        torch.cuda.synchronize()
        solve_ms = (time.perf_counter() - t) * 1e3
        res[name] = {"solve_ms": round(solve_ms, 2), "iterations": info.iterations,
                     "true_relres": info.true_relres, "apply_ms": round(apply_ms, 3),
                     "first_apply_ms": round(first * 1e3, 2)}
        log("[bench] CG+%s: %.1f ms, %d it, application %.3f ms (first %.1f ms)"
            % (name, solve_ms, info.iterations, apply_ms, first * 1e3))
    if vals is not None:
        from oracle import oracle as O
        rp, cl = A.pattern()
        bh = b.cpu().numpy()
        t = time.perf_counter()
        O.cg(rp.astype(np.int64), cl.astype(np.int64) + 1, vals, bh, prec=O.DILU, max_it=10,
             toler=1e-300, fast=True)
        per_it = (time.perf_counter() - t) / 10
        res["cpu_dilu_ms_per_iteration"] = round(per_it * 1e3, 2)
        res["cpu_dilu_ms_per_solve"] = round(per_it * res["dilu"]["iterations"] * 1e3, 1)
        log("[bench] cpu CG+DILU %.1f ms/it" % (per_it * 1e3))
    return res


def bench_batched(args):
    """512 uneven DC2DT3 systems: persistent batched CG (FP64 and mixed,
    autotuned launch shape, graph replay) vs a loop of per-system ofh_cg."""
    import torch
    import ofeli_amd as oa
    from ofeli_amd import workloads as W

    t0 = time.time()
    d = W.config2()
    log("[bench2] config2 generated in %.1fs" % (time.time() - t0))
    ctx = oa.Context(0)
    P = oa.AssemblyPlan(ctx, d["coord"], d["elem"], d["code"])
    A = P.A
    n, nnz = A.n, A.nnz
    sys_rows = d["sys_rows"]
    nsys = len(sys_rows) - 1
    b = torch.zeros(n, dtype=torch.float64, device="cuda")
    P.build(oa.DC2DT3, coef=torch.as_tensor(d["coef"], device="cuda"),
            f=torch.as_tensor(d["f"], device="cuda"), matrix=True, b=b)
    torch.cuda.synchronize()
    log("[bench2] %d systems, n=%d nnz=%d" % (nsys, n, nnz))
    B = oa.Batch(A, sys_rows)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    kw = dict(max_it=20000, toler=args.toler)
    shape64 = B.autotune(b, x, mixed=False, **kw)
    shape32 = B.autotune(b, x, mixed=True, **kw)
    log("[bench2] autotune fp64 %s mixed %s" % (shape64, shape32))

    def run(mixed):
        B.set_config(*(shape32 if mixed else shape64)[:2])
        xs = torch.zeros(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t = time.perf_counter()
        it, st, rr, info = B.cg(b, xs, mixed=mixed, **kw)
        torch.cuda.synchronize()
        return time.perf_counter() - t, it, st, rr, xs

    for _ in range(args.warmup):
        run(False)
        run(True)
    t64, t32 = [], []
    for k in range(args.steps):
        r64 = run(False)
        r32 = run(True)
        t64.append(r64[0])
        t32.append(r32[0])
        log("[bench2] step %d: batched fp64 %.2f ms | mixed %.2f ms" % (k, r64[0] * 1e3, r32[0] * 1e3))
    _, it64, st64, rr64, x64 = r64
    _, it32, st32, rr32, x32 = r32

    # baseline: one ofh_cg per system (the reference's loop of LinearSolver
    # solves, each on the GPU)
    rp, col = A.pattern()
    val = A.values()
    mats = []
    for s in range(nsys):
        r0, r1 = int(sys_rows[s]), int(sys_rows[s + 1])
        k0, k1 = int(rp[r0]), int(rp[r1])
        mats.append(oa.CsrMatrix(ctx, r1 - r0, rp[r0:r1 + 1] - k0, col[k0:k1] - r0, val[k0:k1]))

    def loop():
        xs = torch.zeros(n, dtype=torch.float64, device="cuda")
        its = []
        torch.cuda.synchronize()
        t = time.perf_counter()
        for s in range(nsys):
            r0, r1 = int(sys_rows[s]), int(sys_rows[s + 1])
            info = oa.cg(mats[s], b[r0:r1], xs[r0:r1], prec=oa.DIAG_PREC, **kw)
            its.append(info.iterations)
        torch.cuda.synchronize()
        return time.perf_counter() - t, np.array(its), xs

    loop()
    tl = []
    for _ in range(max(1, min(args.steps, 2))):
        tl_s, it_loop, x_loop = loop()
        tl.append(tl_s)
    ms_loop = 1e3 * float(np.median(tl))
    err64 = float(torch.linalg.norm(x64 - x_loop) / torch.linalg.norm(x_loop))
    err32 = float(torch.linalg.norm(x32 - x_loop) / torch.linalg.norm(x_loop))
    ms64 = 1e3 * float(np.median(t64))
    ms32 = 1e3 * float(np.median(t32))
    sizes = np.diff(sys_rows)
    nnz_s = np.array([rp[sys_rows[s + 1]] - rp[sys_rows[s]] for s in range(nsys)])
    # algorithmic bytes per system and iteration.  Streamed form: one SpMV
    # (12 B/nnz + 4 B/row pointer + x, y) + the CG vector streams.  Resident
    # form: the matrix (12 B/nnz FP64, 8 B/nnz FP32) + x read and written
    # (FP64) + dinv; the direction and the row table stay in LDS, r in
    # registers.  Mixed counts every iteration as an FP32 one.
    def it_bytes(shape, mixed):
        if shape[1] == 0:
            return 8 * nnz_s + 4 * sizes if mixed else 12 * nnz_s + 24 * sizes
        return 8 * nnz_s + 48 * sizes if mixed else 12 * nnz_s + 108 * sizes
    bytes64 = float(np.sum(it64.astype(np.float64) * it_bytes(shape64, False)))
    bytes32 = float(np.sum(it32.astype(np.float64) * it_bytes(shape32, True)))
    eff_gbs = bytes32 / (ms32 * 1e-3) / 1e9
    eff64_gbs = bytes64 / (ms64 * 1e-3) / 1e9

    def kname(shape, mixed):
        if shape[1] == 0:
            return "k_batch_cgmx_res (resident batched mixed CG)" if mixed else \
                "k_batch_cg64_res (resident batched CG, FP64)"
        return ("k_batch_cgmx<%d,%d>" if mixed else "k_batch_cg64<%d,%d>") % tuple(shape[:2])

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        O.lib(fast=True)
        sample = list(range(nsys))  # the whole batch: a few seconds on one core
        t = time.perf_counter()
        work = 0.0
        for s in sample:
            r0, r1 = int(sys_rows[s]), int(sys_rows[s + 1])
            k0, k1 = int(rp[r0]), int(rp[r1])
            O.cg((rp[r0:r1 + 1] - k0).astype(np.int64), (col[k0:k1] - r0).astype(np.int64) + 1,
                 val[k0:k1], b[r0:r1].cpu().numpy(), prec=O.DIAG, max_it=20000, toler=args.toler,
                 fast=True)
            work += float(it_loop[s]) * (nnz_s[s] + 10 * sizes[s])
        cpu_s = time.perf_counter() - t
        total = float(np.sum(it_loop * (nnz_s + 10 * sizes)))
        cpu = {"value": round(cpu_s * total / work * 1e3, 1),
               "unit": "ms per batch of FP64 CG+DIAG solves (extrapolated)", "cores": 1,
               "kind": "port",
               "sample": "%d of the %d systems solved one after another by the oracle -Ofast "
                         "on 1 host core in %.2f s" % (len(sample), nsys, cpu_s)}
        log("[bench2] cpu sample %.2fs -> %.0f ms" % (cpu_s, cpu["value"]))

    out = {
        "metric": METRIC, "value": round(ms32, 3),
        "unit": "ms per batch of 512 solves (batched mixed CG, rtol %g)" % args.toler,
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms32, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "512 DC2DT3 steady heat systems, n x n cells, n in [16,96]",
                   "rows": n, "nnz": nnz, "systems": nsys, "toler": args.toler,
                   "precond": "DIAG_PREC"},
        "batched_fp64_ms": round(ms64, 3), "loop_fp64_ms": round(ms_loop, 3),
        "speedup_batched_fp64_vs_loop": round(ms_loop / ms64, 2),
        "speedup_batched_mixed_vs_loop": round(ms_loop / ms32, 2),
        "shape_fp64": list(shape64[:2]), "shape_mixed": list(shape32[:2]),
        "max_iterations_fp64": int(it64.max()), "max_iterations_mixed": int(it32.max()),
        "all_converged": bool((st64 >= 0).all() and (st32 >= 0).all()),
        "max_true_relres_mixed": float(rr32.max()),
        "batched_fp64_err_vs_loop": err64, "batched_mixed_err_vs_loop": err32,
        "roofline": {"bound": "hbm", "achieved": round(eff_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(eff_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": kname(shape32, True),
                     "note": "effective: algorithmic bytes / time; each system's matrix is "
                             "re-read from L2/MALL in every iteration of its solve",
                     "bytes_per_launch": bytes32, "avg_ms": round(ms32, 3)},
        "roofline_fp64": {"achieved": round(eff64_gbs, 1), "unit": "GB/s",
                          "frac": round(eff64_gbs / HBM_PEAK_GBS, 4),
                          "kernel": kname(shape64, False), "bytes_per_launch": bytes64,
                          "avg_ms": round(ms64, 3)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_transient(args):
    """Transient heat (DC2DT3, backward Euler, 512^2 cells, 2000 steps of
    1e-3).  Steps 1..200 run FP64 CG warm-started from u^n (the reference's
    TimeStepping start); their (inputs, solution) pairs train the learned
    guesses.  Steps 201..T are then run from the same state with each start
    vector and solver: warm, extrapolated, learned, projection onto previous
    solutions, projection corrected by the network; FP64 and mixed CG."""
    import torch
    import ofeli_amd as oa
    from ofeli_amd import workloads as W
    from ofeli_amd.transient import HeatBE

    d = W.config3(args.n3)
    ctx = oa.Context(0)
    t0 = time.time()
    S = HeatBE(ctx, d["coord"], d["elem"], d["code"], d["coef"], d["dt"])
    torch.cuda.synchronize()
    n, nnz = S.A.n, S.A.nnz
    log("[bench3] n=%d nnz=%d setup %.2fs" % (n, nnz, time.time() - t0))
    T, K, PM = args.t_steps, args.train_steps, args.proj_m
    xy = torch.as_tensor(d["coord"], device="cuda")
    # every step's nodal source, resident in HBM before timing
    F = torch.empty((T, xy.shape[0]), dtype=torch.float64, device="cuda")
    for s_ in range(T):
        F[s_] = W.source_field(xy, d["path"], (s_ + 1) * d["dt"])
    torch.cuda.synchronize()
    tol = args.toler

    # phase A: FP64 warm-started steps, training samples
    proj = oa.ProjectionGuess(S.A, m=PM, minres=True)
    xg = torch.empty_like(S.b)
    samples, samples_p, last = [], [], []
    tA = time.perf_counter()
    itA = []
    for s_ in range(K):
        S.rhs(F[s_])
        S.start("warm")
        x_n, x_p, b = S.x.clone(), S.xp.clone(), S.b.clone()
        proj.guess(S.b, xg)
        info = S.solve(mixed=False, toler=tol)
        itA.append(info.iterations)
        xs = S.x0.clone()
        samples.append((x_n, x_p, b, xs))
        samples_p.append((xg.clone(), x_n, b, xs))
        S.advance()
        proj.update(S.x)
        last = (last + [xs])[-PM:]
    torch.cuda.synchronize()
    tA = time.perf_counter() - tA
    log("[bench3] steps 1..%d fp64 warm: %.2f ms/step, mean %.1f it" % (K, tA / K * 1e3, np.mean(itA)))
    st = S.state()
    proj.close()

    def train(net, data, tag):
        x0buf = torch.empty_like(S.b)
        rng = np.random.default_rng(0)
        t0 = time.perf_counter()
        hist = []
        nsteps = args.epochs * len(data)
        k_ = 0
        for ep in range(args.epochs):
            tot = 0.0
            for i in rng.permutation(len(data)):
                lr = args.lr * 0.5 * (1 + np.cos(np.pi * k_ / nsteps))  # cosine decay
                tot += net.train_step(*data[i], x0buf, lr=lr)
                k_ += 1
            hist.append(round(tot / len(data), 5))
        torch.cuda.synchronize()
        log("[bench3] %s: loss per epoch %s (%.1fs)" % (tag, hist, time.perf_counter() - t0))
        return hist, time.perf_counter() - t0

    net_x = oa.LearnedGuess(S.A, layers=4, seed=0, extrapolate=True)
    hist_x, tr_x = train(net_x, samples, "net on extrapolation")
    net_p = oa.LearnedGuess(S.A, layers=4, seed=1, extrapolate=False)
    hist_p, tr_p = train(net_p, samples_p, "net on projection")
    del samples, samples_p

    def make_proj(minres=False):
        P = oa.ProjectionGuess(S.A, m=PM, minres=minres)
        for x in last:
            P.update(x)
        return P

    def run(make, mixed, judge=0):
        guess = make()
        S.set_state(st)
        S.judge_every, S.use_net, S.judged, S.net_steps = judge, True, 0, 0
        torch.cuda.synchronize()
        its, rr = [], []
        t = time.perf_counter()
        for s_ in range(K, T):
            info = S.step(F[s_], guess=guess, mixed=mixed, toler=tol)
            its.append(info.iterations)
            rr.append(info.relres if not mixed else info.true_relres)
        torch.cuda.synchronize()
        dt_ = time.perf_counter() - t
        return dt_ / (T - K) * 1e3, np.array(its), max(rr), S.x.clone()

    variants = [("fp64_warm", lambda: "warm", False),
                ("fp64_extrap", lambda: "extrapolate", False),
                ("fp64_learned", lambda: net_x, False),
                ("fp64_proj", make_proj, False),
                ("fp64_proj_mr", lambda: make_proj(True), False),
                ("fp64_proj_mr_learned", lambda: (make_proj(True), net_p), False),
                ("mixed_warm", lambda: "warm", True),
                ("mixed_proj_mr", lambda: make_proj(True), True),
                ("mixed_proj_mr_learned", lambda: (make_proj(True), net_p), True),
                # the projection judges the network every 25th step (HeatBE.judge_every)
                ("mixed_proj_mr_learned_judged", lambda: (make_proj(True), net_p), True, 25)]
    res = {}
    judge_stats = {}
    for v in variants:
        name, make, mx = v[:3]
        ms, its, rr, xf = run(make, mx, v[3] if len(v) > 3 else 0)
        res[name] = (ms, its, rr, xf)
        if len(v) > 3:
            judge_stats = {"judged_steps": S.judged, "steps_started_by_the_network": S.net_steps,
                           "steps": T - K, "every": v[3], "keep_if_residual_below": S.judge_keep}
        log("[bench3] %-19s %.3f ms/step  mean it %.1f  max relres %.2e" % (name, ms, its.mean(), rr))
    if judge_stats:
        log("[bench3] judge: %s" % judge_stats)
    # cost of the start vectors alone
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    S.set_state(st)
    S.rhs(F[K])
    P = make_proj(True)
    timings = {}
    for tag, fn in (("learned", lambda: net_x.guess(S.x, S.xp, S.b, S.x0)),
                    ("proj_guess", lambda: P.guess(S.b, S.x0)),
                    ("proj_update", lambda: P.update(S.x))):
        fn()
        ev0.record()
        for _ in range(20):
            fn()
        ev1.record()
        torch.cuda.synchronize()
        timings[tag] = round(ev0.elapsed_time(ev1) / 20, 4)
    base = res["fp64_warm"]
    xb = base[3]
    errs = {k: float(torch.linalg.norm(v[3] - xb) / torch.linalg.norm(xb)) for k, v in res.items()}

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        O.lib(fast=True)
        S.set_state(st)
        S.rhs(F[K])
        rp, col = S.A.pattern()
        vals = S.A.values()
        bh, xh = S.b.cpu().numpy(), S.x.cpu().numpy()
        t = time.perf_counter()
        ret, _, _ = O.cg(rp.astype(np.int64), col.astype(np.int64) + 1, vals, bh, x0=xh,
                         prec=O.DIAG, max_it=20000, toler=tol, fast=True)
        cpu_s = time.perf_counter() - t
        cpu = {"value": round(cpu_s * 1e3, 2), "unit": "ms per FP64 CG+DIAG solve (warm start)",
               "cores": 1, "kind": "port",
               "sample": "the solve of step %d (%d iterations) by the oracle -Ofast on 1 host core"
                         % (K + 1, ret)}
        log("[bench3] cpu %.1f ms (%d it)" % (cpu_s * 1e3, ret))

    head = "mixed_proj_mr_learned_judged"
    ms_h = res[head][0]
    out = {
        "metric": METRIC, "value": round(ms_h, 4),
        "unit": "ms per time step (RHS assembly + projected start, corrected by the network "
                "while the projection judges it worth it, + mixed CG to rtol %g)" % tol,
        "n_gpus": 1, "steps": T - K, "warmup": K, "ms_per_step": round(ms_h, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64+bf16",
        "data": "synthetic",
        "config": {"workload": "transient DC2DT3 backward Euler %dx%d, dt 1e-3, moving source"
                               % (args.n3, args.n3), "rows": n, "nnz": nnz, "time_steps": T,
                   "train_steps": K, "toler": tol, "precond": "DIAG_PREC", "proj_vectors": PM},
        "baseline": "fp64_warm", "headline": head, "judge": judge_stats,
        "variants_ms_per_step": {k: round(v[0], 4) for k, v in res.items()},
        "variants_mean_iterations": {k: round(float(v[1].mean()), 2) for k, v in res.items()},
        "variants_max_relres": {k: float(v[2]) for k, v in res.items()},
        "final_state_rel_err_vs_fp64_warm": errs,
        "speedup_vs_fp64_warm": {k: round(base[0] / v[0], 3) for k, v in res.items()},
        "train_seconds": round(tr_x + tr_p, 2),
        "train_loss_per_epoch": {"net_extrapolation": hist_x, "net_projection": hist_p},
        "start_vector_ms": timings, "phaseA_ms_per_step": round(tA / K * 1e3, 4),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_cantilever(args):
    """Cantilever (Elas2DQ4 plane stress, 2048 x 256 quads, 1,052,672 free
    DOFs) under 256 right-edge load cases that combine 8 basis tractions.
    Per-case FP64 CG from zero (the reference solves each case by itself) is
    timed on the first cases; the whole set is then solved with the start
    vector projected on the previous solutions, FP64 and mixed, and with the
    learned direction on top (trained on the first cases).  SSOR_PREC, the
    preconditioner the reference's run names, is timed per application, per
    CG+SSOR solve from zero (--ssor-cases) and per case from the projected
    start once the basis holds the first solutions (--ssor-proj-cases)."""
    import torch
    import ofeli_amd as oa
    from ofeli_amd import workloads as W

    t0 = time.time()
    d = W.config4(args.nx4, args.ny4, ncases=args.cases)
    log("[bench4] workload %.1fs" % (time.time() - t0))
    ctx = oa.Context(0)
    t0 = time.time()
    P = oa.AssemblyPlan(ctx, d["coord"], d["elem"], d["code"])
    A = P.A
    n, nnz = A.n, A.nnz
    torch.cuda.synchronize()
    log("[bench4] plan n=%d nnz=%d %.2fs" % (n, nnz, time.time() - t0))
    E = torch.as_tensor(d["coef"], device="cuda")
    b0 = torch.zeros(n, dtype=torch.float64, device="cuda")
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    P.build(oa.ELAS2DQ4, coef=E, prm=d["prm"], matrix=True, b=b0)
    ev0.record()
    P.build(oa.ELAS2DQ4, coef=E, prm=d["prm"], matrix=True, b=b0)
    ev1.record()
    torch.cuda.synchronize()
    asm_ms = ev0.elapsed_time(ev1)
    B = torch.zeros((len(d["basis"]), n), dtype=torch.float64, device="cuda")
    for k, tr in enumerate(d["basis"]):
        P.traction(d["sides"], torch.as_tensor(tr, device="cuda"), B[k])
    Bc = torch.as_tensor(d["alpha"], device="cuda") @ B  # every case's RHS, resident in HBM
    torch.cuda.synchronize()
    ncases = Bc.shape[0]
    tol = args.toler
    kw = dict(max_it=200000, toler=tol)

    # exact SGS (SSOR_PREC, omega = 1): per application and per CG iteration
    r = Bc[0].clone()
    z = torch.zeros_like(r)
    oa.ssor_solve(A, r, z)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(3):
        oa.ssor_solve(A, r, z)
    torch.cuda.synchronize()
    ssor_ms = (time.perf_counter() - t) / 3 * 1e3
    xs_ = torch.zeros_like(r)
    t = time.perf_counter()
    inf = oa.cg(A, Bc[0], xs_, prec=oa.SSOR_PREC, max_it=20, toler=1e-30)
    torch.cuda.synchronize()
    cg_ssor_it_ms = (time.perf_counter() - t) * 1e3 / max(inf.iterations, 1)
    log("[bench4] SSOR application %.2f ms, CG+SSOR %.2f ms/iteration" % (ssor_ms, cg_ssor_it_ms))

    def per_case(fn, cases, prec=oa.DIAG_PREC):
        its, ms, xs = [], [], []
        for c in cases:
            x = torch.zeros(n, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            t = time.perf_counter()
            info = fn(A, Bc[c], x, prec=prec, **kw)
            torch.cuda.synchronize()
            ms.append((time.perf_counter() - t) * 1e3)
            its.append(info.iterations)
            xs.append(x)
            log("[bench4]   case %d: %d it, %.1f ms%s" % (c, its[-1], ms[-1],
                                                       " (FP64 fallback)" if info.fallback else ""))
        return np.array(ms), np.array(its), xs

    nb = min(args.base_cases, ncases)
    ms64, it64, x64 = per_case(oa.cg, range(nb))
    log("[bench4] fp64 zero start: %.1f ms/case, %.0f it" % (ms64.mean(), it64.mean()))
    ms32, it32, x32 = per_case(oa.cg_mixed, range(nb))
    log("[bench4] mixed zero start: %.1f ms/case, %.0f it" % (ms32.mean(), it32.mean()))

    def all_cases(mixed, net=None, collect=0, tag="", prec=oa.DIAG_PREC, upto=None):
        proj = oa.ProjectionGuess(A, m=16, minres=True)
        x0 = torch.empty(n, dtype=torch.float64, device="cuda")
        xg = torch.empty_like(x0)
        its, samples, sols = [], [], []
        fn = oa.cg_mixed if mixed else oa.cg
        nc = ncases if upto is None else min(upto, ncases)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for c in range(nc):
            b = Bc[c]
            if net is None:
                proj.guess(b, x0)
            else:
                proj.guess(b, xg)
                net.guess(xg, xg, b, x0)
            if c < collect:
                g_ = x0.clone()
            # a start that passes toler ends the case; any other case is
            # solved to toler/100 and joins the projection basis
            info = oa.solve_with_projection(A, b, x0, proj, fn=fn, prec=prec, toler=tol,
                                            tighten=1e-2, max_it=kw["max_it"], guess=False)
            its.append(info.iterations if info.nb_it != 0 else 0)
            if c < collect:
                samples.append((g_, g_, b, x0.clone()))
            if c < nb:
                sols.append(x0.clone())
            if c % 32 == 31 or c < 2:
                log("[bench4]   %s case %d: %d it (%.1f s)" % (tag, c, its[-1], time.perf_counter() - t))
        torch.cuda.synchronize()
        dt_ = time.perf_counter() - t
        proj.close()
        return dt_ * 1e3 / nc, np.array(its), samples, sols

    res = {}
    ms, its, samples, sols = all_cases(False, collect=args.train_cases, tag="fp64 proj")
    res["fp64_proj"] = (ms, its, sols)
    log("[bench4] fp64 projection: %.2f ms/case, it first %s, rest mean %.1f"
        % (ms, its[:10].tolist(), its[10:].mean()))
    ms, its, _, sols = all_cases(True, tag="mixed proj")
    res["mixed_proj"] = (ms, its, sols)
    log("[bench4] mixed projection: %.2f ms/case, rest mean %.1f" % (ms, its[10:].mean()))
    net = oa.LearnedGuess(A, layers=4, seed=0)
    x0buf = torch.empty(n, dtype=torch.float64, device="cuda")
    rng = np.random.default_rng(0)
    t = time.perf_counter()
    hist = []
    for ep in range(args.epochs):
        tot = 0.0
        for i in rng.permutation(len(samples)):
            tot += net.train_step(*samples[i], x0buf, lr=args.lr)
        hist.append(round(tot / max(len(samples), 1), 5))
    train_s = time.perf_counter() - t
    log("[bench4] training on %d cases: %s (%.1fs)" % (len(samples), hist, train_s))
    ms, its, _, sols = all_cases(True, net=net, tag="mixed proj+net")
    res["mixed_proj_learned"] = (ms, its, sols)
    log("[bench4] mixed projection+learned: %.2f ms/case, rest mean %.1f" % (ms, its[10:].mean()))

    def err(xa, xb):
        return max(float(torch.linalg.norm(a - b_) / torch.linalg.norm(b_)) for a, b_ in zip(xa, xb))

    errs = {"mixed_zero": err(x32, x64)}
    for k, v in res.items():
        errs[k] = err(v[2], x64)

    # SSOR_PREC, the preconditioner the reference's elasticity run names: the
    # exact sweeps keep the reference's sequential order, so each application
    # walks the matrix's level chain (8,702 levels per sweep on 2048 x 256)
    ssor = None
    if args.ssor_cases or args.ssor_proj_cases:
        ssor = {"ms_per_application": round(ssor_ms, 2),
                "cg_ms_per_iteration": round(cg_ssor_it_ms, 2)}
        if args.ssor_cases:
            mss, its_s, xs_s = per_case(oa.cg, range(min(args.ssor_cases, ncases)), prec=oa.SSOR_PREC)
            ssor.update({"fp64_zero_ms_per_case": round(float(mss.mean()), 1),
                         "fp64_zero_iterations": float(its_s.mean()),
                         "fp64_zero_cases": len(its_s),
                         "max_rel_err_vs_fp64_diag_zero": err(xs_s, x64[:len(xs_s)])})
            log("[bench4] fp64 CG+SSOR zero start: %.1f ms/case, %.0f it"
                % (mss.mean(), its_s.mean()))
        if args.ssor_proj_cases:
            # the regime of all but the first few cases: a basis of the first
            # nb solutions (from the FP64 DIAG projection run), then each of the
            # next cases from its projected start, SSOR and DIAG alike
            basis = res["fp64_proj"][2]
            cases = range(len(basis), min(len(basis) + args.ssor_proj_cases, ncases))

            def proj_run(prec):
                proj = oa.ProjectionGuess(A, m=16, minres=True)
                for xb in basis:
                    proj.update(xb)
                x0 = torch.empty(n, dtype=torch.float64, device="cuda")
                its = []
                torch.cuda.synchronize()
                t = time.perf_counter()
                for c in cases:
                    proj.guess(Bc[c], x0)
                    info = oa.solve_with_projection(A, Bc[c], x0, proj, fn=oa.cg, prec=prec,
                                                    toler=tol, tighten=1e-2,
                                                    max_it=kw["max_it"], guess=False)
                    its.append(info.iterations if info.nb_it != 0 else 0)
                    log("[bench4]   %s projected case %d: %d it (%.1f s)"
                        % ("ssor" if prec == oa.SSOR_PREC else "diag", c, its[-1],
                           time.perf_counter() - t))
                torch.cuda.synchronize()
                dt_ = time.perf_counter() - t
                proj.close()
                return dt_ * 1e3 / max(len(cases), 1), its

            ms_sp, its_sp = proj_run(oa.SSOR_PREC)
            ms_dp, its_dp = proj_run(oa.DIAG_PREC)
            ssor.update({"proj_cases": "%d..%d, basis of the first %d solutions"
                                       % (cases.start, cases.stop - 1, len(basis)),
                         "fp64_proj_ms_per_case": round(ms_sp, 1), "fp64_proj_iterations": its_sp,
                         "diag_fp64_proj_ms_per_case": round(ms_dp, 1),
                         "diag_fp64_proj_iterations": its_dp})
            log("[bench4] fp64 CG+SSOR projected, cases %d..%d: %.1f ms/case (DIAG %.1f)"
                % (cases.start, cases.stop - 1, ms_sp, ms_dp))

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        O.lib(fast=True)
        rp, col = A.pattern()
        vals = A.values()
        iters = args.cpu_iters
        t = time.perf_counter()
        O.cg(rp.astype(np.int64), col.astype(np.int64) + 1, vals, Bc[0].cpu().numpy(),
             prec=O.DIAG, max_it=iters, toler=1e-300, fast=True)
        cpu_s = time.perf_counter() - t
        per_it = cpu_s / iters
        cpu = {"value": round(per_it * it64.mean() * 1e3, 1),
               "unit": "ms per FP64 CG+DIAG solve from zero (extrapolated)", "cores": 1,
               "kind": "port",
               "sample": "%d CG+DIAG iterations of case 0 by the oracle -Ofast on 1 host core, "
                         "%.1f ms/iteration x %.0f GPU iterations" % (iters, per_it * 1e3, it64.mean())}
        log("[bench4] cpu %.1f ms/it" % (per_it * 1e3))
        if ssor and "fp64_zero_iterations" in ssor:
            k = max(2, iters // 6)
            t = time.perf_counter()
            O.cg(rp.astype(np.int64), col.astype(np.int64) + 1, vals, Bc[0].cpu().numpy(),
                 prec=O.SSOR, max_it=k, toler=1e-300, fast=True)
            per_s = (time.perf_counter() - t) / k
            cpu["ssor"] = {"ms_per_iteration": round(per_s * 1e3, 1),
                           "ms_per_solve_extrapolated": round(per_s * ssor["fp64_zero_iterations"] * 1e3),
                           "sample": "%d CG+SSOR iterations of case 0 by the oracle -Ofast on 1 core "
                                     "x %.0f GPU iterations" % (k, ssor["fp64_zero_iterations"])}
            log("[bench4] cpu CG+SSOR %.1f ms/it" % (per_s * 1e3))

    head = "mixed_proj_learned"
    base_ms = float(ms64.mean())
    out = {
        "metric": METRIC, "value": round(res[head][0], 3),
        "unit": "ms per load case (projected + learned start, mixed CG+DIAG to rtol %g), "
                "mean over %d cases" % (tol, ncases),
        "n_gpus": 1, "steps": ncases, "warmup": 0, "ms_per_step": round(res[head][0], 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64+bf16",
        "data": "synthetic",
        "config": {"workload": "Elas2DQ4 cantilever %dx%d quads, plane stress, %d load cases"
                               % (args.nx4, args.ny4, ncases), "rows": n, "nnz": nnz,
                   "toler": tol, "precond": "DIAG_PREC", "train_cases": args.train_cases},
        "baseline": "fp64 CG+DIAG from zero per case (first %d cases)" % nb,
        "fp64_zero_ms_per_case": round(base_ms, 2), "fp64_zero_iterations": float(it64.mean()),
        "mixed_zero_ms_per_case": round(float(ms32.mean()), 2),
        "mixed_zero_iterations": float(it32.mean()),
        "variants_ms_per_case": {k: round(v[0], 3) for k, v in res.items()},
        "variants_iterations_first10": {k: v[1][:10].tolist() for k, v in res.items()},
        "variants_mean_iterations_after10": {k: round(float(v[1][10:].mean()), 2) for k, v in res.items()},
        "speedup_vs_fp64_zero": {k: round(base_ms / v[0], 2) for k, v in res.items()},
        "max_rel_err_vs_fp64_zero_first_cases": errs,
        "ssor_ms_per_application": round(ssor_ms, 2), "cg_ssor_ms_per_iteration": round(cg_ssor_it_ms, 2),
        "assembly_ms": round(asm_ms, 3), "train_seconds": round(train_s, 2), "train_loss": hist,
        "ssor_prec": ssor,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
