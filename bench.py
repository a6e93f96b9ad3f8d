"""Benchmark of the sparse linear-solve path on one MI355X.

Workload: the 3D variable-coefficient Poisson problem (unit cube, 96^3 cells,
6 Kuhn tets per cell: 857,375 free DOFs, 12,645,163 nonzeros; k_e lognormal
from a 64-mode Gaussian field; f = 1; Dirichlet 0), assembled on the GPU and
solved to rtol 1e-8 with CG + DIAG_PREC.  One step = one solve from x0 = 0.
`value` is the wall time per solve of the mixed-precision CG (FP32 inner,
FP64 residual replacement, FP64-verified); the plain FP64 CG time, the
speedup and the SpMV roofline are reported beside it.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 96]
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
    ap.add_argument("--cpu-iters", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(see the sharding design in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

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

    # SpMV roofline (dominant kernel of every CG iteration)
    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    for _ in range(5):
        oa.spmv(A, x, y)
    ev0.record()
    for _ in range(args.spmv_reps):
        oa.spmv(A, x, y)
    ev1.record()
    torch.cuda.synchronize()
    spmv_ms = ev0.elapsed_time(ev1) / args.spmv_reps
    b_spmv64 = 12 * nnz + 4 * (n + 1) + 16 * n
    spmv_gbs = b_spmv64 / (spmv_ms * 1e-3) / 1e9
    log("[bench] spmv64 %.4f ms  %.0f GB/s (32-bit columns)" % (spmv_ms, spmv_gbs))
    A.compress_columns(True)
    ev0.record()
    for _ in range(args.spmv_reps):
        oa.spmv(A, x, y)
    ev1.record()
    torch.cuda.synchronize()
    spmv_c16_ms = ev0.elapsed_time(ev1) / args.spmv_reps
    A.compress_columns(False)
    log("[bench] spmv64 %.4f ms (16-bit column offsets)" % spmv_c16_ms)
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
    xf, yf = x.float(), y.float()
    oa.spmv_f32(A, xf, yf)
    ev0.record()
    for _ in range(args.spmv_reps):
        oa.spmv_f32(A, xf, yf)
    ev1.record()
    torch.cuda.synchronize()
    spmv32_ms = ev0.elapsed_time(ev1) / args.spmv_reps
    b_spmv32 = 8 * nnz + 4 * (n + 1) + 8 * n
    log("[bench] spmv32 %.4f ms  %.0f GB/s" % (spmv32_ms, b_spmv32 / (spmv32_ms * 1e-3) / 1e9))

    def solve(fn):
        xs = torch.zeros(n, dtype=torch.float64, device="cuda")
        info = fn(A, b, xs, prec=oa.DIAG_PREC, max_it=20000, toler=args.toler)
        return info, xs

    for _ in range(args.warmup):
        solve(oa.cg)
        solve(oa.cg_mixed)
    t64, t32 = [], []
    for k in range(args.steps):
        torch.cuda.synchronize()
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
    cg_iter_us = ms64 * 1e3 / max(it64, 1)
    b_cg64_iter = b_spmv64 + 88 * n

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        vals = A.values()
        rp, cl = A.pattern()
        rp64 = rp.astype(np.int64)
        ci64 = cl.astype(np.int64) + 1
        bh = b.cpu().numpy()
        O.lib(fast=True)
        t = time.perf_counter()
        O.cg(rp64, ci64, vals, bh, prec=O.DIAG, max_it=args.cpu_iters, toler=1e-300, fast=True)
        cpu_s = time.perf_counter() - t
        per_it = cpu_s / args.cpu_iters
        cpu = {"value": per_it * it64 * 1e3, "unit": "ms per FP64 CG+DIAG solve (extrapolated)",
               "cores": 1, "kind": "port",
               "sample": "%d CG+DIAG iterations of this system on 1 host core (oracle -Ofast), "
                         "%.1f ms/iteration x %d GPU iterations" % (args.cpu_iters, per_it * 1e3, it64)}
        log("[bench] cpu %.1f ms/it" % (per_it * 1e3))

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
        "plan_setup_s": round(setup_s, 3), "cg64_us_per_iter": round(cg_iter_us, 2),
        "cg64_iter_GBs": round(b_cg64_iter / (cg_iter_us * 1e-6) / 1e9, 1),
        "spmv32_ms": round(spmv32_ms, 4), "spmv64_col16_ms": round(spmv_c16_ms, 4),
        "device_copy_GBs": round(copy_gbs, 1),
        "roofline": {"bound": "hbm", "achieved": round(spmv_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(spmv_gbs / HBM_PEAK_GBS, 4),
                     # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE runs recorded in
                     # profiles/spmv_cg_poisson3d_96cube.md (96^3 system only)
                     "traffic": 171.4e6 if args.n == 96 else None,
                     "kernel": "k_spmv<double> (CSR-stream FP64)",
                     "bytes_per_launch": b_spmv64, "avg_ms": round(spmv_ms, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
