#!/usr/bin/env python3
"""Benchmark of the sparse solve path on one MI355X.

Workload (default): 3-D 7-point Poisson on 128^3 (n = 2,097,152,
nnz = 14,581,760), x_true ~ U[-1,1] (PCG64 seed 0), b = A x_true, float64
CSR, stopping test ||r|| / ||b|| < 1e-10.

A step is one complete solve of A x = b with the preconditioner already
built, b and x resident in device memory. The headline method is PCG with
the smoothed-aggregation AMG V-cycle (2 pre / 2 post symmetric multicolour
Gauss-Seidel sweeps, weight 1.0) in float32 inside float64 iterative
refinement; plain float64 IC(0)-PCG, float64 SA-AMG-PCG and float64
Jacobi-PCG are timed beside it (one step each). Every result is checked:
true float64 relative residual and relative error against x_true.

Also reported: the float64 CSR SpMV on the same matrix (HIP events on the
library's stream) as the roofline kernel, and the CPU oracle's plain
IC(0)-PCG on this host as the CPU baseline.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cpp-fem-solver_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def spmv_bytes(nnz, n, m, s=8):
    return nnz * (s + 4) + 4 * (n + 1) + m * s + n * s


def measured_traffic():
    """HBM bytes per float64 SpMV launch as read by rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy), from the committed
    profile summary; None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "spmv_pmc_poisson3d_128.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1].strip())
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128, help="grid side of the 3-D Poisson workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--methods", default="all", help="comma list: saamg_mp,saamg_fp64,ic0_fp64,jacobi_fp64")
    ap.add_argument("--spmv-probe", action="store_true",
                    help="only run a 512 MiB device copy (counter calibration) and the float64 SpMV; "
                         "for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import femgpu as F
    from femgpu import gen

    t0 = time.time()
    N = args.n
    G = gen.poisson3d(N)
    xt, b = gen.rhs(G, 0)
    dev_name, cus = F.device_info()
    A = F.CsrMatrix.from_gen(G)
    bd = F.DeviceArray.from_host(b)
    xd = F.DeviceArray(G.n)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    log(f"[bench] {dev_name} {cus} CUs; workload poisson3d N={N} n={G.n} nnz={G.nnz} gen {time.time() - t0:.1f}s")

    if args.spmv_probe:
        import torch
        src = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        ms = A.bench_spmv(fp32=False, reps=10)
        print(json.dumps({"copy_bytes_each_way": src.numel() * 8, "spmv_bytes": spmv_bytes(G.nnz, G.n, G.m, 8),
                          "spmv_ms": ms}))
        return

    # ---- roofline kernel: float64 CSR SpMV -------------------------------
    spmv_ms = A.bench_spmv(fp32=False, reps=200)
    spmv32_ms = A.bench_spmv(fp32=True, reps=200)
    B64 = spmv_bytes(G.nnz, G.n, G.m, 8)
    B32 = spmv_bytes(G.nnz, G.n, G.m, 4)
    spmv_gbs = B64 / (spmv_ms * 1e-3) / 1e9
    spmv32_gbs = B32 / (spmv32_ms * 1e-3) / 1e9
    log(f"[bench] spmv fp64 {spmv_ms * 1e3:.1f} us {spmv_gbs:.0f} GB/s; fp32 {spmv32_ms * 1e3:.1f} us {spmv32_gbs:.0f} GB/s")

    sa = dict(smoother=F.SMOOTH_SGS, npre=2, npost=2, omega=1.0, p_omega=0.0, coarse_exact=1)
    methods = {
        "saamg_mp": ("saamg", sa, F.FP32_IR),
        "saamg_fp64": ("saamg", sa, F.FP64),
        "ic0_fp64": ("ic0", {}, F.FP64),
        "jacobi_fp64": ("diag", {}, F.FP64),
    }
    want = list(methods) if args.methods == "all" else args.methods.split(",")
    results = {}
    headline = None
    for name in want:
        kind, prm, prec = methods[name]
        t = time.time()
        M = F.Preconditioner(A, kind, fp32=int(prec == F.FP32_IR), **prm)
        setup_ms = (time.time() - t) * 1e3
        S = F.Solver(A, M, solver="pcg", precision=prec, max_iter=5000, device_ptrs=True, inner_tol=1e-5)
        is_head = name == "saamg_mp"
        warm = args.warmup if is_head else 1
        steps = args.steps if is_head else 1
        for _ in range(warm):
            S.solve_device(bd.address, xd.address)
        times = []
        st = None
        for _ in range(steps):
            t = time.perf_counter()
            st = S.solve_device(bd.address, xd.address)
            times.append((time.perf_counter() - t) * 1e3)
        x = xd.download()
        relres = float(np.linalg.norm(b - G.matvec(x)) / np.linalg.norm(b))
        err = float(np.linalg.norm(x - xt) / np.linalg.norm(xt))
        info = M.info()
        results[name] = {
            "solve_ms": float(np.median(times)), "setup_ms": round(setup_ms, 2),
            "iterations": st["iterations"], "refinements": st["refinements"], "relres": relres, "rel_err": err,
            "levels": info["levels"],
        }
        log(f"[bench] {name}: setup {setup_ms:.1f} ms solve {np.median(times):.2f} ms it {st['iterations']} "
            f"relres {relres:.2e} err {err:.2e}")
        if is_head:
            headline = (float(np.mean(times)), times)
        S.close()
        M.close()

    # ---- CPU baseline: the oracle's float64 IC(0)-PCG on this host -------
    cpu = None
    if not args.no_cpu:
        import oracle_api as O
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        O.set_threads(threads)
        Ao = O.Csr.from_gen(G)
        t = time.time()
        pc = O.Pc(Ao, O.K_IC0)
        xo, ito = O.solve(Ao, 0, pc, b, tol=1e-10, max_iter=5000)
        cpu_ms = (time.time() - t) * 1e3
        cpu = {"value": round(cpu_ms, 1), "unit": "ms per solve (setup + iterations)", "cores": threads,
               "kind": "port",
               "sample": f"one full IC(0)-PCG solve of the 128^3 Poisson system to 1e-10 ({ito + 1} iterations); "
                         f"SpMV/dot on {threads} OpenMP threads, triangular sweeps serial as in the reference"}
        log(f"[bench] cpu oracle ic0-pcg {cpu_ms:.0f} ms it {ito}")

    if headline is None:
        headline = (results[want[0]]["solve_ms"], [results[want[0]]["solve_ms"]])
    value = headline[0]
    out = {
        "metric": "solve time to 1e-10 rel-resid (learned/MP vs plain FP64), SpMV HBM GB/s",
        "value": round(value, 3),
        "unit": "ms per solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(value, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-refinement",
        "data": "synthetic (x_true ~ U[-1,1], PCG64 seed 0; b = A x_true)",
        "config": {"workload": f"3-D 7-point Poisson {N}^3 (n={G.n}, nnz={G.nnz}), PCG + SA-AMG "
                               "(SGS 2/2, w=1.0) FP32 inner / FP64 refinement, tol 1e-10",
                   "methods": results,
                   "spmv_fp32_us": round(spmv32_ms * 1e3, 2), "spmv_fp32_gbs": round(spmv32_gbs, 1),
                   "device": dev_name},
        "roofline": {"bound": "hbm", "kernel": "k_spmv<double,0> (CSR-stream SpMV, float64)",
                     "achieved": round(spmv_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(spmv_gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                     "bytes_per_launch": B64, "avg_launch_us": round(spmv_ms * 1e3, 2)},
        "cpu_baseline": cpu,
    }
    if results.get("ic0_fp64") and results.get("saamg_mp"):
        out["config"]["speedup_mp_vs_fp64_ic0"] = round(results["ic0_fp64"]["solve_ms"] / results["saamg_mp"]["solve_ms"], 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
