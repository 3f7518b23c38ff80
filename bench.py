#!/usr/bin/env python3
"""Benchmark of the sparse solve path on one MI355X.

Workload (default): 3-D 7-point Poisson on 128^3 (n = 2,097,152,
nnz = 14,581,760), x_true ~ U[-1,1] (PCG64 seed 0), b = A x_true, float64
CSR, stopping test ||r|| / ||b|| < 1e-10.

A step is one complete solve of A x = b with the preconditioner already
built, b and x resident in device memory. The headline method is PCG with
the smoothed-aggregation AMG V-cycle (2 pre / 2 post symmetric multicolour
Gauss-Seidel sweeps, weight 1.0, coarsening down to at most 300 rows, whose
dense exact solve ends the V-cycle) in float32 inside float64 iterative
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


def spmv_bytes_packed(nnz, n, m, s=8):
    """Bytes the CSR-stream kernel moves with 16-bit packed column offsets
    (2 bytes per entry plus a 4-byte base per 64 entries)."""
    return nnz * (s + 2) + 4 * ((nnz + 63) // 64) + 4 * (n + 1) + m * s + n * s


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


def c4_leg(F, gen, args, log):
    """Learned SPAI on the variable-coefficient 96^3 diffusion system
    (kappa = exp(g), g a Gaussian random field, sigma 1, correlation length 8
    cells). The network is trained on the GPU on `c4_train` seeded fields of
    32^3 cells (same field statistics; the network only sees 7-point
    stencils) and tested on `c4_test` held-out 96^3 seeds: learned SPAI PCG
    in float32-inner/float64-refinement form against plain float64 IC(0)-PCG
    (every 4th seed), float64 Jacobi-PCG and SA-AMG PCG. Solve times are
    single timed solves after one warm-up, b and x in HBM."""
    from femgpu import learned as Lr
    t0 = time.time()
    fields = Lr.make_fields(32, range(args.c4_train))
    gen_s = time.time() - t0
    net = F.Mlp(seed=0)
    t0 = time.time()
    for f in fields:
        f.errors = Lr.error_vectors(f) if args.c4_errors == "ic0" else np.empty((0, f.G.n))
    err_s = time.time() - t0
    steps = 3 * len(fields)
    hist, train_s = Lr.train(net, fields, steps=steps)
    del fields
    log(f"[bench] c4 training: {args.c4_train} fields, {steps} Adam steps in {train_s:.2f} s "
        f"(field generation {gen_s:.1f} s, error vectors {err_s:.1f} s), loss {hist[:10].mean():.3f} -> {hist[-50:].mean():.3f}")
    N = 96
    sa = dict(smoother=F.SMOOTH_SGS, npre=2, npost=2, omega=1.0, p_omega=0.0, coarse_exact=1, fp32=1)
    rows = {k: [] for k in ("learned_mp", "learned_fp64", "jacobi_fp64", "ic0_fp64", "saamg_mp")}
    want = list(rows) if args.c4_methods == "all" else args.c4_methods.split(",")
    worst = {"relres": 0.0, "rel_err": 0.0}
    fallbacks = 0
    for t, seed in enumerate(range(256, 256 + args.c4_test)):
        G = gen.vardiff3d(N, seed)[0]
        xt, b = gen.rhs(G, seed)
        A = F.CsrMatrix.from_gen(G)
        bd = F.DeviceArray.from_host(b)
        xd = F.DeviceArray(G.n)

        def run(name, M, prec):
            S = F.Solver(A, M, solver="pcg", precision=prec, max_iter=20000, device_ptrs=True, tol=1e-10)
            S.solve_device(bd.address, xd.address)
            t1 = time.perf_counter()
            st = S.solve_device(bd.address, xd.address)
            ms = (time.perf_counter() - t1) * 1e3
            x = xd.download()
            err = float(np.linalg.norm(x - xt) / np.linalg.norm(xt))
            worst["relres"] = max(worst["relres"], st["true_relres"])
            worst["rel_err"] = max(worst["rel_err"], err)
            rows[name].append((seed, ms, st["iterations"]))
            S.close()
            return st

        Ml = F.LearnedPreconditioner(A, (N, N, N), net, fallback="ic0", budget=2000)
        if "learned_mp" in want:
            fallbacks += run("learned_mp", Ml, F.FP32_IR)["fallback"]
        if "learned_fp64" in want:
            run("learned_fp64", Ml, F.FP64)
        if "jacobi_fp64" in want:
            run("jacobi_fp64", F.Preconditioner(A, "diag"), F.FP64)
        if t % 4 == 0:
            if "ic0_fp64" in want:
                run("ic0_fp64", F.Preconditioner(A, "ic0"), F.FP64)
            if "saamg_mp" in want:
                run("saamg_mp", F.Preconditioner(A, "saamg", **sa), F.FP32_IR)
        Ml.close()
        A.close()
    med = {k: (float(np.median([r[1] for r in v])), float(np.median([r[2] for r in v]))) for k, v in rows.items() if v}
    ic0_seeds = [r[0] for r in rows["ic0_fp64"]]
    lm = {r[0]: r[1] for r in rows["learned_mp"]}
    speed = float(np.median([r[1] / lm[r[0]] for r in rows["ic0_fp64"]])) if ic0_seeds and lm else None
    for k, (ms, it) in med.items():
        log(f"[bench] c4 {k}: median {ms:.2f} ms, {it:.0f} iterations")
    if speed:
        log(f"[bench] c4 learned MP vs FP64 IC0-PCG: {speed:.1f}x (median over seeds {ic0_seeds})")
    return {"workload": "3-D log-normal diffusion 96^3 (n=884736), sigma 1, correlation length 8 cells, tol 1e-10",
            "train": {"fields": args.c4_train, "field_cells": "32^3", "steps": steps, "seconds": round(train_s, 2),
                      "loss_first": round(float(hist[:10].mean()), 4), "loss_last": round(float(hist[-50:].mean()), 4)},
            "test_seeds": f"256..{255 + args.c4_test}",
            "median_ms": {k: round(v[0], 3) for k, v in med.items()},
            "median_iterations": {k: v[1] for k, v in med.items()},
            "speedup_learned_mp_vs_ic0_fp64": round(speed, 2) if speed else None,
            "fallbacks": int(fallbacks), "worst_true_relres": worst["relres"], "worst_rel_err": worst["rel_err"]}


def c3_leg(F, gen, args, log):
    """Nonsymmetric 2-D upwind convection-diffusion, 1024^2 cells, cell Peclet
    10, velocity (cos 2 pi y, sin 2 pi x): BiCGSTAB and GMRES(30) with the
    level-scheduled ILU(0) in float64, tol 1e-10, b and x in HBM; one timed
    solve each after the preconditioner is built."""
    G = gen.convdiff2d(args.c3_n)
    xt, b = gen.rhs(G, 0)
    A = F.CsrMatrix.from_gen(G)
    bd = F.DeviceArray.from_host(b)
    xd = F.DeviceArray(G.n)
    t = time.time()
    M = F.Preconditioner(A, "ilu0")
    setup_ms = (time.time() - t) * 1e3
    out = {"workload": f"2-D upwind convection-diffusion {args.c3_n}^2 (n={G.n}, nnz={G.nnz}), cell Peclet 10, "
                       "ILU(0), float64, tol 1e-10",
           "ilu0_setup_ms": round(setup_ms, 1), "ilu0_levels": M.info()["levels"]}
    if args.c3_applies > 0:
        # only preconditioner applications (for kernel counters), no solves
        r = np.random.default_rng(0).standard_normal(G.n)
        for _ in range(args.c3_applies):
            M.apply(r)
        out["applies"] = args.c3_applies
        return out
    for name, solver in (("bicgstab_ilu0", "bicgstab"), ("gmres30_ilu0", "gmres")):
        S = F.Solver(A, M, solver=solver, precision=F.FP64, max_iter=args.c3_max_iter, restart=30, device_ptrs=True,
                     tol=1e-10)
        t = time.perf_counter()
        st = S.solve_device(bd.address, xd.address)
        ms = (time.perf_counter() - t) * 1e3
        x = xd.download()
        err = float(np.linalg.norm(x - xt) / np.linalg.norm(xt))
        out[name] = {"solve_ms": round(ms, 1), "iterations": st["iterations"], "converged": st["converged"],
                     "relres": st["true_relres"], "rel_err": err}
        log(f"[bench] c3 {name}: {ms:.0f} ms, {st['iterations']} iterations, converged {st['converged']}, "
            f"relres {st['true_relres']:.2e}")
        S.close()
    return out


def rows_leg(F, gen, args, log):
    """The other rows of the solve path, one GPU, FP64 unless stated: one
    timed solve per method after a warm-up solve (preconditioner built, b and
    x in HBM; block-matrix solves: the library's own timer around the solve,
    which for a new solver object includes capturing its graphs), with
    iterations, true relative residual and setup time; the block matvec's
    average time over 50 device calls (each synchronised)."""
    out = {}

    def csr_solve(name, A, b, kind, solver="pcg", prec=F.FP64, restart=30, max_iter=20000, **prm):
        bd = F.DeviceArray.from_host(b)
        xd = F.DeviceArray(A.rows)
        t = time.time()
        M = F.Preconditioner(A, kind, fp32=int(prec == F.FP32_IR), **prm) if kind != "none" else None
        setup_ms = (time.time() - t) * 1e3
        S = F.Solver(A, M, solver=solver, precision=prec, max_iter=max_iter, device_ptrs=True, restart=restart)
        S.solve_device(bd.address, xd.address)
        F.device_synchronize()
        t = time.perf_counter()
        st = S.solve_device(bd.address, xd.address)
        F.device_synchronize()
        ms = (time.perf_counter() - t) * 1e3
        out[name] = {"solve_ms": round(ms, 2), "setup_ms": round(setup_ms, 1), "iterations": st["iterations"],
                     "relres": st["true_relres"]}
        log(f"[bench] rows {name}: setup {setup_ms:.1f} ms solve {ms:.2f} ms it {st['iterations']} "
            f"relres {st['true_relres']:.2e}")
        S.close()
        if M is not None:
            M.close()

    G = gen.poisson3d(128)
    _, b = gen.rhs(G, 0)
    A = F.CsrMatrix.from_gen(G)
    csr_solve("poisson128_ssor_pcg", A, b, "ssor", omega=1.0)
    csr_solve("poisson128_ic0_pcg_mp", A, b, "ic0", prec=F.FP32_IR)
    A.close()
    # C/F AMG (host-built hierarchy)
    G = gen.poisson3d(64)
    _, b = gen.rhs(G, 0)
    A = F.CsrMatrix.from_gen(G)
    csr_solve("poisson64_cf_amg_pcg", A, b, "amg")
    A.close()
    G = gen.convdiff2d(256)
    _, b = gen.rhs(G, 0)
    A = F.CsrMatrix.from_gen(G)
    csr_solve("convdiff256_bicgstab_ilu0", A, b, "ilu0", solver="bicgstab")
    csr_solve("convdiff256_gmres30_ilu0", A, b, "ilu0", solver="gmres", restart=30)
    csr_solve("convdiff256_bicgstab_ilu0_mp", A, b, "ilu0", solver="bicgstab", prec=F.FP32_IR)
    A.close()
    # block::BlockMatrix: 3x3 blocks from the reference test's generator, 128^2 block rows
    S = gen.block_test_scalar(128)
    Ac = F.CsrMatrix.from_gen(S)
    B = F.BsrMatrix.condense(Ac, 3, 128 * 128, 128 * 128)
    x = gen.drand(np.arange(S.n))
    xd, yd = F.DeviceArray.from_host(x), F.DeviceArray(S.n)
    B.matvec_device(xd.address, yd.address)
    t = time.perf_counter()
    for _ in range(50):
        B.matvec_device(xd.address, yd.address)
    mv_us = (time.perf_counter() - t) / 50 * 1e6
    byt = B.nnzb * (72 + 4) + 4 * (B.nbr + 1) + 2 * 8 * S.n
    out["bsr3_matvec_49152"] = {"us_per_call_incl_sync": round(mv_us, 1), "nnzb": int(B.nnzb), "bytes": int(byt)}
    log(f"[bench] rows bsr3 matvec ({B.nbr} block rows, {B.nnzb} blocks): {mv_us:.1f} us per synchronised call")
    bb = 1.0 + 0.5 * gen.drand(np.arange(S.n) + 7)
    for name, kind, solver in (("bsr3_bjacobi_minicg", F.PC_BJACOBI, "minicg"),
                               ("bsr3_bilu0_gmres30", F.PC_BILU0, "gmres")):
        M = F.Preconditioner(B, kind)
        F.solve(B, M, bb, solver=solver, tol=1e-8, max_iter=3000, restart=30)
        _, st = F.solve(B, M, bb, solver=solver, tol=1e-8, max_iter=3000, restart=30)
        out[name] = {"solve_ms_library_timer": round(st["solve_ms"], 2), "iterations": st["iterations"],
                     "relres": st["true_relres"]}
        log(f"[bench] rows {name}: {st['solve_ms']:.2f} ms (library timer) it {st['iterations']} "
            f"relres {st['true_relres']:.2e}")
        M.close()
    B.close()
    Ac.close()
    return out


def c5_leg(F, gen, args, log):
    """Batch of independent SPD systems (half 2-D Poisson, side U{32..256};
    half 3x3-block 7-point systems, side U{16..48}; 4 right-hand sides each;
    seed 0) solved by block-Jacobi PCG in float32 inside float64 refinement
    in one persistent launch, schedule from the autotuner; device time of the
    launch, every right-hand side checked against tol 1e-10."""
    t0 = time.time()
    sysl = gen.batch_c5(0, count=args.c5_count)
    gen_s = time.time() - t0
    B = F.Batch([s["A"] for s in sysl], nrhs=4)
    B.set_rhs([s["B"] for s in sysl])
    t0 = time.time()
    best, trials = B.tune(divisor=8)
    tune_s = time.time() - t0
    B.set_schedule(**best)
    B.solve()
    times = []
    for _ in range(3):
        st, ms = B.solve()
        times.append(ms)
    X = B.solution()
    worst = max(s["max_relres"] for s in st)
    conv = sum(s["converged"] for s in st)
    err = max(float(np.linalg.norm(x - s["X"]) / np.linalg.norm(s["X"])) for x, s in zip(X, sysl))
    its = [s["iterations"] for s in st]
    unknowns = sum(int(x.size) for x in X)
    log(f"[bench] c5: {len(sysl)} systems ({unknowns} unknowns x rhs), schedule {best}, "
        f"solve {np.median(times):.1f} ms, {conv}/{4 * len(sysl)} converged, worst relres {worst:.2e}")
    B.close()
    return {"workload": f"{len(sysl)} SPD systems (half 2-D Poisson side 32..256, half 3x3-block 7-point side 16..48), "
                        "4 rhs each, block-Jacobi PCG FP32 inner / FP64 refinement, tol 1e-10, one persistent launch",
            "solve_ms_median": round(float(np.median(times)), 2), "unknowns_times_rhs": unknowns,
            "schedule": best, "tune_seconds": round(tune_s, 2), "trial_ms": [round(float(t), 2) for t in trials],
            "converged": int(conv), "rhs_total": 4 * len(sysl), "worst_true_relres": worst,
            "worst_rel_err": err, "inner_iterations": {"min": int(min(its)), "median": float(np.median(its)),
                                                       "max": int(max(its))},
            "generation_seconds": round(gen_s, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128, help="grid side of the 3-D Poisson workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--methods", default="all",
                    help="comma list: saamg_mp,saamg_fp64,ic0_fp64,jacobi_fp64 ('none': no 128^3 solves)")
    ap.add_argument("--no-c4", action="store_true", help="skip the learned-preconditioner leg")
    ap.add_argument("--c4-train", type=int, default=256, help="training fields (32^3 cells each)")
    ap.add_argument("--c4-test", type=int, default=32, help="held-out 96^3 fields")
    ap.add_argument("--c4-errors", choices=("ic0", "none"), default="ic0",
                    help="training probes besides random signs: IC(0)-PCG error vectors, or none")
    ap.add_argument("--c4-methods", default="all",
                    help="comma list of 96^3 comparisons: learned_mp,learned_fp64,jacobi_fp64,ic0_fp64,saamg_mp")
    ap.add_argument("--no-c5", action="store_true", help="skip the batched-solve leg")
    ap.add_argument("--c3", action="store_true", help="add the convection-diffusion ILU(0) leg (about 30 s)")
    ap.add_argument("--c3-n", type=int, default=1024, help="grid side of the convection-diffusion leg")
    ap.add_argument("--c3-max-iter", type=int, default=3000)
    ap.add_argument("--c3-applies", type=int, default=0,
                    help="only apply the ILU(0) preconditioner this many times (no solves)")
    ap.add_argument("--c5-count", type=int, default=512, help="systems in the batch")
    ap.add_argument("--rows", action="store_true",
                    help="add a timing of the other solve-path rows (SSOR, IC0 mixed precision, C/F AMG, "
                         "BiCGSTAB/GMRES + ILU(0), block matvec, block-Jacobi, block ILU(0))")
    ap.add_argument("--spmv-probe", action="store_true",
                    help="only run a 512 MiB device copy (counter calibration) and the float64 SpMV; "
                         "for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "runs on one GPU only; sharding across GPUs is designed, not implemented",
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
    # each launch alone after 1 GiB of clean reads evicted L2 and the 256 MB
    # Infinity Cache (the back-to-back figure includes its hits: the float64
    # matrix is 183 MB)
    spmv_cold_ms = A.bench_spmv(fp32=False, reps=20, cold=True)
    B64 = spmv_bytes(G.nnz, G.n, G.m, 8)
    packed = A.pack_columns()
    B64_moved = spmv_bytes_packed(G.nnz, G.n, G.m, 8) if packed else B64
    B32 = spmv_bytes(G.nnz, G.n, G.m, 4)
    spmv_gbs = B64 / (spmv_ms * 1e-3) / 1e9
    spmv32_gbs = B32 / (spmv32_ms * 1e-3) / 1e9
    spmv_cold_gbs = B64 / (spmv_cold_ms * 1e-3) / 1e9
    log(f"[bench] spmv fp64 {spmv_ms * 1e3:.1f} us {spmv_gbs:.0f} GB/s (cold {spmv_cold_ms * 1e3:.1f} us "
        f"{spmv_cold_gbs:.0f} GB/s); fp32 {spmv32_ms * 1e3:.1f} us {spmv32_gbs:.0f} GB/s")

    sa = dict(smoother=F.SMOOTH_SGS, npre=2, npost=2, omega=1.0, p_omega=0.0, coarse_exact=1, coarse_size=300)
    methods = {
        "saamg_mp": ("saamg", sa, F.FP32_IR),
        "saamg_fp64": ("saamg", sa, F.FP64),
        "ic0_fp64": ("ic0", {}, F.FP64),
        "jacobi_fp64": ("diag", {}, F.FP64),
    }
    want = list(methods) if args.methods == "all" else [] if args.methods == "none" else args.methods.split(",")
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
        # exactly `steps` solves between two device synchronisations (one
        # process, one GPU: no barrier across ranks is needed)
        times = []
        st = None
        F.device_synchronize()
        t_all = time.perf_counter()
        for _ in range(steps):
            t = time.perf_counter()
            st = S.solve_device(bd.address, xd.address)
            times.append((time.perf_counter() - t) * 1e3)
        F.device_synchronize()
        total_ms = (time.perf_counter() - t_all) * 1e3
        x = xd.download()
        relres = float(np.linalg.norm(b - G.matvec(x)) / np.linalg.norm(b))
        err = float(np.linalg.norm(x - xt) / np.linalg.norm(xt))
        info = M.info()
        results[name] = {
            "solve_ms": float(np.median(times)), "setup_ms": round(setup_ms, 2),
            "iterations": st["iterations"], "refinements": st["refinements"], "relres": relres, "rel_err": err,
            "levels": info["levels"],
        }
        if kind == "saamg":
            L = info["levels"]
            results[name]["level_rows"] = [int(v) for v in info["rows"][:L]]
            results[name]["level_nnz"] = [int(v) for v in info["nnz"][:L]]
            log(f"[bench] {name}: levels rows {results[name]['level_rows']} nnz {results[name]['level_nnz']}")
        pc_bytes = M.apply_bytes(fp32=prec == F.FP32_IR)
        if pc_bytes > 0:
            # one inner PCG iteration: q = A p (with p.q), x += a p and r -= a q
            # (6 vectors), z = M r, r.z (2), p = z + b p (3). With the diagonal
            # the update kernels form r / d themselves: d read in both, r in
            # the p update, no z and no separate r.z. The SA-AMG V-cycle with
            # Gauss-Seidel post-smoothing yields r.z from its last passes. The
            # solve time also holds the float64 refinement steps.
            s_b = 4 if prec == F.FP32_IR else 8
            rz_fused = kind == "saamg" and prm.get("smoother") == F.SMOOTH_SGS and prm.get("npost", 1) >= 1
            vec = (9 if rz_fused else 11) * G.n * s_b + (0 if kind == "diag" else pc_bytes)
            it_bytes = spmv_bytes(G.nnz, G.n, G.m, s_b) + vec
            gbs = st["iterations"] * it_bytes / (np.median(times) * 1e-3) / 1e9
            results[name]["solve_level"] = {
                "bytes_per_iteration": int(it_bytes), "precond_bytes": int(pc_bytes) if kind != "diag" else 0,
                "achieved_GBs": round(gbs, 1), "peak_GBs": 8000.0, "frac": round(gbs / 8000.0, 4)}
            log(f"[bench] {name}: {it_bytes / 1e6:.1f} MB per iteration (preconditioner {pc_bytes / 1e6:.1f} MB), "
                f"solve-level {gbs:.0f} GB/s ({gbs / 80:.1f}% of 8 TB/s)")
        log(f"[bench] {name}: setup {setup_ms:.1f} ms solve {np.median(times):.2f} ms it {st['iterations']} "
            f"relres {relres:.2e} err {err:.2e}")
        if is_head:
            headline = (total_ms / steps, times)
        S.close()
        M.close()

    # ---- learned preconditioner leg: 96^3 log-normal diffusion -----------
    c4 = None if args.no_c4 else c4_leg(F, gen, args, log)

    # ---- batched leg: 512 independent systems in one persistent launch ----
    c5 = None if args.no_c5 else c5_leg(F, gen, args, log)
    c3 = c3_leg(F, gen, args, log) if args.c3 else None
    rows = rows_leg(F, gen, args, log) if args.rows else None

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

    if headline is None and results:
        first = next(iter(results.values()))["solve_ms"]
        headline = (first, [first])
    elif headline is None:
        headline = (None, [])
    value = headline[0]
    out = {
        "metric": "solve time to 1e-10 rel-resid (learned/MP vs plain FP64), SpMV HBM GB/s",
        "value": round(value, 3) if value is not None else None,
        "unit": "ms per solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(value, 3) if value is not None else None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-refinement",
        "data": "synthetic (x_true ~ U[-1,1], PCG64 seed 0; b = A x_true)",
        "config": {"workload": f"3-D 7-point Poisson {N}^3 (n={G.n}, nnz={G.nnz}), PCG + SA-AMG "
                               "(SGS 2/2, w=1.0, coarsest <= 300 rows) FP32 inner / FP64 refinement, tol 1e-10",
                   "methods": results,
                   "spmv_fp32_us": round(spmv32_ms * 1e3, 2), "spmv_fp32_gbs": round(spmv32_gbs, 1),
                   "device": dev_name},
        "roofline": {"bound": "hbm",
                     "kernel": "k_spmv<double,0,%s> (CSR-stream SpMV, float64%s)"
                               % ("true" if packed else "false", ", 16-bit packed column offsets" if packed else ""),
                     "achieved": round(spmv_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(spmv_gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                     "bytes_per_launch": B64, "avg_launch_us": round(spmv_ms * 1e3, 2),
                     "bytes_moved_per_launch": B64_moved,
                     "achieved_moved": round(B64_moved / (spmv_ms * 1e-3) / 1e9, 1),
                     "timing": "200 back-to-back launches, HIP events on the library's stream",
                     "achieved_cold": round(spmv_cold_gbs, 1), "frac_cold": round(spmv_cold_gbs / HBM_PEAK_GBS, 4),
                     "avg_launch_us_cold": round(spmv_cold_ms * 1e3, 2)},
        "cpu_baseline": cpu,
    }
    if c4 is not None:
        out["config"]["c4"] = c4
    if c5 is not None:
        out["config"]["c5"] = c5
    if c3 is not None:
        out["config"]["c3"] = c3
    if rows is not None:
        out["config"]["rows"] = rows
    if results.get("ic0_fp64") and results.get("saamg_mp"):
        out["config"]["speedup_mp_vs_fp64_ic0"] = round(results["ic0_fp64"]["solve_ms"] / results["saamg_mp"]["solve_ms"], 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
