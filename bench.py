#!/usr/bin/env python3
"""Benchmark of the sparse solve path on one MI355X.

Workload (default): 3-D 7-point Poisson on 128^3 (n = 2,097,152,
nnz = 14,581,760), x_k ~ U[-1,1] (PCG64 seed k), b_k = A x_k, float64 CSR,
stopping test ||r|| / ||b|| < 1e-10.

A step is one complete solve of A x = b_k for a right-hand side not solved
before, with the preconditioner already built and b and x resident in device
memory; `value` is the time of the timed steps between two device
synchronisations divided by their number (`solve_ms` per method is the
median of the per-step times). The headline method is
PCG with the smoothed-aggregation AMG V-cycle (2 pre / 2 post symmetric
multicolour Gauss-Seidel sweeps, weight 1.0, coarsening down to at most 300
rows, whose dense exact solve ends the V-cycle) in float32 inside float64
iterative refinement; plain float64 IC(0)-PCG, float64 SA-AMG-PCG and float64
Jacobi-PCG are timed beside it. Per method: setup time, first solve after
the setup, solve time, their sum (the reference's solver call builds its
preconditioner), the headline's time for a repeated right-hand side, the
true float64 relative residual, and the relative error against x_true and
against the oracle's float64 solve of b_0, with the parity bar
cond(A) * 2 * tol.

Also reported: the float64 CSR SpMV on the same matrix (HIP events on the
library's stream) as the roofline kernel, the oracle's float64 PCG with the
same SA-AMG on this host's cores as the CPU baseline (setup included), and
the legs: learned preconditioner (96^3 log-normal diffusion), batch of 512
systems, convection-diffusion 1024^2 with ILU(0), each with its oracle check.

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
    """Learned factorized SPAI (M = G^T G) on the variable-coefficient 96^3
    diffusion system (kappa = exp(g), g a Gaussian random field, sigma 1,
    correlation length 8 cells). The network is trained on the GPU on
    `c4_train` seeded fields of 32^3 cells (same field statistics; the network
    only sees 7-point stencils) and tested on `c4_test` held-out 96^3 seeds:
    learned SPAI PCG in float32-inner/float64-refinement form against plain
    float64 Jacobi-PCG (the judge of the learned solve's iteration count) and,
    every 4th seed, float64 IC(0)-PCG, SA-AMG PCG and the non-learned
    approximate inverses of the same patterns (Neumann, exact FSAI). Solve
    times are single timed solves after one warm-up, b and x in HBM."""
    from femgpu import learned as Lr
    t0 = time.time()
    fields = Lr.make_fields(32, range(args.c4_train))
    gen_s = time.time() - t0
    net = F.Mlp(seed=0)
    t0 = time.time()
    for f in fields:
        f.errors = Lr.error_vectors(f) if args.c4_errors == "solvers" else np.empty((0, f.G.n))
    err_s = time.time() - t0
    steps = 3 * len(fields)
    hist, train_s = Lr.train(net, fields, steps=steps)
    del fields
    log(f"[bench] c4 training: {args.c4_train} fields, {steps} Adam steps in {train_s:.2f} s "
        f"(field generation {gen_s:.1f} s, error vectors {err_s:.1f} s), loss {hist[:10].mean():.3f} -> {hist[-50:].mean():.3f}")
    N = 96
    sa = dict(smoother=F.SMOOTH_SGS, npre=2, npost=2, omega=1.0, p_omega=0.0, coarse_exact=1, fp32=1)
    rows = {k: [] for k in ("learned_mp", "learned_fp64", "jacobi_fp64", "ic0_fp64", "saamg_mp", "neumann_mp",
                            "fsai_mp")}
    want = list(rows) if args.c4_methods == "all" else args.c4_methods.split(",")
    worst = {"relres": 0.0, "rel_err": 0.0}
    fallbacks = 0
    first_x = {}
    oracle_err = {}
    for t, seed in enumerate(range(256, 256 + args.c4_test)):
        G, kappa = gen.vardiff3d(N, seed)
        xt, b = gen.rhs(G, seed)
        A = F.CsrMatrix.from_gen(G)
        bd = F.DeviceArray.from_host(b)
        xd = F.DeviceArray(G.n)

        def run(name, make, prec):
            F.device_synchronize()
            t1 = time.perf_counter()
            M = make()
            F.device_synchronize()
            setup = (time.perf_counter() - t1) * 1e3
            S = F.Solver(A, M, solver="pcg", precision=prec, max_iter=20000, device_ptrs=True, tol=1e-10)
            S.solve_device(bd.address, xd.address)
            t1 = time.perf_counter()
            st = S.solve_device(bd.address, xd.address)
            ms = (time.perf_counter() - t1) * 1e3
            x = xd.download()
            err = float(np.linalg.norm(x - xt) / np.linalg.norm(xt))
            worst["relres"] = max(worst["relres"], st["true_relres"])
            worst["rel_err"] = max(worst["rel_err"], err)
            rows[name].append((seed, ms, st["iterations"], setup))
            if t == 0:
                first_x[name] = x
            S.close()
            return M, st

        # the plain preconditioned solve (Jacobi) judges the learned one: a
        # learned solve needing more iterations than it is a regression and
        # falls back to IC(0)
        judge = None
        if "jacobi_fp64" in want:
            _, stj = run("jacobi_fp64", lambda: F.Preconditioner(A, "diag"), F.FP64)
            judge = stj["inner_iterations"]

        def learned():
            M = F.LearnedPreconditioner(A, (N, N, N), net, fallback="ic0", budget=4000)
            if judge:
                M.reference(judge)
            return M

        Ml = None
        if "learned_mp" in want:
            Ml, st = run("learned_mp", learned, F.FP32_IR)
            fallbacks += st["fallback"]
        if "learned_fp64" in want:
            run("learned_fp64", (lambda: Ml) if Ml is not None else learned, F.FP64)
        if t % 4 == 0:
            if "ic0_fp64" in want:
                run("ic0_fp64", lambda: F.Preconditioner(A, "ic0"), F.FP64)
            if "saamg_mp" in want:
                run("saamg_mp", lambda: F.Preconditioner(A, "saamg", **sa), F.FP32_IR)
            # non-learned approximate inverses: the Neumann form of A's own
            # pattern and cost, and the exact FSAI of the learned G's pattern
            # and cost (host-built matrices, device applies)
            if "neumann_mp" in want:
                Mn = F.CsrMatrix.from_gen(Lr.neumann_spai(G))
                run("neumann_mp", lambda: F.SpaiPreconditioner(A, Mn), F.FP32_IR)
            if "fsai_mp" in want:
                Gf = F.CsrMatrix.from_gen(Lr.fsai_exact(G))
                run("fsai_mp", lambda: F.SpaiPreconditioner(A, Gf, factorized=True), F.FP32_IR)
        if t == 0 and not args.no_cpu:
            # the oracle's float64 IC(0)-PCG solve of the first held-out seed
            import oracle_api as O
            O.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
            Ao = O.Csr.from_gen(G)
            xo, _ = O.solve(Ao, 0, O.Pc(Ao, O.K_IC0), b, tol=1e-10, max_iter=20000)
            oracle_err.update({k: float(np.linalg.norm(x - xo) / np.linalg.norm(xo)) for k, x in first_x.items()})
            import parity
            oracle_err["bar"] = max(1e-8, 2e-10 * parity.cond_bound_vardiff(N, kappa))
            oracle_err["bar_rule"] = ("cond(A) * 2 * 1e-10 with cond(A) <= 12 max(kappa) / (min(kappa) "
                                      "lambda_min(7-point Laplacian)), a rigorous upper bound")
            oracle_err["seed"] = seed
        if Ml is not None:
            Ml.close()
        A.close()
    med = {k: (float(np.median([r[1] for r in v])), float(np.median([r[2] for r in v])),
               float(np.median([r[3] for r in v]))) for k, v in rows.items() if v}
    ic0_seeds = [r[0] for r in rows["ic0_fp64"]]
    lm = {r[0]: r[1] for r in rows["learned_mp"]}
    speed = float(np.median([r[1] / lm[r[0]] for r in rows["ic0_fp64"]])) if ic0_seeds and lm else None
    jm = {r[0]: r[1] for r in rows["jacobi_fp64"]}
    vs_jac = float(np.median([lm[k] / jm[k] for k in lm if k in jm])) if lm and jm else None
    for k, (ms, it, su) in med.items():
        log(f"[bench] c4 {k}: median {ms:.2f} ms, {it:.0f} iterations, setup {su:.1f} ms")
    if speed:
        log(f"[bench] c4 learned MP vs FP64 IC0-PCG: {speed:.1f}x (median over seeds {ic0_seeds})")
    return {"workload": "3-D log-normal diffusion 96^3 (n=884736), sigma 1, correlation length 8 cells, tol 1e-10",
            "train": {"fields": args.c4_train, "field_cells": "32^3", "steps": steps, "seconds": round(train_s, 2),
                      "loss_first": round(float(hist[:10].mean()), 4), "loss_last": round(float(hist[-50:].mean()), 4)},
            "test_seeds": f"256..{255 + args.c4_test}",
            "median_ms": {k: round(v[0], 3) for k, v in med.items()},
            "median_iterations": {k: v[1] for k, v in med.items()},
            "median_setup_ms": {k: round(v[2], 2) for k, v in med.items()},
            "rel_err_vs_oracle": oracle_err,
            "speedup_learned_mp_vs_ic0_fp64": round(speed, 2) if speed else None,
            "learned_mp_time_over_jacobi_fp64": round(vs_jac, 3) if vs_jac else None,
            "fallbacks": int(fallbacks), "worst_true_relres": worst["relres"], "worst_rel_err": worst["rel_err"]}


# 1-norm condition number of the convection-diffusion matrix, estimated with
# scipy.sparse.linalg.onenormest on A and on A^-1 through a SuperLU
# factorization (about 35 s at 1024^2; --c3-cond recomputes it)
C3_KAPPA1 = {1024: 2.90e5, 256: 2.02e4}


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
    xs = {}
    for name, solver in (("bicgstab_ilu0", "bicgstab"), ("gmres30_ilu0", "gmres")):
        S = F.Solver(A, M, solver=solver, precision=F.FP64, max_iter=args.c3_max_iter, restart=30, device_ptrs=True,
                     tol=1e-10)
        t = time.perf_counter()
        st = S.solve_device(bd.address, xd.address)
        ms = (time.perf_counter() - t) * 1e3
        x = xd.download()
        xs[name] = x
        err = float(np.linalg.norm(x - xt) / np.linalg.norm(xt))
        out[name] = {"solve_ms": round(ms, 1), "setup_plus_solve_ms": round(setup_ms + ms, 1),
                     "iterations": st["iterations"], "converged": st["converged"],
                     "relres": st["true_relres"], "rel_err_vs_xtrue": err}
        log(f"[bench] c3 {name}: {ms:.0f} ms, {st['iterations']} iterations, converged {st['converged']}, "
            f"relres {st['true_relres']:.2e}")
        S.close()
    out["gmres30_note"] = ("restarted GMRES(30) + ILU(0) stagnates near 1e-9 on this matrix; the oracle's "
                           "GMRES(30) (array.cpp:4190-4401 restated) gives relres 1.004e-9 after 3000 iterations too")
    if not args.no_cpu:
        # the oracle's float64 BiCGSTAB + ILU(0) solve of the same system
        import oracle_api as O
        O.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        Ao = O.Csr.from_gen(G)
        t = time.time()
        xo, ito = O.solve(Ao, 1, O.Pc(Ao, O.K_ILU0), b, tol=1e-10, max_iter=args.c3_max_iter)
        out["oracle_bicgstab_ilu0"] = {"ms": round((time.time() - t) * 1e3, 1), "iterations": ito,
                                       "relres": float(np.linalg.norm(b - G.matvec(xo)) / np.linalg.norm(b))}
        kappa1 = C3_KAPPA1.get(args.c3_n)
        if args.c3_cond or kappa1 is None:
            import scipy.sparse.linalg as spla
            lu = spla.splu(G.to_scipy().tocsc())
            Ainv = spla.LinearOperator(lu.shape, matvec=lu.solve, rmatvec=lambda v: lu.solve(v, trans="T"))
            kappa1 = float(spla.onenormest(G.to_scipy()) * spla.onenormest(Ainv))
        ro = out["oracle_bicgstab_ilu0"]["relres"]
        for name, x in xs.items():
            out[name]["rel_err_vs_oracle"] = float(np.linalg.norm(x - xo) / np.linalg.norm(xo))
            out[name]["bar"] = kappa1 * (out[name]["relres"] + ro)
            log(f"[bench] c3 {name}: rel err vs oracle BiCGSTAB {out[name]['rel_err_vs_oracle']:.2e} "
                f"(bar {out[name]['bar']:.1e})")
        out["bar_rule"] = (f"kappa_1(A) * (relres + oracle relres), kappa_1 = {kappa1:.3g} estimated by "
                           "scipy's onenormest on A and on its LU solve (bench.py --c3-cond recomputes it)")
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
                     "refinements": st["refinements"], "converged": st["converged"], "relres": st["true_relres"]}
        log(f"[bench] rows {name}: setup {setup_ms:.1f} ms solve {ms:.2f} ms it {st['iterations']} "
            f"(refinements {st['refinements']}) converged {st['converged']} relres {st['true_relres']:.2e}")
        S.close()
        if M is not None:
            M.close()

    # C1: 2-D Poisson 64^2, Jacobi PCG (cache-resident: time only)
    G = gen.poisson2d(64)
    _, b = gen.rhs(G, 0)
    A = F.CsrMatrix.from_gen(G)
    csr_solve("c1_poisson64_2d_jacobi_pcg", A, b, "diag", max_iter=5000)
    A.close()
    G = gen.poisson3d(128)
    _, b = gen.rhs(G, 0)
    A = F.CsrMatrix.from_gen(G)
    csr_solve("poisson128_ssor_pcg", A, b, "ssor", omega=1.0)
    csr_solve("poisson128_ic0_pcg_mp", A, b, "ic0", prec=F.FP32_IR)
    # C2 with the C/F AMG (device setup), FP64 and FP32 inner / FP64 refinement
    csr_solve("poisson128_cf_amg_pcg", A, b, "amg")
    csr_solve("poisson128_cf_amg_pcg_mp", A, b, "amg", prec=F.FP32_IR)
    A.close()
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
    # C3 at full size in FP32 inner / FP64 refinement
    G = gen.convdiff2d(args.c3_n)
    _, b = gen.rhs(G, 0)
    A = F.CsrMatrix.from_gen(G)
    csr_solve(f"convdiff{args.c3_n}_bicgstab_ilu0_mp", A, b, "ilu0", solver="bicgstab", prec=F.FP32_IR,
              max_iter=args.c3_max_iter)
    csr_solve(f"convdiff{args.c3_n}_gmres30_ilu0_mp", A, b, "ilu0", solver="gmres", restart=30, prec=F.FP32_IR,
              max_iter=args.c3_max_iter)
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
    # algorithmic bytes of the batch solve: per inner iteration of a system
    # its float32 matrix (values and uint32 column indices, row pointers),
    # the inverse diagonal blocks twice (the update and the p passes) and 11
    # float32 vector transfers per row and right-hand side (q = A p: p, q;
    # update: d, r read and written, p, q; p update: r, p read and written);
    # per refinement step and once at the end the float64 matrix and 40 bytes
    # per row and right-hand side (b, x, r; x += d)
    total_bytes = 0.0
    for sd, s_ in zip(sysl, st):
        Am = sd["A"]
        bs = getattr(Am, "bs", 1)
        nb = Am.nbr if bs > 1 else Am.n
        nnzb = Am.nnzb if bs > 1 else Am.nnz
        n = nb * bs
        m32 = nnzb * (bs * bs * 4 + 4) + 4 * (nb + 1) + 2 * nb * bs * bs * 4
        m64 = nnzb * (bs * bs * 8 + 4) + 4 * (nb + 1)
        total_bytes += s_["iterations"] * (m32 + 44.0 * n * 4) + (s_["refinements"] + 1) * (m64 + 40.0 * n * 4)
    ms_med = float(np.median(times))
    gbs = total_bytes / (ms_med * 1e-3) / 1e9
    log(f"[bench] c5: {len(sysl)} systems ({unknowns} unknowns x rhs), schedule {best}, "
        f"solve {ms_med:.1f} ms, {conv}/{4 * len(sysl)} converged, worst relres {worst:.2e}, "
        f"{total_bytes / 1e9:.2f} GB algorithmic -> {gbs:.0f} GB/s")
    oracle_cmp = None
    if not args.no_cpu:
        # the oracle's float64 PCG (Jacobi / block-Jacobi, mini_cg) of the
        # first right-hand side of the largest system of each kind
        import oracle_api as O
        O.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        oracle_cmp = {}
        for kind in ("poisson2d", "bsr3"):
            k = max((i for i, sd in enumerate(sysl) if sd["kind"] == kind), key=lambda i: X[i][0].size)
            sd = sysl[k]
            if kind == "poisson2d":
                Ao = O.Csr.from_gen(sd["A"])
                xo, _ = O.solve(Ao, 0, O.Pc(Ao, O.K_DIAG), sd["B"][0], tol=1e-10, max_iter=20000)
            else:
                Ab = sd["A"]
                Bo = O.Bsr(3, Ab.nbr, Ab.nbc, Ab.row_ptr, Ab.col_idx, Ab.values)
                xo, _ = Bo.solve(0, O.Pc.block(Bo, 0), sd["B"][0], tol=1e-10, max_iter=20000)
            import parity
            cond = (parity.cond_poisson(sd["side"], 2) if kind == "poisson2d"
                    else parity.cond_bound_diagdom(sd["A"].to_scipy()))
            oracle_cmp[f"system_{k}_{kind}_side_{sd['side']}"] = {
                "rel_err": float(np.linalg.norm(X[k][0] - xo) / np.linalg.norm(xo)),
                "bar": max(1e-8, 2e-10 * cond),
                "bar_rule": "cond * 2e-10; 2-D Laplacian cond in closed form" if kind == "poisson2d"
                else "cond * 2e-10; cond bounded by Gershgorin (strict diagonal dominance)"}
        log(f"[bench] c5: rel err vs oracle {oracle_cmp}")
    B.close()
    return {"workload": f"{len(sysl)} SPD systems (half 2-D Poisson side 32..256, half 3x3-block 7-point side 16..48), "
                        "4 rhs each, block-Jacobi PCG FP32 inner / FP64 refinement, tol 1e-10, one persistent launch",
            "solve_ms_median": round(ms_med, 2), "unknowns_times_rhs": unknowns,
            "roofline": {"bound": "hbm", "algorithmic_bytes": int(total_bytes), "achieved": round(gbs, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)},
            "rel_err_vs_oracle": oracle_cmp,
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
    ap.add_argument("--c4-errors", choices=("solvers", "none"), default="solvers",
                    help="training probes besides random signs: error vectors of IC(0)/ILU(0)/SA-AMG PCG, or none")
    ap.add_argument("--c4-methods", default="all",
                    help="comma list of 96^3 comparisons: learned_mp,learned_fp64,jacobi_fp64,ic0_fp64,saamg_mp,"
                         "neumann_mp,fsai_mp")
    ap.add_argument("--no-c5", action="store_true", help="skip the batched-solve leg")
    ap.add_argument("--no-c3", action="store_true", help="skip the convection-diffusion ILU(0) leg (about 20 s)")
    ap.add_argument("--c3-n", type=int, default=1024, help="grid side of the convection-diffusion leg")
    ap.add_argument("--c3-max-iter", type=int, default=3000)
    ap.add_argument("--c3-cond", action="store_true",
                    help="estimate the convection-diffusion matrix's condition number (SuperLU, about 35 s)")
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
    # the device's streaming-read ceiling for the same bytes, warm and cold
    stream_warm_gbs = B64_moved / (F.bench_stream(B64_moved, False, 20) * 1e-3) / 1e9
    stream_cold_gbs = B64_moved / (F.bench_stream(B64_moved, True, 10) * 1e-3) / 1e9
    log(f"[bench] streaming read of {B64_moved / 1e6:.0f} MB: warm {stream_warm_gbs:.0f} GB/s, "
        f"cold {stream_cold_gbs:.0f} GB/s")
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
    # right-hand sides b_k = A x_k with x_k ~ U[-1,1] from PCG64(k); b_0 is the
    # one every method solves first (its solution is checked against x_0 and
    # against the oracle's float64 solve); each timed step solves another one
    nrhs = 1 + max(args.warmup, 1) + args.steps
    rhs = [(xt, b)] + [gen.rhs(G, k) for k in range(1, nrhs)]
    bds = [bd] + [F.DeviceArray.from_host(bk) for _, bk in rhs[1:]]
    results = {}
    xs0 = {}
    headline = None
    for name in want:
        kind, prm, prec = methods[name]
        # setup: the preconditioner and the solver object (its work vectors
        # and, for float32 inner solves, the float32 data it needs)
        F.device_synchronize()
        t = time.perf_counter()
        M = F.Preconditioner(A, kind, fp32=int(prec == F.FP32_IR), **prm)
        S = F.Solver(A, M, solver="pcg", precision=prec, max_iter=5000, device_ptrs=True, inner_tol=1e-5)
        F.device_synchronize()
        setup_ms = (time.perf_counter() - t) * 1e3
        is_head = name == "saamg_mp"
        # first solve (b_0) after the setup: includes capturing the graphs
        F.device_synchronize()
        t = time.perf_counter()
        st0 = S.solve_device(bds[0].address, xd.address)
        F.device_synchronize()
        first_ms = (time.perf_counter() - t) * 1e3
        x0 = xd.download()
        xs0[name] = x0
        relres = float(np.linalg.norm(b - G.matvec(x0)) / np.linalg.norm(b))
        err = float(np.linalg.norm(x0 - xt) / np.linalg.norm(xt))
        warm = args.warmup if is_head else 0
        steps = args.steps if is_head else 1
        for k in range(warm):
            S.solve_device(bds[1 + k].address, xd.address)
        # exactly `steps` solves, each of a right-hand side not solved before,
        # between two device synchronisations (one process, one GPU: no
        # barrier across ranks is needed)
        times = []
        F.device_synchronize()
        t_all = time.perf_counter()
        for k in range(steps):
            t = time.perf_counter()
            st = S.solve_device(bds[1 + warm + k].address, xd.address)
            times.append((time.perf_counter() - t) * 1e3)
        F.device_synchronize()
        total_ms = (time.perf_counter() - t_all) * 1e3
        xk = xd.download()
        xt_k = rhs[warm + steps][0]
        err_k = float(np.linalg.norm(xk - xt_k) / np.linalg.norm(xt_k))
        solve_ms = float(np.median(times))
        info = M.info()
        results[name] = {
            "solve_ms": round(solve_ms, 3), "setup_ms": round(setup_ms, 2),
            "setup_plus_solve_ms": round(setup_ms + solve_ms, 2), "first_solve_ms": round(first_ms, 2),
            "iterations": st["iterations"], "refinements": st["refinements"], "relres": relres,
            "rel_err_vs_xtrue": err, "rel_err_vs_xtrue_last_step": err_k, "levels": info["levels"],
        }
        if is_head:
            # the same system solved again with the same right-hand side: the
            # solver replays the chunk lengths it recorded for it
            rep = []
            for _ in range(max(3, min(args.steps, 10))):
                F.device_synchronize()
                t = time.perf_counter()
                S.solve_device(bds[0].address, xd.address)
                F.device_synchronize()
                rep.append((time.perf_counter() - t) * 1e3)
            results[name]["repeated_rhs_ms"] = round(float(np.median(rep)), 3)
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
            gbs = st["iterations"] * it_bytes / (solve_ms * 1e-3) / 1e9
            results[name]["solve_level"] = {
                "bytes_per_iteration": int(it_bytes), "precond_bytes": int(pc_bytes) if kind != "diag" else 0,
                "achieved_GBs": round(gbs, 1), "peak_GBs": 8000.0, "frac": round(gbs / 8000.0, 4)}
            log(f"[bench] {name}: {it_bytes / 1e6:.1f} MB per iteration (preconditioner {pc_bytes / 1e6:.1f} MB), "
                f"solve-level {gbs:.0f} GB/s ({gbs / 80:.1f}% of 8 TB/s)")
        log(f"[bench] {name}: setup {setup_ms:.1f} ms, first solve {first_ms:.1f} ms, solve {solve_ms:.2f} ms "
            f"(fresh right-hand sides), it {st['iterations']}, relres {relres:.2e}, err vs x_true {err:.2e}")
        if is_head:
            log(f"[bench] {name}: per-step ms {[round(t, 3) for t in times]}")
            headline = (total_ms / steps, times)
        S.close()
        M.close()

    # ---- learned preconditioner leg: 96^3 log-normal diffusion -----------
    c4 = None if args.no_c4 else c4_leg(F, gen, args, log)

    # ---- batched leg: 512 independent systems in one persistent launch ----
    c5 = None if args.no_c5 else c5_leg(F, gen, args, log)
    c3 = None if args.no_c3 else c3_leg(F, gen, args, log)
    rows = rows_leg(F, gen, args, log) if args.rows else None

    # ---- CPU baseline: the headline method in the oracle, on this host ---
    # The oracle's float64 PCG with the same SA-AMG (robust aggregation,
    # multicolour SGS 2/2, w = 1, coarsest <= 300 rows), setup included as
    # in the reference's solver call; its solution of b_0 is the float64
    # oracle solve every method's result is compared with.
    cpu = None
    oracle_cmp = None
    if not args.no_cpu:
        import oracle_api as O
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        O.set_threads(threads)
        Ao = O.Csr.from_gen(G)
        t = time.time()
        pc = O.Pc(Ao, O.K_SAAMG, [0.25, 1.0, 2, 2, 20, 4, 300, 0.0, 1, 1])
        cpu_setup_ms = (time.time() - t) * 1e3
        xo, ito = O.solve(Ao, 0, pc, b, tol=1e-10, max_iter=5000)
        cpu_ms = (time.time() - t) * 1e3
        del pc
        t = time.time()
        xo2, ito2 = O.solve(Ao, 0, O.Pc(Ao, O.K_IC0), b, tol=1e-10, max_iter=5000)
        cpu_ic0_ms = (time.time() - t) * 1e3
        cpu = {"value": round(cpu_ms, 1), "unit": "ms per solve (setup + iterations)", "cores": threads,
               "kind": "port",
               "sample": f"one full SA-AMG PCG solve (the headline method, float64) of the {N}^3 Poisson system "
                         f"to 1e-10: setup {cpu_setup_ms:.0f} ms + {ito + 1} iterations; the oracle's float64 "
                         f"IC(0)-PCG (setup + {ito2 + 1} iterations) took {cpu_ic0_ms:.0f} ms",
               "ic0_pcg_ms": round(cpu_ic0_ms, 1)}
        # parity bar: two iterates with ||b - A x|| <= tol ||b|| differ by at
        # most cond(A) * 2 tol; 3-D Poisson's condition number in closed form
        import parity
        cond = parity.cond_poisson(N, 3)
        bar = max(1e-8, 2e-10 * cond)
        oracle_cmp = {"oracle_solution": f"oracle float64 SA-AMG PCG of b_0 (relres "
                                         f"{np.linalg.norm(b - G.matvec(xo)) / np.linalg.norm(b):.2e})",
                      "bar": bar, "bar_rule": "max(1e-8, cond(A) * 2 * 1e-10), cond of the 3-D Laplacian in closed form", "oracle_ic0_vs_oracle_sa": float(np.linalg.norm(xo2 - xo) / np.linalg.norm(xo))}
        for name, x0 in xs0.items():
            e = float(np.linalg.norm(x0 - xo) / np.linalg.norm(xo))
            results[name]["rel_err_vs_oracle"] = e
            log(f"[bench] {name}: rel err vs oracle float64 solve {e:.2e} (bar {bar:.1e})")
        log(f"[bench] cpu oracle sa-amg pcg {cpu_ms:.0f} ms (setup {cpu_setup_ms:.0f}) it {ito}; ic0-pcg {cpu_ic0_ms:.0f} ms")

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
                   "step": "one solve of a right-hand side not solved before, preconditioner built, b and x in HBM; "
                           "setup_ms and setup_plus_solve_ms per method give the reference's call semantics "
                           "(its solver call builds the preconditioner)",
                   "methods": results, "oracle_check": oracle_cmp,
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
                     "avg_launch_us_cold": round(spmv_cold_ms * 1e3, 2),
                     "stream_read": {"warm_GBs": round(stream_warm_gbs, 1), "cold_GBs": round(stream_cold_gbs, 1),
                                     "note": "a plain 16-byte-load read of the moved bytes (fem_bench_stream): the "
                                             "device's practical ceiling; SpMV moved-bytes rate over it is "
                                             "frac_of_stream"},
                     "frac_of_stream": round(B64_moved / (spmv_ms * 1e-3) / 1e9 / stream_warm_gbs, 4),
                     "frac_of_stream_cold": round(B64_moved / (spmv_cold_ms * 1e-3) / 1e9 / stream_cold_gbs, 4)},
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
        out["config"]["speedup_saamg_mp_vs_ic0_fp64"] = {
            "solve": round(results["ic0_fp64"]["solve_ms"] / results["saamg_mp"]["solve_ms"], 2),
            "setup_plus_solve": round(results["ic0_fp64"]["setup_plus_solve_ms"] /
                                      results["saamg_mp"]["setup_plus_solve_ms"], 2)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
