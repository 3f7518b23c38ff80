"""GPU checks of the batched persistent solver (fem_batch_*): every system
reaches the float64 1e-10 residual for every right-hand side, solutions
match the oracle's block-Jacobi / Jacobi PCG solves within the parity bar,
results are deterministic for a schedule, the autotuner returns a schedule
that solves the batch, and bad input is rejected."""
import numpy as np
import pytest

import femgpu as F
import oracle_api as O
from femgpu import gen
from parity import cond_poisson, rel, tol_x

pytestmark = pytest.mark.gpu


def small_batch(seed=3, nrhs=4):
    return gen.batch_c5(seed, count=12, poisson_sides=(8, 24), bsr_sides=(4, 8), nrhs=nrhs)


def _scalar(A):
    return A.to_csr() if hasattr(A, "bs") else A


@pytest.mark.parametrize("nrhs", [4, 1, 3])
def test_batch_solves_every_system(nrhs):
    sysl = small_batch(nrhs=nrhs)
    B = F.Batch([s["A"] for s in sysl], nrhs=nrhs)
    B.set_rhs([s["B"] for s in sysl])
    stats, ms = B.solve(tol=1e-10)
    X = B.solution()
    assert ms > 0
    for s, st, x in zip(sysl, stats, X):
        S = _scalar(s["A"]).to_scipy()
        assert st["converged"] == nrhs and st["max_relres"] < 1e-10
        for c in range(nrhs):
            rr = np.linalg.norm(s["B"][c] - S @ x[c]) / np.linalg.norm(s["B"][c])
            assert rr < 1e-10
            assert rel(x[c], s["X"][c]) < tol_x(_scalar(s["A"]))


def test_batch_matches_oracle_pcg():
    sysl = small_batch(seed=5)
    B = F.Batch([s["A"] for s in sysl], nrhs=4)
    B.set_rhs([s["B"] for s in sysl])
    B.solve(tol=1e-10)
    X = B.solution()
    for s, x in zip(sysl, X):
        G = _scalar(s["A"])
        Ao = O.Csr.from_gen(G)
        if s["kind"] == "poisson2d":
            xo, _ = O.solve(Ao, 0, O.Pc(Ao, O.K_DIAG), s["B"][0], tol=1e-10, max_iter=5000)
        else:
            Ab = s["A"]
            Bo = O.Bsr(3, Ab.nbr, Ab.nbc, Ab.row_ptr, Ab.col_idx, Ab.values)
            xo, _ = Bo.solve(0, O.Pc.block(Bo, 0), s["B"][0], tol=1e-10, max_iter=5000)
        assert rel(x[0], xo) < tol_x(G)


def test_batch_c5_size_extremes_match_oracle():
    """Systems at the ends of the batched workload's size ranges (2-D Poisson
    sides 32 and 256, 3x3-block sides 16 and 48: up to 331,776 unknowns),
    solved together in one launch, each right-hand side against the float64
    residual bar, the first one against the oracle's PCG solve."""
    sysl = []
    for k, side in enumerate((32, 256)):
        A = gen.poisson2d(side)
        X, Bm = gen.rhs(A, 1000 + k, 4)
        sysl.append(dict(kind="poisson2d", A=A, X=X, B=Bm))
    for k, t in enumerate((16, 48)):
        Ab = gen.bsr7(t, 2000 + k)
        X = np.random.default_rng(3000 + k).uniform(-1.0, 1.0, (4, Ab.nbr * 3))
        sysl.append(dict(kind="bsr3", A=Ab, X=X, B=np.ascontiguousarray((Ab.to_scipy() @ X.T).T)))
    B = F.Batch([s["A"] for s in sysl], nrhs=4)
    B.set_rhs([s["B"] for s in sysl])
    stats, _ = B.solve(tol=1e-10)
    X = B.solution()
    for s, st, x in zip(sysl, stats, X):
        print(f"{s['kind']} n={x[0].size}: inner iterations {st['iterations']}, refinements {st['refinements']}, "
              f"relres {st['max_relres']:.2e}", flush=True)
        assert st["converged"] == 4 and st["max_relres"] < 1e-10
        G = _scalar(s["A"])
        Ao = O.Csr.from_gen(G)
        cond = None
        if s["kind"] == "poisson2d":
            xo, _ = O.solve(Ao, 0, O.Pc(Ao, O.K_DIAG), s["B"][0], tol=1e-10, max_iter=20000)
            cond = cond_poisson(int(round(np.sqrt(G.n))), 2)
        else:
            Ab = s["A"]
            Bo = O.Bsr(3, Ab.nbr, Ab.nbc, Ab.row_ptr, Ab.col_idx, Ab.values)
            xo, _ = Bo.solve(0, O.Pc.block(Bo, 0), s["B"][0], tol=1e-10, max_iter=20000)
        assert rel(x[0], xo) < tol_x(G, cond=cond)


def test_batch_deterministic_and_schedules():
    sysl = small_batch(seed=7)
    B = F.Batch([s["A"] for s in sysl], nrhs=4)
    B.set_rhs([s["B"] for s in sysl])
    s1, _ = B.solve()
    x1 = B.solution()
    s2, _ = B.solve()
    x2 = B.solution()
    assert all(np.array_equal(a, b) for a, b in zip(x1, x2))
    assert s1 == s2
    best, trials = B.tune(divisor=4)
    assert best["block"] in (256, 512, 1024) and best["order"] in (0, 1) and best["split"] >= 0
    assert np.all(trials > 0)
    for blk in (512, 1024):
        B.set_schedule(block=blk, order=1)
        st, _ = B.solve()
        X = B.solution()
        for s, t, x in zip(sysl, st, X):
            assert t["converged"] == 4
            assert rel(x[0], s["X"][0]) < tol_x(_scalar(s["A"]))


@pytest.mark.parametrize("block", [256, 512])
def test_split_right_hand_sides_give_identical_results(block):
    """Queuing systems as one work item per right-hand side changes the
    vector layout and the schedule, not the arithmetic: solutions, residuals
    and convergence are bit for bit those of the unsplit solve."""
    sysl = small_batch(seed=3)
    B = F.Batch([s["A"] for s in sysl], nrhs=4)
    B.set_rhs([s["B"] for s in sysl])
    B.set_schedule(block=block, order=0, split=0)
    s0, _ = B.solve()
    x0 = B.solution()
    for split in (2, len(sysl)):
        B.set_schedule(block=block, order=0, split=split)
        s1, _ = B.solve()
        x1 = B.solution()
        assert all(np.array_equal(a, b) for a, b in zip(x0, x1))
        for a, b in zip(s0, s1):
            assert a["converged"] == b["converged"] and a["refinements"] == b["refinements"]
            assert a["max_relres"] == b["max_relres"]
            # unsplit, the right-hand sides of a system step together: its
            # count is the sum over refinements of the per-step maximum
            assert b["iterations"] <= a["iterations"]
    with pytest.raises(F.InvalidArgument):
        B.set_schedule(block=block, split=-1)


def test_batch_errors():
    P = gen.poisson2d(4)
    with pytest.raises(F.InvalidArgument):
        F.Batch([P], nrhs=5)
    bad = gen.BSR(2, 2, 2, np.array([0, 1, 2], np.uint32), np.array([0, 1], np.uint32), np.ones(8))
    with pytest.raises(F.InvalidArgument):
        F.Batch([bad])
    nodiag = gen.CSR(2, 2, np.array([0, 1, 2], np.uint32), np.array([1, 0], np.uint32), np.ones(2))
    with pytest.raises(F.SolverFailure):
        F.Batch([nodiag])
    oor = gen.CSR(2, 2, np.array([0, 1, 2], np.uint32), np.array([0, 7], np.uint32), np.ones(2))
    with pytest.raises(F.OutOfRange):
        F.Batch([oor])
