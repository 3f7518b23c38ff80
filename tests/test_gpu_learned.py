"""GPU checks of the learned factorized sparse approximate inverse M = G^T G
against the oracle (oracle/spai_oracle.cpp with bfloat16 operand rounding,
the same places the device kernels round): features, network outputs, loss
and gradient; then training, the gain over the untrained network, over
Jacobi and over the non-learned approximate inverse of A's pattern and
apply cost (the Neumann form D^-1 (2I - A D^-1)), the SPD apply, the
budget and regression fallbacks and the float32 refinement path."""
import numpy as np
import pytest

import femgpu as F
import oracle_api as O
from femgpu import gen
from femgpu import learned as Lr

pytestmark = pytest.mark.gpu


def _params(seed, scale=0.3):
    return (np.random.default_rng(seed).standard_normal(F.MLP_NPARAM) * scale).astype(np.float32)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("n", [1, 16, 210, 1000])
def test_mlp_forward_matches_oracle(n):
    P = _params(0)
    X = np.random.default_rng(n).uniform(0, 1.5, (n, 6))
    net = F.Mlp(P)
    out = net.forward(X.astype(np.float32))
    ref = O.mlp_forward(P, X.astype(np.float32).astype(np.float64), bf16=True)
    assert np.abs(out - ref).max() <= 2e-3 * max(1.0, np.abs(ref).max())
    # the bfloat16 emulation matters: without it the error is much larger
    exact = O.mlp_forward(P, X.astype(np.float32).astype(np.float64))
    if n >= 16:
        assert np.abs(out - exact).max() > np.abs(out - ref).max()


def test_params_roundtrip_and_init():
    P = _params(1)
    net = F.Mlp(P)
    assert np.array_equal(net.params, P)
    d = F.Mlp(seed=3).params
    assert np.all(np.isfinite(d)) and np.abs(d).max() > 0


@pytest.mark.parametrize("grid", [(7, 6, 5), (8, 8, 8)])
def test_features_match_oracle(grid):
    G = gen.vardiff3d(grid[0], seed=2, ell=2.0)[0] if grid[0] == grid[1] == grid[2] else None
    if G is None:
        from test_oracle_spai import _axis_matrix
        G = _axis_matrix(*grid)
    A = F.CsrMatrix.from_gen(G)
    f = F.spai_features(A, grid)
    fo, _ = O.spai_features(G, grid)
    np.testing.assert_allclose(f, fo, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("grid,k", [((8, 8, 8), 4), ((7, 6, 5), 3)])
def test_loss_and_gradient_match_oracle(grid, k):
    if grid[0] == grid[1] == grid[2]:
        G = gen.vardiff3d(grid[0], seed=4, ell=2.0)[0]
    else:
        from test_oracle_spai import _axis_matrix
        G = _axis_matrix(*grid)
    A = F.CsrMatrix.from_gen(G)
    P = _params(5, 0.2)
    V = np.random.default_rng(6).choice([-1.0, 1.0], (k, G.n))
    net = F.Mlp(P)
    L, g = net.loss_grad(A, grid, V)
    Lo, go = O.spai_loss_grad(G, grid, P, V, bf16=True)
    assert abs(L - Lo) <= 1e-3 * abs(Lo)
    assert _rel(g.astype(np.float64), go) < 5e-3
    # and the gradient is that of the exact loss up to the bfloat16 rounding
    _, gx = O.spai_loss_grad(G, grid, P, V, bf16=False)
    assert _rel(g.astype(np.float64), gx) < 0.1


def test_gradient_is_deterministic():
    G = gen.vardiff3d(8, seed=1, ell=2.0)[0]
    A = F.CsrMatrix.from_gen(G)
    V = np.random.default_rng(0).choice([-1.0, 1.0], (3, G.n))
    net = F.Mlp(_params(7, 0.2))
    L1, g1 = net.loss_grad(A, (8, 8, 8), V)
    L2, g2 = net.loss_grad(A, (8, 8, 8), V)
    assert L1 == L2 and np.array_equal(g1, g2)


@pytest.fixture(scope="module")
def trained():
    fields = Lr.make_fields(16, range(12))
    net = F.Mlp(seed=0)
    hist, secs = Lr.train(net, fields, steps=360)
    return net, hist


def test_training_reduces_loss(trained):
    """The training loss falls; it keeps a floor from the smooth error
    vectors (no 7-point M inverts A on smooth modes), so the rough-probe
    loss on an unseen field is the sharper measure."""
    net, hist = trained
    assert hist[-40:].mean() < 0.85 * hist[:10].mean()
    f = Lr.make_fields(16, [950])[0]
    V = np.random.default_rng(3).choice([-1.0, 1.0], (8, f.G.n))
    L0, _ = F.Mlp(seed=0).loss_grad(f.A, f.grid, V)
    L1, _ = net.loss_grad(f.A, f.grid, V)
    assert L1 < 0.5 * L0, (L0, L1)


def test_learned_spai_gain_on_heldout_fields(trained):
    """Held-out seeds at a larger grid: PCG with the learned M needs fewer
    iterations than with the untrained network, with Jacobi, and with the
    Neumann approximate inverse of A's own pattern and apply cost, and
    reaches the same 1e-10 residual with the oracle's solution. The
    non-learned FSAI of G's pattern (exact local solves) is printed beside."""
    net, _ = trained
    N = 24
    for seed in (900, 901):
        G = gen.vardiff3d(N, seed)[0]
        A = F.CsrMatrix.from_gen(G)
        xt, b = gen.rhs(G, seed)

        def its(M):
            x, st = F.solve(A, M, b, solver="pcg", tol=1e-10, max_iter=5000)
            assert st["converged"] == 1 and st["fallback"] == 0 and st["true_relres"] < 1.5e-10
            return x, st["iterations"]

        x, il = its(F.LearnedPreconditioner(A, (N, N, N), net))
        _, i0 = its(F.LearnedPreconditioner(A, (N, N, N), F.Mlp(seed=0)))
        _, ij = its(F.Preconditioner(A, "diag"))
        Mn = F.CsrMatrix.from_gen(Lr.neumann_spai(G))
        _, inn = its(F.SpaiPreconditioner(A, Mn))
        Gf = F.CsrMatrix.from_gen(Lr.fsai_exact(G))
        _, ifs = its(F.SpaiPreconditioner(A, Gf, factorized=True))
        print(f"seed {seed}: learned {il}, untrained {i0}, jacobi {ij}, neumann {inn}, fsai {ifs}", flush=True)
        assert il < 0.8 * i0 and il < 0.7 * ij and il < inn
        xo, _ = O.solve(O.Csr.from_gen(G), 0, O.Pc(O.Csr.from_gen(G), O.K_IC0), b)
        assert _rel(x, xo) < 1e-8


def test_explicit_spai_apply_matches_scipy():
    """fem_precond_create_spai: z = M r and, factorized, z = G^T (G r) (one
    fused pass on the device) against scipy, on the Neumann M and the exact
    FSAI G of a log-normal field; and the FSAI property (G A)_ij = 0 on the
    strictly lower pattern with a unit diagonal of G A G^T."""
    N = 12
    G = gen.vardiff3d(N, 31)[0]
    A = F.CsrMatrix.from_gen(G)
    r = np.random.default_rng(5).standard_normal(G.n)
    Mn = Lr.neumann_spai(G)
    z = F.SpaiPreconditioner(A, F.CsrMatrix.from_gen(Mn)).apply(r)
    zr = Mn.to_scipy() @ r
    assert _rel(z, zr) < 1e-14
    Gf = Lr.fsai_exact(G)
    z = F.SpaiPreconditioner(A, F.CsrMatrix.from_gen(Gf), factorized=True).apply(r)
    Gs = Gf.to_scipy()
    assert _rel(z, Gs.T @ (Gs @ r)) < 1e-14
    Ad = G.to_scipy().toarray()
    Gd = Gs.toarray()
    GA = Gd @ Ad
    assert np.abs(GA[np.tril(Ad, -1) != 0]).max() < 1e-12
    np.testing.assert_allclose(np.diag(Gd @ Ad @ Gd.T), 1.0, rtol=1e-12)


def _long_row_factor(A):
    """Lower-triangular G = D^-1/2 plus a full last row and a full first
    column (small entries): rows of G and of G^T longer than a block's LDS
    stage (4096 float32 / 2048 float64 products)."""
    import scipy.sparse as sp
    As = A.to_scipy()
    n = As.shape[0]
    rng = np.random.default_rng(11)
    rows = [np.arange(n), np.full(n - 1, n - 1), np.arange(1, n - 1)]
    cols = [np.arange(n), np.arange(n - 1), np.zeros(n - 2, dtype=np.int64)]
    vals = [1.0 / np.sqrt(As.diagonal()), 1e-3 * rng.standard_normal(n - 1), 1e-3 * rng.standard_normal(n - 2)]
    Gs = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    Gs.sort_indices()
    return F.CsrMatrix(n, n, Gs.indptr, Gs.indices, Gs.data)


@pytest.mark.parametrize("prec", [F.FP64, F.FP32_IR], ids=["fp64", "fp32_ir"])
@pytest.mark.parametrize("case", ["fsai_vardiff24", "long_rows_poisson80"])
def test_fused_factorized_pcg_matches_six_launch_form(monkeypatch, case, prec):
    """PCG with M = G^T G runs three launches per iteration (t = G (r - alpha q)
    with the p.q tail; x/r update and p = G^T t + beta p with the ||r|| and
    r.z tails). Its iterates equal those of the six-launch form
    (FEM_PCG_FUSED=0) bit for bit, and the solve reaches 1e-10."""
    if case == "fsai_vardiff24":
        Gg = gen.vardiff3d(24, 905)[0]
        Gf = F.CsrMatrix.from_gen(Lr.fsai_exact(Gg))
    else:
        Gg = gen.poisson2d(80)
        Gf = _long_row_factor(Gg)
    A = F.CsrMatrix.from_gen(Gg)
    _, b = gen.rhs(Gg, 4)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("FEM_PCG_FUSED", mode)
        M = F.SpaiPreconditioner(A, Gf, factorized=True)
        x, st = F.solve(A, M, b, solver="pcg", precision=prec, tol=1e-10, max_iter=20000)
        assert st["converged"] == 1 and st["true_relres"] < 1.5e-10, (mode, st)
        out[mode] = (x, st["iterations"])
        M.close()
    print(f"{case}: {out['1'][1]} iterations", flush=True)
    assert out["1"][1] == out["0"][1]
    assert np.array_equal(out["1"][0], out["0"][0])


def test_learned_apply_is_spd_operator(trained):
    net, _ = trained
    G = gen.vardiff3d(10, seed=77, ell=3.0)[0]
    A = F.CsrMatrix.from_gen(G)
    M = F.LearnedPreconditioner(A, (10, 10, 10), net)
    rng = np.random.default_rng(0)
    u, v = rng.standard_normal(G.n), rng.standard_normal(G.n)
    a, b = u @ M.apply(v), v @ M.apply(u)
    assert abs(a - b) <= 1e-12 * abs(a)
    Mm = np.column_stack([M.apply(e) for e in np.eye(G.n)])
    assert np.linalg.eigvalsh(0.5 * (Mm + Mm.T)).min() > 0


def test_fallback_when_budget_runs_out(trained):
    net, _ = trained
    N = 16
    G = gen.vardiff3d(N, 905)[0]
    A = F.CsrMatrix.from_gen(G)
    xt, b = gen.rhs(G, 1)
    M = F.LearnedPreconditioner(A, (N, N, N), net, fallback="ic0", budget=5)
    x, st = F.solve(A, M, b, solver="pcg", tol=1e-10, max_iter=2000)
    assert st["fallback"] == 1 and st["converged"] == 1 and st["true_relres"] < 1.5e-10
    xi, si = F.solve(A, F.Preconditioner(A, "ic0"), b, solver="pcg", tol=1e-10, max_iter=2000)
    assert np.array_equal(x, xi) and st["iterations"] == si["iterations"]
    # a generous budget keeps the learned solve
    M2 = F.LearnedPreconditioner(A, (N, N, N), net, fallback="ic0", budget=2000)
    _, st2 = F.solve(A, M2, b, solver="pcg", tol=1e-10, max_iter=2000)
    assert st2["fallback"] == 0


def test_fallback_when_iteration_count_regresses(trained):
    """A learned solve that would need more iterations than the fallback's
    recorded count is a regression: it stops there and the fallback solves
    the system. The count is recorded when the fallback runs, or given."""
    net, _ = trained
    N = 16
    G = gen.vardiff3d(N, 906)[0]
    A = F.CsrMatrix.from_gen(G)
    _, b = gen.rhs(G, 4)
    xl, sl = F.solve(A, F.LearnedPreconditioner(A, (N, N, N), net), b, solver="pcg", tol=1e-10, max_iter=2000)
    _, si = F.solve(A, F.Preconditioner(A, "ic0"), b, solver="pcg", tol=1e-10, max_iter=2000)
    assert si["iterations"] < sl["iterations"]      # IC(0) needs fewer iterations here
    M = F.LearnedPreconditioner(A, (N, N, N), net, fallback="ic0")
    x1, s1 = F.solve(A, M, b, solver="pcg", tol=1e-10, max_iter=2000)
    assert s1["fallback"] == 0                       # no count known yet
    M.reference(si["inner_iterations"])
    x2, s2 = F.solve(A, M, b, solver="pcg", tol=1e-10, max_iter=2000)
    assert s2["fallback"] == 1 and s2["converged"] == 1 and s2["true_relres"] < 1.5e-10
    # the fallback's run recorded its own count: the next solve falls back too
    _, s3 = F.solve(A, M, b, solver="pcg", tol=1e-10, max_iter=2000)
    assert s3["fallback"] == 1 and s3["iterations"] == si["iterations"]
    # a count above the learned solve's keeps it
    M.reference(sl["inner_iterations"] + 10)
    _, s4 = F.solve(A, M, b, solver="pcg", tol=1e-10, max_iter=2000)
    assert s4["fallback"] == 0


def test_learned_fp32_refinement(trained):
    net, _ = trained
    N = 20
    G = gen.vardiff3d(N, 907)[0]
    A = F.CsrMatrix.from_gen(G)
    xt, b = gen.rhs(G, 2)
    M = F.LearnedPreconditioner(A, (N, N, N), net)
    x, st = F.solve(A, M, b, solver="pcg", precision=F.FP32_IR, tol=1e-10, max_iter=5000)
    assert st["converged"] == 1 and st["true_relres"] < 1e-10
    assert _rel(x, xt) < 1e-7


def test_errors():
    net = F.Mlp(seed=0)
    G = gen.poisson2d(6)
    A = F.CsrMatrix.from_gen(G)
    with pytest.raises(F.InvalidArgument):
        F.LearnedPreconditioner(A, (5, 5, 1), net)           # wrong size
    with pytest.raises(F.InvalidArgument):
        F.LearnedPreconditioner(A, (4, 9, 1), net)           # entries outside the stencil
    G3 = gen.poisson3d(4)
    A3 = F.CsrMatrix.from_gen(G3)
    with pytest.raises(F.InvalidArgument):
        F.LearnedPreconditioner(A3, (4, 4, 4), net, fallback="ssor")
    with pytest.raises(F.InvalidArgument):
        net.train_step(A3, (4, 4, 4), np.ones((1, 64)), lr=-1.0)
