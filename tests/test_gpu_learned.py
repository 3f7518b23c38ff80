"""GPU checks of the learned sparse approximate inverse against the oracle
(oracle/spai_oracle.cpp with bfloat16 operand rounding, the same places the
device kernels round): features, network outputs, loss and gradient; then
training, the gain over the preconditioner of the same apply cost without
the network (Jacobi), the fallback and the float32 refinement path."""
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


def test_learned_spai_beats_jacobi_on_heldout_fields(trained):
    """Held-out seeds at a larger grid: PCG with the learned M needs clearly
    fewer iterations than Jacobi (the same apply cost without the network),
    and reaches the same 1e-10 residual with the same solution."""
    net, _ = trained
    N = 24
    for seed in (900, 901):
        G = gen.vardiff3d(N, seed)[0]
        A = F.CsrMatrix.from_gen(G)
        xt, b = gen.rhs(G, seed)
        M = F.LearnedPreconditioner(A, (N, N, N), net)
        x, st = F.solve(A, M, b, solver="pcg", tol=1e-10, max_iter=5000)
        xj, sj = F.solve(A, F.Preconditioner(A, "diag"), b, solver="pcg", tol=1e-10, max_iter=5000)
        assert st["converged"] == 1 and st["fallback"] == 0 and st["true_relres"] < 1.5e-10
        assert st["iterations"] < 0.75 * sj["iterations"], (st["iterations"], sj["iterations"])
        xo, _ = O.solve(O.Csr.from_gen(G), 0, O.Pc(O.Csr.from_gen(G), O.K_IC0), b)
        assert _rel(x, xo) < 1e-8


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
