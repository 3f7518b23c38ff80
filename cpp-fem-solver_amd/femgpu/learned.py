"""Training and evaluation of the learned sparse approximate inverse.

The network and every kernel run on the device through the C ABI
(fem_mlp_*, fem_spai_*); this module only draws the fields and probe
vectors and loops over Adam steps, and builds the non-learned baselines.

Probe vectors per field: random +-1 vectors (redrawn every step) and error
vectors e = x* - x_j of the library's own preconditioned CG runs stopped
after j iterations (IC(0) after 2 and 8, ILU(0) after 4 and 16, SA-AMG after
1 and 2; drawn once per field), each scaled to norm sqrt(n). The loss
1/(n k) sum ||M A v - v||^2 over both kinds is the Frobenius-norm SPAI
objective estimated on rough vectors and on the smooth errors the library's
solvers leave.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import gen

__all__ = ["Field", "make_fields", "error_vectors", "neumann_spai", "fsai_exact", "train", "pcg_iterations"]


@dataclass
class Field:
    G: gen.CSR
    A: object            # femgpu.CsrMatrix
    grid: tuple
    seed: int
    errors: np.ndarray = field(default=None)   # (k_e, n)


def make_fields(N: int, seeds, sigma=1.0, ell=8.0):
    """Log-normal diffusion matrices on N^3 cells (gen.vardiff3d)."""
    from . import CsrMatrix
    out = []
    for s in seeds:
        G, _ = gen.vardiff3d(N, s, sigma, ell)
        out.append(Field(G, CsrMatrix.from_gen(G), (N, N, N), int(s)))
    return out


# (preconditioner, iterations) of the runs whose errors become probes
ERROR_RUNS = (("ic0", 2), ("ic0", 8), ("ilu0", 4), ("ilu0", 16), ("saamg", 1), ("saamg", 2))


def error_vectors(f: Field, runs=ERROR_RUNS, seed=0):
    """e = x* - x_j after j PCG iterations with the library's IC(0), ILU(0)
    and SA-AMG preconditioners (one random x* per run), scaled to norm
    sqrt(n)."""
    from . import Preconditioner, solve
    rng = np.random.default_rng(10_000 + 31 * f.seed + seed)
    n = f.G.n
    pcs = {}
    E = np.empty((len(runs), n))
    for t, (kind, j) in enumerate(runs):
        if kind not in pcs:
            pcs[kind] = Preconditioner(f.A, kind)
        xs = rng.uniform(-1.0, 1.0, n)
        b = f.G.matvec(xs)
        x, _ = solve(f.A, pcs[kind], b, solver="pcg", tol=1e-30, max_iter=j)
        e = xs - x
        E[t] = e * (np.sqrt(n) / np.linalg.norm(e))
    return E


def neumann_spai(G: gen.CSR) -> gen.CSR:
    """The non-learned approximate inverse of A's own pattern and apply cost:
    the truncated Neumann series M = D^-1 (2 I - A D^-1)."""
    import scipy.sparse as sp
    A = G.to_scipy().tocsr()
    Dinv = sp.diags(1.0 / A.diagonal())
    M = (Dinv @ (2.0 * sp.identity(G.n) - A @ Dinv)).tocsr()
    M.sort_indices()
    return gen.CSR(G.n, G.n, M.indptr.astype(np.uint32), M.indices.astype(np.uint32), M.data.astype(np.float64))


def fsai_exact(G: gen.CSR) -> gen.CSR:
    """The non-learned factorized approximate inverse of the learned G's
    pattern and cost (Kolotilina & Yeremin's FSAI on A's lower pattern): row i
    of G solves A[J, J] g = e_i on J = {j <= i : a_ij != 0}, scaled to
    g_ii = 1 / sqrt(g_ii), so that G A G^T has a unit diagonal."""
    import scipy.sparse as sp
    A = G.to_scipy().tocsr()
    L = sp.tril(A).tocsr()
    L.sort_indices()
    n = G.n
    cnt = np.diff(L.indptr)
    vals = np.empty(L.nnz)
    for k in np.unique(cnt):
        rows = np.flatnonzero(cnt == k)
        J = np.stack([L.indices[L.indptr[rows] + t] for t in range(k)], axis=1)   # (m, k), ascending: i last
        sub = np.empty((rows.size, k, k))
        for a in range(k):
            for b in range(k):
                sub[:, a, b] = np.asarray(A[J[:, a], J[:, b]]).ravel()
        rhs = np.zeros((rows.size, k, 1))
        rhs[:, k - 1, 0] = 1.0
        g = np.linalg.solve(sub, rhs)[:, :, 0]
        g /= np.sqrt(g[:, k - 1])[:, None]
        for t in range(k):
            vals[L.indptr[rows] + t] = g[:, t]
    return gen.CSR(n, n, L.indptr.astype(np.uint32), L.indices.astype(np.uint32), vals)


def train(net, fields, steps=600, k_rand=6, lr=2e-3, lr_end=2e-4, seed=0, log=None):
    """Adam on the stochastic SPAI loss, one field per step (round robin).
    Returns (loss history, seconds)."""
    rng = np.random.default_rng(seed)
    for f in fields:
        if f.errors is None:
            f.errors = error_vectors(f)
    hist = []
    t0 = time.perf_counter()
    for step in range(steps):
        f = fields[step % len(fields)]
        n = f.G.n
        V = np.empty((k_rand + len(f.errors), n))
        V[:k_rand] = rng.choice([-1.0, 1.0], size=(k_rand, n))
        V[k_rand:] = f.errors
        # cosine decay of the learning rate
        a = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * step / max(1, steps - 1)))
        hist.append(net.train_step(f.A, f.grid, V, lr=a))
        if log and (step % 100 == 0 or step == steps - 1):
            log(f"[learned] step {step} loss {hist[-1]:.4f}")
    return np.array(hist), time.perf_counter() - t0


def pcg_iterations(A, M, b, tol=1e-10, max_iter=5000, **kw):
    from . import solve
    x, st = solve(A, M, b, solver="pcg", tol=tol, max_iter=max_iter, **kw)
    return x, st
