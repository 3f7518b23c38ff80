"""Training and evaluation of the learned sparse approximate inverse.

The network and every kernel run on the device through the C ABI
(fem_mlp_*, fem_spai_*); this module only draws the fields and probe
vectors and loops over Adam steps.

Probe vectors per field: random +-1 vectors (redrawn every step) and error
vectors e = x* - x_j of the library's own IC(0)-PCG runs stopped after j
iterations (j = 2, 4, 8, 16; drawn once per field), each scaled to norm
sqrt(n). The loss 1/(n k) sum ||M A v - v||^2 over both kinds is the
Frobenius-norm SPAI objective estimated on rough and smooth vectors.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import gen

__all__ = ["Field", "make_fields", "error_vectors", "train", "pcg_iterations"]


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


def error_vectors(f: Field, iters=(2, 4, 8, 16), seed=0, ic0=None):
    """e = x* - x_j after j iterations of the library's IC(0)-PCG, scaled to
    norm sqrt(n); one random x* per j."""
    from . import Preconditioner, solve
    M = ic0 if ic0 is not None else Preconditioner(f.A, "ic0")
    rng = np.random.default_rng(10_000 + 31 * f.seed + seed)
    n = f.G.n
    E = np.empty((len(iters), n))
    for t, j in enumerate(iters):
        xs = rng.uniform(-1.0, 1.0, n)
        b = f.G.matvec(xs)
        x, _ = solve(f.A, M, b, solver="pcg", tol=1e-30, max_iter=j)
        e = xs - x
        E[t] = e * (np.sqrt(n) / np.linalg.norm(e))
    return E


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
