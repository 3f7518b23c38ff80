"""Shared parity bars for solution comparisons.

Two iterates that both satisfy ||b - A x|| <= tol ||b|| can differ by up to
cond(A) * 2 tol relative. The solution bar is therefore 1e-8 relative, or
that bound when cond(A) * 2e-10 exceeds it (computed exactly for test sizes
where the dense condition number is affordable)."""
import numpy as np

TOL_X = 1e-8


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def tol_x(G, tol=1e-10):
    if G.n > 5000:
        return TOL_X
    c = np.linalg.cond(G.to_scipy().toarray())
    return max(TOL_X, 2.0 * tol * c)
