"""Shared parity bars for solution comparisons.

Two iterates that both satisfy ||b - A x|| <= tol ||b|| can differ by up to
cond(A) * 2 tol relative. The solution bar is therefore 1e-8 relative, or
that bound when cond(A) * 2e-10 exceeds it (computed exactly for test sizes
where the dense condition number is affordable)."""
import numpy as np

TOL_X = 1e-8


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def tol_x(G, tol=1e-10, cond=None):
    """cond: cond(A) where it is known in closed form (any size)."""
    if cond is None:
        if G.n > 5000:
            return TOL_X
        cond = np.linalg.cond(G.to_scipy().toarray())
    return max(TOL_X, 2.0 * tol * cond)


# ---- condition numbers and bounds for the large workloads -------------------
# The solution bar of two iterates with relative residuals r1, r2 is
# cond(A) (r1 + r2); these give cond(A) (or an upper bound) without a
# factorization.
def cond_poisson(N, dims):
    """Dirichlet 5-point (dims=2) / 7-point (dims=3) Laplacian on N^dims
    points: lambda = sum_d (2 - 2 cos(k_d pi / (N + 1)))."""
    c = np.cos(np.pi / (N + 1))
    return (dims * (2 + 2 * c)) / (dims * (2 - 2 * c))


def cond_bound_vardiff(N, kappa):
    """Upper bound for the variable-coefficient diffusion matrix of
    femgpu.gen.vardiff3d: A - k_min L is a sum of positive semidefinite face
    and boundary terms (k_min = min kappa bounds every harmonic-mean face and
    boundary coefficient from below), so lambda_min(A) >= k_min lambda_min(L)
    with L the 7-point Dirichlet Laplacian; lambda_max(A) <= 2 max_i a_ii
    (Gershgorin: each off-diagonal row sum is at most the diagonal)."""
    lmin_L = 3 * (2 - 2 * np.cos(np.pi / (N + 1)))
    kmin = float(np.min(kappa))
    # a_ii <= 6 max kappa
    return (2 * 6 * float(np.max(kappa))) / (kmin * lmin_L)


def cond_bound_diagdom(A):
    """Symmetric, strictly diagonally dominant A (scipy sparse): Gershgorin
    bounds lambda_min >= min_i (a_ii - sum_j |a_ij|), lambda_max <= max_i
    (a_ii + sum_j |a_ij|)."""
    A = A.tocsr()
    d = A.diagonal()
    off = np.asarray(abs(A).sum(axis=1)).ravel() - np.abs(d)
    lo = float(np.min(d - off))
    if lo <= 0:
        return float("inf")
    return float(np.max(d + off)) / lo
