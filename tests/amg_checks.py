"""Invariant checks of AMG hierarchies, shared by the oracle tests and the
GPU tests. They hold for any C/F split with direct interpolation, so they pin
the interpolation and the Galerkin products without fixing how points are
chosen."""
import numpy as np
import scipy.sparse as sp


def cf_interpolation_checks(A, P, st, theta=0.25):
    """Invariants of a C/F level with direct interpolation, independent of
    how the split was chosen: C rows of P are unit rows numbering the C points
    in row order; every F point has a strong C neighbour and its P row lists
    exactly its strong C neighbours; an F row of zero row sum whose strong
    neighbours are all C interpolates constants (row sum of P = 1). Returns
    the number of F rows the last check covered."""
    A = A.tocsr()
    P = P.tocsr()
    n = A.shape[0]
    assert set(np.unique(st)) <= {1, 2}
    cpts = np.flatnonzero(st == 1)
    assert P.shape == (n, cpts.size)
    mapc = -np.ones(n, np.int64)
    mapc[cpts] = np.arange(cpts.size)
    covered = 0
    for i in range(n):
        lo, hi = A.indptr[i], A.indptr[i + 1]
        cols, vals = A.indices[lo:hi], A.data[lo:hi]
        off = cols != i
        mx = np.abs(vals[off]).max() if off.any() else 0.0
        strong = set(cols[off & (np.abs(vals) >= theta * mx)].tolist())
        pc = P.indices[P.indptr[i]:P.indptr[i + 1]]
        pv = P.data[P.indptr[i]:P.indptr[i + 1]]
        if st[i] == 1:
            assert pc.tolist() == [mapc[i]] and pv.tolist() == [1.0]
            continue
        sc = sorted(mapc[j] for j in strong if st[j] == 1)
        assert sc, f"F point {i} has no strong C neighbour"
        assert sorted(pc.tolist()) == sc
        if abs(vals.sum()) < 1e-14 * np.abs(vals).max() and all(st[j] == 1 for j in strong):
            assert abs(pv.sum() - 1.0) < 1e-13
            covered += 1
    return covered


def galerkin_error(A, P, Ac):
    """max |P^T A P - A_c| relative to max |A_c|."""
    D = (P.T @ A @ P - Ac).toarray()
    return float(np.abs(D).max() / max(np.abs(Ac.toarray()).max(), 1e-300))


def cf_vcycle(As, Ps, Rs, r, omega=2 / 3, nu1=1, nu2=1):
    """One V-cycle from zero with damped Jacobi smoothing (weight omega on
    the diagonal), restriction by R, prolongation by P and a dense solve on
    the last level: the cycle of array.cpp:3344-3389 written from the level
    operators."""
    def cyc(k, f):
        A = As[k]
        if k == len(As) - 1:
            return np.linalg.solve(A.toarray(), f)
        d = A.diagonal()
        u = np.zeros_like(f)
        for _ in range(nu1):
            u = u + omega * (f - A @ u) / d
        uc = cyc(k + 1, Rs[k] @ (f - A @ u))
        u = u + Ps[k] @ uc
        for _ in range(nu2):
            u = u + omega * (f - A @ u) / d
        return u
    return cyc(0, np.asarray(r, dtype=np.float64))
