"""Writes golden.json: small known-answer cases for the sparse solve path,
each fixed by hand or by a closed form (never by the oracle or the library):

* coo_known      COO input with duplicates, explicit zeros, a duplicate pair
                 that cancels and unsorted order; the CSR the SMatrix COO
                 constructor (array.cpp:2072-2182) must produce, written out
                 by hand.
* spmv_integer   integer-valued sparse matrix and vector: y = A x is exact in
                 float64 (computed in int64).
* ic0_tridiag    IC(0) of tridiag(-1, 2, -1) is its Cholesky factor:
                 L_ii = sqrt((i+2)/(i+1)), L_i,i-1 = -sqrt(i/(i+1)).
* tridiag_solve  IC(0) and ILU(0) of a tridiagonal matrix are exact
                 factorizations, so one preconditioner apply solves A z = r
                 (z from the Thomas algorithm in exact-ish float64).
* pcg_eigen      2-D Poisson (5-point, N = 10) with b a sum of four
                 eigenvectors sin(k pi i/(N+1)) sin(l pi j/(N+1)) of distinct
                 eigenvalues: CG (unpreconditioned or Jacobi, whose D is 4 I)
                 converges at the fourth step, so the reference's returned
                 index (the last non-converged iteration, array.cpp:3445-3601)
                 is 2, and x = sum v_kl / lambda_kl.
* redblack_order the stable red-black renumbering of a 4 x 5 grid
                 (array.cpp:2935-3067): red cells (i+j even) first in row
                 order, then black.

Run: python tests/golden/make_golden.py (writes tests/golden/golden.json)."""
import json
import math
import os

import numpy as np


def csr_from_dense(D):
    rp, ci, v = [0], [], []
    for i in range(D.shape[0]):
        for j in range(D.shape[1]):
            if D[i, j] != 0:
                ci.append(j)
                v.append(float(D[i, j]))
        rp.append(len(ci))
    return rp, ci, v


def main():
    out = {}
    # --- COO constructor ---------------------------------------------------
    entries = [(2, 1, 1.5), (0, 4, 2.0), (2, 1, -1.5), (1, 0, 0.0), (3, 3, 1.0), (0, 0, -1.0), (3, 3, 2.0),
               (0, 4, 0.25), (1, 2, 0.5), (1, 2, 0.125)]
    out["coo_known"] = {
        "rows": 4, "cols": 5,
        "r": [e[0] for e in entries], "c": [e[1] for e in entries], "v": [e[2] for e in entries],
        # row 0: (0, -1), (4, 2 + 0.25); row 1: (2, 0.5 + 0.125), the 0.0 dropped;
        # row 2: 1.5 - 1.5 = 0 dropped; row 3: (3, 1 + 2)
        "row_ptr": [0, 2, 3, 3, 4], "col_idx": [0, 4, 2, 3], "values": [-1.0, 2.25, 0.625, 3.0],
    }
    # --- exact SpMV ----------------------------------------------------------
    rng = np.random.default_rng(11)
    D = rng.integers(-9, 10, size=(40, 30)) * (rng.random((40, 30)) < 0.2)
    x = rng.integers(-20, 21, size=30)
    rp, ci, v = csr_from_dense(D)
    out["spmv_integer"] = {"rows": 40, "cols": 30, "row_ptr": rp, "col_idx": ci, "values": v,
                           "x": [float(t) for t in x], "y": [float(t) for t in (D.astype(np.int64) @ x)]}
    # --- IC(0) of tridiag(-1, 2, -1) -----------------------------------------
    n = 10
    T = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    rp, ci, v = csr_from_dense(T)
    Lrp, Lci, Lv = [0], [], []
    for i in range(n):
        if i > 0:
            Lci.append(i - 1)
            Lv.append(-math.sqrt(i / (i + 1)))
        Lci.append(i)
        Lv.append(math.sqrt((i + 2) / (i + 1)))
        Lrp.append(len(Lci))
    out["ic0_tridiag"] = {"n": n, "row_ptr": rp, "col_idx": ci, "values": v,
                          "L_row_ptr": Lrp, "L_col_idx": Lci, "L_values": Lv}
    # --- one apply of an exact factorization ------------------------------------
    n = 12
    sub = -1.0 - 0.25 * np.arange(1, n)
    sup = -0.5 * np.ones(n - 1)
    dia = 4.0 + 0.1 * np.arange(n)
    A = np.diag(dia) + np.diag(sub, -1) + np.diag(sup, 1)
    S = 0.5 * (A + A.T)
    r = np.cos(np.arange(n) + 0.5)

    def thomas(M, rhs):
        a = np.array([M[i, i - 1] for i in range(1, n)])
        b = np.diag(M).copy()
        c = np.array([M[i, i + 1] for i in range(n - 1)])
        d = rhs.copy()
        for i in range(1, n):
            w = a[i - 1] / b[i - 1]
            b[i] -= w * c[i - 1]
            d[i] -= w * d[i - 1]
        z = np.zeros(n)
        z[-1] = d[-1] / b[-1]
        for i in range(n - 2, -1, -1):
            z[i] = (d[i] - c[i] * z[i + 1]) / b[i]
        return z

    rpA, ciA, vA = csr_from_dense(A)
    rpS, ciS, vS = csr_from_dense(S)
    out["tridiag_solve"] = {"n": n, "r": list(r),
                            "nonsym": {"row_ptr": rpA, "col_idx": ciA, "values": vA, "z": list(thomas(A, r))},
                            "sym": {"row_ptr": rpS, "col_idx": ciS, "values": vS, "z": list(thomas(S, r))}}
    # --- CG on an eigenvector combination -------------------------------------
    N = 10
    i = np.arange(N)[:, None] + 1
    j = np.arange(N)[None, :] + 1
    b = np.zeros((N, N))
    xs = np.zeros((N, N))
    for (k, l, c) in ((1, 1, 1.0), (1, 3, -0.5), (2, 5, 0.75), (4, 4, 0.25)):
        vkl = np.sin(k * math.pi * i / (N + 1)) * np.sin(l * math.pi * j / (N + 1))
        lam = 4 - 2 * math.cos(k * math.pi / (N + 1)) - 2 * math.cos(l * math.pi / (N + 1))
        b += c * lam * vkl
        xs += c * vkl
    out["pcg_eigen"] = {"N": N, "b": list(b.ravel()), "x": list(xs.ravel()), "distinct_eigenvalues": 4,
                        "returned_iterations": 2}
    # --- red-black renumbering ---------------------------------------------------
    ny, nx = 4, 5
    cells = [(a, c) for a in range(ny) for c in range(nx)]
    red = [a * nx + c for a, c in cells if (a + c) % 2 == 0]
    black = [a * nx + c for a, c in cells if (a + c) % 2 == 1]
    order = red + black
    new_id = [0] * (ny * nx)
    for pos, old in enumerate(order):
        new_id[old] = pos
    out["redblack_order"] = {"ny": ny, "nx": nx, "new_id": new_id, "colour_ptr": [0, len(red), ny * nx]}

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
