"""The HIP path against the committed golden fixtures (tests/golden/golden.json,
hand-derived cases and closed forms), through the C ABI."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import femgpu as F

pytestmark = pytest.mark.gpu

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def test_coo_constructor_known_answer():
    c = G["coo_known"]
    A = F.CsrMatrix.from_coo(c["rows"], c["cols"], np.array(c["v"]), np.array(c["r"], np.uint32),
                             np.array(c["c"], np.uint32))
    rp, ci, v = A.export()
    assert rp.tolist() == c["row_ptr"] and ci.tolist() == c["col_idx"] and v.tolist() == c["values"]


def test_spmv_exact_integers():
    c = G["spmv_integer"]
    A = F.CsrMatrix(c["rows"], c["cols"], c["row_ptr"], c["col_idx"], c["values"])
    assert A.spmv(np.array(c["x"])).tolist() == c["y"]


@pytest.mark.parametrize("key,kind", [("sym", "ic0"), ("nonsym", "ilu0"), ("sym", "ilu0")])
def test_exact_factorizations_solve_tridiagonal(key, kind):
    c = G["tridiag_solve"]
    m = c[key]
    A = F.CsrMatrix(c["n"], c["n"], m["row_ptr"], m["col_idx"], m["values"])
    np.testing.assert_allclose(F.Preconditioner(A, kind).apply(np.array(c["r"])), m["z"], rtol=1e-13)


def test_ic0_tridiagonal_closed_form_through_apply():
    """The IC(0) factor is the closed-form Cholesky factor: applying M^{-1} to
    L L^T e equals e."""
    c = G["ic0_tridiag"]
    n = c["n"]
    A = F.CsrMatrix(n, n, c["row_ptr"], c["col_idx"], c["values"])
    L = sp.csr_matrix((c["L_values"], c["L_col_idx"], c["L_row_ptr"]), shape=(n, n))
    e = np.linspace(-1.0, 1.0, n)
    np.testing.assert_allclose(F.Preconditioner(A, "ic0").apply(L @ (L.T @ e)), e, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("pc", ["none", "diag"])
def test_pcg_on_eigenvector_combination(pc):
    c = G["pcg_eigen"]
    N = c["N"]
    T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(N, N))
    M = (sp.kron(sp.eye(N), T) + sp.kron(T, sp.eye(N))).tocsr()
    M.eliminate_zeros()
    M.sort_indices()
    A = F.CsrMatrix(M.shape[0], M.shape[1], M.indptr, M.indices, M.data)
    P = None if pc == "none" else F.Preconditioner(A, "diag")
    x, st = F.solve(A, P, np.array(c["b"]), solver="pcg", tol=1e-10)
    assert st["iterations"] == c["returned_iterations"]
    assert st["inner_iterations"] == c["distinct_eigenvalues"]
    np.testing.assert_allclose(x, c["x"], rtol=0, atol=1e-11 * np.abs(c["x"]).max())
