"""The oracle against the committed golden fixtures (tests/golden/golden.json,
written by tests/golden/make_golden.py from hand-derived cases and closed
forms)."""
import json
import os

import numpy as np
import scipy.sparse as sp

import oracle_api as O

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def test_fixture_generator_is_reproducible(tmp_path):
    """The committed file is what the generator writes."""
    import subprocess
    import sys
    src = os.path.join(os.path.dirname(__file__), "golden", "make_golden.py")
    code = open(src).read().replace('os.path.dirname(os.path.abspath(__file__))', repr(str(tmp_path)))
    subprocess.run([sys.executable, "-c", code], check=True, capture_output=True)
    assert json.load(open(tmp_path / "golden.json")) == G


def test_coo_constructor_known_answer():
    c = G["coo_known"]
    A = O.Csr.from_coo(c["rows"], c["cols"], np.array(c["v"]), np.array(c["r"], np.uint32), np.array(c["c"], np.uint32))
    rp, ci, v = A.export()
    assert rp.tolist() == c["row_ptr"] and ci.tolist() == c["col_idx"] and v.tolist() == c["values"]


def test_spmv_exact_integers():
    c = G["spmv_integer"]
    A = O.Csr(c["rows"], c["cols"], np.array(c["row_ptr"], np.uint32), np.array(c["col_idx"], np.uint32),
              np.array(c["values"]))
    assert A.spmv(np.array(c["x"])).tolist() == c["y"]


def test_ic0_tridiagonal_closed_form():
    c = G["ic0_tridiag"]
    A = O.Csr(c["n"], c["n"], np.array(c["row_ptr"], np.uint32), np.array(c["col_idx"], np.uint32),
              np.array(c["values"]))
    rp, ci, v = O.Pc(A, O.K_IC0).ic0_factor()
    assert rp.tolist() == c["L_row_ptr"] and ci.tolist() == c["L_col_idx"]
    np.testing.assert_allclose(v, c["L_values"], rtol=1e-14)


def test_exact_factorizations_solve_tridiagonal():
    c = G["tridiag_solve"]
    for key, kind in (("sym", O.K_IC0), ("nonsym", O.K_ILU0), ("sym", O.K_ILU0)):
        m = c[key]
        A = O.Csr(c["n"], c["n"], np.array(m["row_ptr"], np.uint32), np.array(m["col_idx"], np.uint32),
                  np.array(m["values"]))
        np.testing.assert_allclose(O.Pc(A, kind).apply(np.array(c["r"])), m["z"], rtol=1e-13)


def _poisson2d(N):
    T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(N, N))
    A = (sp.kron(sp.eye(N), T) + sp.kron(T, sp.eye(N))).tocsr()
    A.eliminate_zeros()
    A.sort_indices()
    return A


def test_pcg_on_eigenvector_combination():
    c = G["pcg_eigen"]
    M = _poisson2d(c["N"])
    A = O.Csr(M.shape[0], M.shape[1], M.indptr, M.indices, M.data)
    b = np.array(c["b"])
    for name in ("none", "jacobi"):
        x, it = O.solve_named(A, 0, name, b, tol=1e-10)
        assert it == c["returned_iterations"], name
        np.testing.assert_allclose(x, c["x"], rtol=0, atol=1e-11 * np.abs(c["x"]).max())


def test_redblack_order_known_answer():
    c = G["redblack_order"]
    ny, nx = c["ny"], c["nx"]
    Tx = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(nx, nx))
    Ty = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(ny, ny))
    M = (sp.kron(sp.eye(ny), Tx) + sp.kron(Ty, sp.eye(nx))).tocsr()
    M.eliminate_zeros()
    M.sort_indices()
    A = O.Csr(M.shape[0], M.shape[1], M.indptr, M.indices, M.data)
    nid, cp = O.Pc(A, O.K_SSOR, [1.2]).ssor_order()
    assert nid.tolist() == c["new_id"] and cp.tolist() == c["colour_ptr"]
