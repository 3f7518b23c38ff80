"""GPU parity: CSR SpMV, dot, and the Krylov solvers with the none / diag /
IC0 / ILU0 preconditioners against the CPU oracle, through the C ABI.

Bars: SpMV and the triangular sweeps are bit-exact (same products, same
summation order, no fused multiply-add); solves reach the reference's FP64
relative residual 1e-10 with solutions within 1e-8 relative of the oracle's.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import femgpu as F
import oracle_api as O
from femgpu import gen
from parity import cond_poisson, tol_x

pytestmark = pytest.mark.gpu

TOL_X = 1e-8


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_device_is_gfx950():
    name, cus = F.device_info()
    assert "gfx950" in name and cus > 0


@pytest.mark.parametrize("case", ["poisson3d", "convdiff", "random_long_rows", "rect_empty_rows"])
def test_spmv_bitexact(case):
    rng = np.random.default_rng(11)
    if case == "poisson3d":
        G = gen.poisson3d(20)
        rp, ci, v, n, m = G.row_ptr, G.col_idx, G.values, G.n, G.m
    elif case == "convdiff":
        G = gen.convdiff2d(64)
        rp, ci, v, n, m = G.row_ptr, G.col_idx, G.values, G.n, G.m
    elif case == "random_long_rows":
        M = sp.random(300, 5000, density=0.02, random_state=rng, format="lil")
        M[7, :] = rng.uniform(-1, 1, 5000)          # one row longer than an LDS stage
        M[200, :4000] = rng.uniform(-1, 1, 4000)
        M = M.tocsr()
        M.sort_indices()
        rp, ci, v, n, m = M.indptr, M.indices, M.data, 300, 5000
    else:
        M = sp.random(1000, 700, density=0.01, random_state=rng, format="csr")
        M.sort_indices()
        rp, ci, v, n, m = M.indptr, M.indices, M.data, 1000, 700
    A = F.CsrMatrix(n, m, rp, ci, v)
    Ao = O.Csr(n, m, rp, ci, v)
    x = rng.uniform(-1, 1, m)
    y = A.spmv(x)
    yo = Ao.spmv(x)
    if case == "random_long_rows":
        # rows longer than the LDS stage use a tree reduction: not bit-exact
        np.testing.assert_allclose(y, yo, rtol=1e-13, atol=1e-13)
    else:
        assert np.array_equal(y, yo)


def _spmv32(A, x):
    """float32 SpMV through fem_spmv_device on raw float32 device buffers."""
    import ctypes as C
    n, m = A.rows, A.cols
    xd, yd = F.DeviceArray((m + 1) // 2), F.DeviceArray((n + 1) // 2)
    x32 = np.ascontiguousarray(x, dtype=np.float32)
    assert F._hiprt().hipMemcpy(xd.ptr, x32.ctypes.data_as(C.c_void_p), m * 4, 1) == 0
    F._check(F.lib().fem_spmv_device(A._h, xd.ptr, yd.ptr, 1))
    return yd.download().view(np.float32)[:n].copy()


@pytest.mark.parametrize("case", ["poisson3d_48", "long_rows", "wide_columns"])
def test_spmv_packed_columns(case):
    """16-bit packed column offsets: used where every 64-entry chunk spans
    < 65536 columns; float64 results stay bit-exact with the oracle and the
    float32 ones identical to the unpacked kernel's."""
    rng = np.random.default_rng(17)
    if case == "poisson3d_48":
        G = gen.poisson3d(48)        # nnz 774,144; z-neighbours 2304 rows away
        rp, ci, v, n, m = G.row_ptr, G.col_idx, G.values, G.n, G.m
        packable, exact = True, True
    elif case == "long_rows":
        M = sp.random(2000, 40000, density=0.002, random_state=rng, format="lil")
        M[5, :30000] = rng.uniform(-1, 1, 30000)    # rows longer than an LDS stage
        M[1500, 9000:40000] = rng.uniform(-1, 1, 31000)
        M = M.tocsr()
        M.sort_indices()
        rp, ci, v, n, m = M.indptr, M.indices, M.data, 2000, 40000
        packable, exact = True, False
    else:
        M = sp.random(3000, 400000, density=0.0002, random_state=rng, format="csr")
        M.sort_indices()
        rp, ci, v, n, m = M.indptr, M.indices, M.data, 3000, 400000
        packable, exact = False, True
    assert len(ci) >= 65536
    A = F.CsrMatrix(n, m, rp, ci, v)
    assert A.pack_columns() == packable
    x = rng.uniform(-1, 1, m)
    yo = O.Csr(n, m, rp, ci, v).spmv(x)
    y = A.spmv(x)
    y32 = _spmv32(A, x)
    if exact:
        assert np.array_equal(y, yo)
    else:
        np.testing.assert_allclose(y, yo, rtol=1e-13, atol=1e-13 * np.abs(yo).max())
    np.testing.assert_allclose(y32, yo, rtol=1e-4, atol=1e-5 * np.abs(yo).max())
    assert A.pack_columns(0) is False
    assert np.array_equal(A.spmv(x), y)
    assert np.array_equal(_spmv32(A, x), y32)
    assert A.pack_columns(1) == packable
    assert np.array_equal(A.spmv(x), y)


@pytest.mark.parametrize("lanes", [4, 8, 16, 32, 64])
def test_spmv_row_lanes(lanes):
    """The lane-per-row SpMV variants sum in another order: within 1e-13 of
    the oracle; lanes 0 is the bit-exact CSR-stream kernel again."""
    rng = np.random.default_rng(5)
    M = sp.random(3000, 2500, density=0.01, random_state=rng, format="csr")
    M.sort_indices()
    A = F.CsrMatrix(3000, 2500, M.indptr, M.indices, M.data)
    x = rng.uniform(-1, 1, 2500)
    yo = O.Csr(3000, 2500, M.indptr, M.indices, M.data).spmv(x)
    A.set_row_lanes(lanes)
    np.testing.assert_allclose(A.spmv(x), yo, rtol=1e-13, atol=1e-13 * np.abs(yo).max())
    A.set_row_lanes(0)
    assert np.array_equal(A.spmv(x), yo)
    with pytest.raises(F.InvalidArgument):
        A.set_row_lanes(3)


def test_dot():
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(1_000_003), rng.standard_normal(1_000_003)
    assert abs(F.dot(a, b) - O.dot(a, b)) < 1e-11 * np.abs(a * b).sum()


def test_coo_constructor_matches_oracle():
    rng = np.random.default_rng(5)
    nin = 5000
    r = rng.integers(0, 300, nin)
    c = rng.integers(0, 200, nin)
    v = rng.uniform(-1, 1, nin)
    v[::7] = 0.0
    r[1::11] = r[0::11][: r[1::11].size]   # force duplicates
    c[1::11] = c[0::11][: c[1::11].size]
    A = F.CsrMatrix.from_coo(300, 200, v, r, c)
    Ao = O.Csr.from_coo(300, 200, v, r, c)
    for a, b in zip(A.export(), Ao.export()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n,m,nin", [(20000, 15000, 400000), (7, 5, 0), (1, 1, 3)])
def test_coo_device_conversion_large(n, m, nin):
    """COO -> CSR on the device: duplicates (several per key, summed in input
    order), explicit zeros, duplicates that cancel to zero, shuffled input;
    bit-identical to the oracle's restatement of the reference ctor."""
    rng = np.random.default_rng(11)
    r = rng.integers(0, n, nin)
    c = rng.integers(0, m, nin)
    v = rng.uniform(-1, 1, nin)
    if nin:
        v[::13] = 0.0
        k = nin // 4
        r[k:2 * k] = r[:k]
        c[k:2 * k] = c[:k]
        v[k:k + 50] = -v[:50]          # pairs that cancel
        perm = rng.permutation(nin)
        r, c, v = r[perm], c[perm], v[perm]
    A = F.CsrMatrix.from_coo(n, m, v, r, c)
    Ao = O.Csr.from_coo(n, m, v, r, c)
    for a, b in zip(A.export(), Ao.export()):
        assert np.array_equal(a, b)
    x = rng.standard_normal(m)
    assert np.array_equal(A.spmv(x), Ao.spmv(x))


def test_transpose_and_product():
    rng = np.random.default_rng(8)
    M = sp.random(80, 60, density=0.1, random_state=rng, format="csr")
    N = sp.random(60, 70, density=0.1, random_state=rng, format="csr")
    M.sort_indices(); N.sort_indices()
    A = F.CsrMatrix(80, 60, M.indptr, M.indices, M.data)
    B = F.CsrMatrix(60, 70, N.indptr, N.indices, N.data)
    rp, ci, v = A.transpose().export()
    np.testing.assert_array_equal(sp.csr_matrix((v, ci, rp), shape=(60, 80)).toarray(), M.toarray().T)
    C = A @ B
    rp, ci, v = C.export()
    np.testing.assert_allclose(sp.csr_matrix((v, ci, rp), shape=(80, 70)).toarray(), (M @ N).toarray(), atol=1e-14)


@pytest.mark.parametrize("pc", ["none", "diag"])
def test_pcg_c1_poisson64(pc):
    G = gen.poisson2d(64)
    xt, b = gen.rhs(G, 0)
    A = F.CsrMatrix.from_gen(G)
    M = F.Preconditioner(A, pc)
    x, st = F.solve(A, M, b, solver="pcg", tol=1e-10, max_iter=5000)
    xo, ito = O.solve_named(O.Csr.from_gen(G), 0, "jacobi" if pc == "diag" else "none", b, tol=1e-10, max_iter=5000)
    assert st["true_relres"] < 1e-10 * 1.5
    assert rel(x, xo) < TOL_X
    assert abs(st["iterations"] - ito) <= 2


def test_graph_and_plain_launches_agree():
    G = gen.poisson2d(40)
    _, b = gen.rhs(G, 2)
    A = F.CsrMatrix.from_gen(G)
    M = F.Preconditioner(A, "ic0")
    x1, s1 = F.solve(A, M, b, use_graph=True)
    x2, s2 = F.solve(A, M, b, use_graph=False)
    x3, s3 = F.solve(A, M, b, use_graph=True, check_every=3)
    assert np.array_equal(x1, x2) and np.array_equal(x1, x3)
    assert s1["iterations"] == s2["iterations"] == s3["iterations"]


@pytest.mark.parametrize("case", ["vardiff3d", "poisson3d", "unsorted_rows", "poisson3d_64", "poisson2d_1500"])
def test_ic0_apply_bitexact(case):
    """IC(0) factor + apply against the oracle, bit for bit. poisson3d_64 has
    levels of up to 3,072 rows (narrow chains, wide chains of two rows per
    thread, single-level launches); poisson2d_1500 has anti-diagonal levels of
    up to 1,500 rows (wide chains of four rows per thread)."""
    if case == "vardiff3d":
        G = gen.vardiff3d(14, seed=7)[0]
        rp, ci, v, n = G.row_ptr, G.col_idx, G.values, G.n
    elif case == "poisson3d":
        G = gen.poisson3d(24)
        rp, ci, v, n = G.row_ptr, G.col_idx, G.values, G.n
    elif case == "poisson3d_64":
        G = gen.poisson3d(64)
        rp, ci, v, n = G.row_ptr, G.col_idx, G.values, G.n
    elif case == "poisson2d_1500":
        G = gen.poisson2d(1500)
        rp, ci, v, n = G.row_ptr, G.col_idx, G.values, G.n
    else:
        G = gen.poisson2d(12)
        rp, ci, v, n = G.row_ptr.copy(), G.col_idx.copy(), G.values.copy(), G.n
        for i in range(0, n, 3):   # reverse each third row's storage order
            a, e = rp[i], rp[i + 1]
            ci[a:e] = ci[a:e][::-1].copy()
            v[a:e] = v[a:e][::-1].copy()
    A = F.CsrMatrix(n, n, rp, ci, v)
    Ao = O.Csr(n, n, rp, ci, v)
    r = np.random.default_rng(4).standard_normal(n)
    z = F.Preconditioner(A, "ic0").apply(r)
    zo = O.Pc(Ao, O.K_IC0).apply(r)
    assert np.array_equal(z, zo)


def _stencil27(N):
    """27-point stencil on N^3 points: diagonal 26, every neighbour -1 (SPD:
    27 I minus the Kronecker cube of tridiag(1, 1, 1))."""
    T = sp.diags([1.0, 1.0, 1.0], [-1, 0, 1], shape=(N, N))
    A = (27.0 * sp.identity(N ** 3) - sp.kron(sp.kron(T, T), T)).tocsr()
    A.sort_indices()
    return A


def _banded(n, half):
    """SPD band matrix: diagonal 3 half, every other entry within the band -1
    (rows of `half` lower entries, more than the block sweep holds)."""
    offs = list(range(-half, half + 1))
    A = sp.diags([3.0 * half if o == 0 else -1.0 for o in offs], offs, shape=(n, n)).tocsr()
    A.sort_indices()
    return A


@pytest.mark.parametrize("case", ["poisson3d_96", "stencil27_40", "banded_20", "banded_20_wide"])
def test_ic0_apply_sync_free_bitexact(case):
    """Block sweeps (rows of at most 16 entries) and the level-schedule
    fallback, bit for bit against the oracle. poisson3d_96: 109 blocks of
    8,176 rows whose boundaries cut grid planes (7-point rows); stencil27_40:
    13 entries per row (the widest register layout); banded_20: rows of 20
    lower entries and one row per level (level schedule: sync-free segment);
    banded_20_wide: the same band over 20 independent copies (levels of 20
    rows). Two applications in a row (the second reuses the sentinel buffer
    and the ticket counters)."""
    if case == "poisson3d_96":
        G = gen.poisson3d(96)
        rp, ci, v, n = G.row_ptr, G.col_idx, G.values, G.n
    else:
        if case == "stencil27_40":
            M = _stencil27(40)
        elif case == "banded_20":
            M = _banded(3000, 20)
        else:
            M = sp.block_diag([_banded(400, 20)] * 20, format="csr")
            M.sort_indices()
        rp, ci, v, n = M.indptr.astype(np.uint32), M.indices.astype(np.uint32), M.data, M.shape[0]
    A = F.CsrMatrix(n, n, rp, ci, v)
    Ao = O.Csr(n, n, rp, ci, v)
    Mg, Mo = F.Preconditioner(A, "ic0"), O.Pc(Ao, O.K_IC0)
    for seed in (4, 5):
        r = np.random.default_rng(seed).standard_normal(n)
        assert np.array_equal(Mg.apply(r), Mo.apply(r)), seed


def test_ilu0_apply_sync_free_bitexact():
    """ILU(0) of a nonsymmetric 3-D 7-point matrix (lower couplings 1.3, upper
    0.7) on 64^3 points: the unit-lower sweep (no division) and the upper
    sweep both have levels of up to 3,072 rows, swept sync-free."""
    G = gen.poisson3d(64)
    rows = np.repeat(np.arange(G.n), np.diff(G.row_ptr.astype(np.int64)))
    cols = G.col_idx.astype(np.int64)
    v = np.where(cols < rows, G.values * 1.3, np.where(cols > rows, G.values * 0.7, G.values))
    A = F.CsrMatrix(G.n, G.n, G.row_ptr, G.col_idx, v)
    Ao = O.Csr(G.n, G.n, G.row_ptr, G.col_idx, v)
    Mg, Mo = F.Preconditioner(A, "ilu0"), O.Pc(Ao, O.K_ILU0)
    for seed in (6, 7):
        r = np.random.default_rng(seed).standard_normal(G.n)
        assert np.array_equal(Mg.apply(r), Mo.apply(r)), seed


@pytest.mark.parametrize("fp32", [0, 1])
def test_ic0_sync_free_in_graphs(fp32):
    """IC(0)-PCG on the 7-point Laplacian on 64^3 points through captured
    graph chunks and through plain launches: identical iterates (float64),
    or convergence with float32 sweeps inside float64 refinement; the
    float64 iteration count and the solution against the oracle's."""
    G = gen.poisson3d(64)
    _, b = gen.rhs(G, 3)
    A = F.CsrMatrix.from_gen(G)
    M = F.Preconditioner(A, "ic0", fp32=fp32)
    x1, s1 = F.solve(A, M, b, use_graph=True)
    x2, s2 = F.solve(A, M, b, use_graph=False)
    xo, ito = O.solve_named(O.Csr.from_gen(G), 0, "ic0", b)
    assert s1["true_relres"] < 1.5e-10 and s2["true_relres"] < 1.5e-10
    if not fp32:
        assert np.array_equal(x1, x2) and s1["iterations"] == s2["iterations"]
        assert abs(s1["iterations"] - ito) <= 1
    assert rel(x1, xo) < tol_x(G, cond=cond_poisson(64, 3))


def test_ilu0_apply_bitexact():
    G = gen.convdiff2d(50)
    A = F.CsrMatrix.from_gen(G)
    r = np.random.default_rng(6).standard_normal(G.n)
    z = F.Preconditioner(A, "ilu0").apply(r)
    zo = O.Pc(O.Csr.from_gen(G), O.K_ILU0).apply(r)
    assert np.array_equal(z, zo)


def test_ilu0_general_pattern_bitexact():
    rng = np.random.default_rng(9)
    M = sp.random(400, 400, density=0.02, random_state=rng, format="csr") + sp.eye(400) * 5.0
    M = M.tocsr()
    M.sort_indices()
    A = F.CsrMatrix(400, 400, M.indptr, M.indices, M.data)
    r = rng.standard_normal(400)
    z = F.Preconditioner(A, "ilu0").apply(r)
    zo = O.Pc(O.Csr(400, 400, M.indptr, M.indices, M.data), O.K_ILU0).apply(r)
    assert np.array_equal(z, zo)


def test_pcg_ic0_vardiff():
    G = gen.vardiff3d(20, seed=1)[0]
    _, b = gen.rhs(G, 0)
    A = F.CsrMatrix.from_gen(G)
    x, st = F.solve(A, F.Preconditioner(A, "ic0"), b, solver="pcg")
    xo, ito = O.solve_named(O.Csr.from_gen(G), 0, "ic0", b)
    assert st["true_relres"] < 1.5e-10
    assert rel(x, xo) < TOL_X
    assert abs(st["iterations"] - ito) <= 2


@pytest.mark.parametrize("solver", ["bicgstab", "gmres"])
@pytest.mark.parametrize("pc", ["ilu0", "diag", "none"])
def test_nonsymmetric_convdiff(solver, pc):
    G = gen.convdiff2d(48)
    _, b = gen.rhs(G, 5)
    A = F.CsrMatrix.from_gen(G)
    M = F.Preconditioner(A, pc)
    x, st = F.solve(A, M, b, solver=solver, restart=30, max_iter=4000)
    xo, ito = O.solve_named(O.Csr.from_gen(G), 1 if solver == "bicgstab" else 2, pc, b, max_iter=4000, restart=30)
    assert st["true_relres"] < 1e-9
    assert rel(x, xo) < tol_x(G)
    assert abs(st["iterations"] - ito) <= max(3, ito // 20)


@pytest.mark.parametrize("solver,pc", [("pcg", "ic0"), ("pcg", "diag"), ("bicgstab", "ilu0"), ("gmres", "ilu0")])
def test_fp32_refinement_reaches_fp64_tolerance(solver, pc):
    G = gen.vardiff3d(16, seed=2)[0] if solver == "pcg" else gen.convdiff2d(40)
    xt, b = gen.rhs(G, 1)
    A = F.CsrMatrix.from_gen(G)
    M = F.Preconditioner(A, pc, fp32=1)
    x, st = F.solve(A, M, b, solver=solver, precision=F.FP32_IR, max_iter=5000)
    assert st["converged"] == 1
    assert st["true_relres"] < 1e-10
    xref = O.solve_named(O.Csr.from_gen(G), {"pcg": 0, "bicgstab": 1, "gmres": 2}[solver], pc, b, max_iter=5000)[0]
    assert rel(x, xref) < tol_x(G)


def test_error_paths():
    # not SPD -> PCG failure (array.cpp:3549-3552)
    M = sp.csr_matrix(np.array([[1.0, 0.0], [0.0, -1.0]]))
    A = F.CsrMatrix(2, 2, M.indptr, M.indices, M.data)
    with pytest.raises(F.SolverFailure):
        F.solve(A, None, np.array([1.0, 1.0]), solver="pcg")
    # the same system in float32 inner / float64 refinement: a breakdown of
    # the inner solve before it has reduced its residual fails the solve as
    # in float64 (p.Ap = 0 in the first iteration; p.Ap < 0 in the second)
    for bb in ([1.0, 1.0], [1.0, 0.5]):
        with pytest.raises(F.SolverFailure):
            F.solve(A, None, np.array(bb), solver="pcg", precision=F.FP32_IR)
    # zero / negative diagonal -> diag preconditioner failure (array.cpp:2485-2493)
    with pytest.raises(F.SolverFailure):
        F.Preconditioner(A, "diag")
    # IC0 on an indefinite matrix
    M2 = sp.csr_matrix(np.array([[1.0, 2.0], [2.0, 1.0]]))
    with pytest.raises(F.SolverFailure):
        F.Preconditioner(F.CsrMatrix(2, 2, M2.indptr, M2.indices, M2.data), "ic0")
    # b = 0 -> x = 0, iterations 0
    G = gen.poisson2d(8)
    Ag = F.CsrMatrix.from_gen(G)
    x, st = F.solve(Ag, F.Preconditioner(Ag, "diag"), np.zeros(G.n))
    assert not x.any() and st["iterations"] == 0
    # out-of-range column
    with pytest.raises(F.OutOfRange):
        F.CsrMatrix(2, 2, [0, 1, 2], [0, 5], [1.0, 1.0])
    # silent iteration cap: max_iter - 1, no error
    _, b = gen.rhs(G, 0)
    x, st = F.solve(Ag, F.Preconditioner(Ag, "diag"), b, max_iter=5)
    assert st["iterations"] == 4 and st["converged"] == 0


def test_persistent_solver_reuse():
    G = gen.poisson2d(48)
    A = F.CsrMatrix.from_gen(G)
    M = F.Preconditioner(A, "ic0")
    S = F.Solver(A, M, solver="pcg")
    for seed in range(3):
        xt, b = gen.rhs(G, seed)
        x, st = S.solve(b)
        xo, _ = O.solve_named(O.Csr.from_gen(G), 0, "ic0", b)
        assert rel(x, xo) < TOL_X and st["true_relres"] < 1.5e-10


def test_gmres_host_callback():
    G = gen.convdiff2d(24)
    _, b = gen.rhs(G, 3)
    A = F.CsrMatrix.from_gen(G)
    d = G.to_scipy().diagonal()
    x, it = F.gmres_callback(A, lambda r: r / d, b, restart=20, tol=1e-10, max_iter=2000)
    assert np.linalg.norm(b - G.matvec(x)) / np.linalg.norm(b) < 1e-9
