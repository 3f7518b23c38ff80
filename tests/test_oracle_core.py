"""Pins the CPU oracle (oracle/fem_oracle.cpp) against what the mathematics
fixes: dense brute force, closed forms, scipy, factorization invariants and
the reference's own threshold tests. None of these retype the oracle's
formulas: each check is a property a dropped term, wrong index or wrong sign
would break."""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle_api as O
from amg_checks import cf_interpolation_checks
from femgpu import gen


def dense(A):
    rp, ci, v = A.export() if isinstance(A, O.Csr) else (A.row_ptr, A.col_idx, A.values)
    n = len(rp) - 1
    m = A.m if hasattr(A, "m") else int(ci.max()) + 1
    D = np.zeros((n, m))
    for i in range(n):
        for p in range(rp[i], rp[i + 1]):
            D[i, ci[p]] += v[p]
    return D


def rand_sparse(n, m, density, seed, diag_shift=None):
    rng = np.random.default_rng(seed)
    M = sp.random(n, m, density=density, random_state=rng, format="csr")
    M.data = rng.uniform(-1, 1, M.data.size)
    if diag_shift is not None:
        M = M + sp.eye(n) * diag_shift
    M = M.tocsr()
    M.sort_indices()
    return M


def ocsr(M):
    return O.Csr(M.shape[0], M.shape[1], M.indptr, M.indices, M.data)


# ---------------------------------------------------------------- container
def test_coo_semantics():
    # duplicates summed, zeros dropped before and after merging, (i,j) sorted
    v = [1.0, 0.0, 2.0, -2.0, 3.0, 4.0, 5.0]
    r = [2, 0, 1, 1, 0, 2, 0]
    c = [1, 0, 2, 2, 3, 1, 0]
    A = O.Csr.from_coo(3, 4, v, r, c)
    rp, ci, vv = A.export()
    assert list(rp) == [0, 2, 2, 3]
    assert list(ci) == [0, 3, 1]
    assert list(vv) == [5.0, 3.0, 5.0]


def test_coo_out_of_range():
    with pytest.raises(O.OracleError) as e:
        O.Csr.from_coo(2, 2, [1.0], [2], [0])
    assert e.value.code == 2


def test_spmv_dense_bruteforce():
    M = rand_sparse(57, 43, 0.2, 1)
    A = ocsr(M)
    x = np.random.default_rng(2).uniform(-1, 1, 43)
    np.testing.assert_allclose(A.spmv(x), M.toarray() @ x, rtol=1e-14, atol=1e-14)


def test_spmv_integer_exact():
    G = gen.poisson3d(7)
    A = O.Csr.from_gen(G)
    x = np.random.default_rng(0).integers(-100, 100, G.n).astype(float)
    assert np.array_equal(A.spmv(x), dense(G) @ x)


def test_dot():
    a = np.random.default_rng(1).standard_normal(10001)
    b = np.random.default_rng(2).standard_normal(10001)
    import math
    assert abs(O.dot(a, b) - math.fsum(a * b)) < 1e-10


def test_transpose_and_spgemm():
    M = rand_sparse(30, 20, 0.2, 3)
    N = rand_sparse(20, 25, 0.2, 4)
    A, B = ocsr(M), ocsr(N)
    np.testing.assert_allclose(dense(A.transpose()), M.toarray().T)
    np.testing.assert_allclose(dense(A.spgemm(B)), M.toarray() @ N.toarray(), atol=1e-14)
    # a column reached by a single row only (the case the reference's marker
    # reuse can drop) must be present
    P = sp.csr_matrix(np.array([[1.0, 0, 0], [0, 2.0, 0], [0, 0, 3.0]]))
    Q = sp.csr_matrix(np.array([[0, 0, 5.0], [1.0, 0, 0], [1.0, 0, 0]]))
    np.testing.assert_array_equal(dense(ocsr(P).spgemm(ocsr(Q))), P.toarray() @ Q.toarray())


# ---------------------------------------------------------------- IC0 / ILU0
def test_ic0_invariant_on_pattern():
    G = gen.vardiff3d(6, seed=3)[0]
    A = O.Csr.from_gen(G)
    pc = O.Pc(A, O.K_IC0)
    rp, ci, v = pc.ic0_factor()
    L = np.zeros((G.n, G.n))
    for i in range(G.n):
        for p in range(rp[i], rp[i + 1]):
            L[i, ci[p]] = v[p]
    D = dense(G)
    LLt = L @ L.T
    mask = (np.tril(D) != 0)
    np.testing.assert_allclose(LLt[mask], D[mask], rtol=1e-12, atol=1e-12)
    # apply = (L L^T)^{-1} r
    r = np.random.default_rng(0).standard_normal(G.n)
    np.testing.assert_allclose(pc.apply(r), np.linalg.solve(LLt, r), rtol=1e-10)


def test_ic0_tridiagonal_is_cholesky():
    n = 12
    T = np.diag(np.full(n, 2.5)) + np.diag(np.full(n - 1, -1.0), 1) + np.diag(np.full(n - 1, -1.0), -1)
    M = sp.csr_matrix(T)
    pc = O.Pc(ocsr(M), O.K_IC0)
    rp, ci, v = pc.ic0_factor()
    L = np.zeros((n, n))
    for i in range(n):
        for p in range(rp[i], rp[i + 1]):
            L[i, ci[p]] = v[p]
    np.testing.assert_allclose(L, np.linalg.cholesky(T), rtol=1e-14)


def test_ic0_not_spd_raises():
    M = sp.csr_matrix(np.array([[1.0, 2.0], [2.0, 1.0]]))
    with pytest.raises(O.OracleError) as e:
        O.Pc(ocsr(M), O.K_IC0)
    assert e.value.code == 3


def test_ilu0_textbook_on_stencil():
    G = gen.convdiff2d(9)
    A = O.Csr.from_gen(G)
    pc = O.Pc(A, O.K_ILU0)
    (lp, lc, lv), (up, uc, uv) = pc.ilu0_factors()
    n = G.n
    L = np.eye(n)
    U = np.zeros((n, n))
    for i in range(n):
        for p in range(lp[i], lp[i + 1]):
            L[i, lc[p]] = lv[p]
        for p in range(up[i], up[i + 1]):
            U[i, uc[p]] = uv[p]
    D = dense(G)
    mask = D != 0
    np.testing.assert_allclose((L @ U)[mask], D[mask], rtol=1e-12, atol=1e-12)
    r = np.random.default_rng(5).standard_normal(n)
    np.testing.assert_allclose(pc.apply(r), np.linalg.solve(L @ U, r), rtol=1e-10)


def test_ilu0_upper_only_variant_on_general_pattern():
    # The reference updates only U-row entries (array.cpp:2854-2896): for every
    # L entry, L_ik U_kk = a_ik with the ORIGINAL a_ik, and for every U entry,
    # U_ij + sum_k L_ik U_kj = a_ij over k in L-row(i) with (k,j) in U.
    M = rand_sparse(40, 40, 0.12, 7, diag_shift=6.0)
    A = ocsr(M)
    pc = O.Pc(A, O.K_ILU0)
    (lp, lc, lv), (up, uc, uv) = pc.ilu0_factors()
    D = M.toarray()
    U = {}
    for i in range(40):
        for p in range(up[i], up[i + 1]):
            U[(i, uc[p])] = uv[p]
    for i in range(40):
        for p in range(lp[i], lp[i + 1]):
            k = lc[p]
            assert abs(lv[p] * U[(k, k)] - D[i, k]) < 1e-12
        for p in range(up[i], up[i + 1]):
            j = uc[p]
            s = U[(i, j)]
            for q in range(lp[i], lp[i + 1]):
                k = lc[q]
                if (k, j) in U:
                    s += lv[q] * U[(k, j)]
            assert abs(s - D[i, j]) < 1e-12


# ---------------------------------------------------------------- SSOR
def test_ssor_redblack_order_closed_form():
    N = 6
    G = gen.poisson2d(N)
    pc = O.Pc(O.Csr.from_gen(G), O.K_SSOR, [1.2])
    nid, cp = pc.ssor_order()
    assert len(cp) == 3
    par = np.array([(i // N + i % N) % 2 for i in range(N * N)])
    expect = np.empty(N * N, np.int64)
    reds = np.flatnonzero(par == 0)
    blacks = np.flatnonzero(par == 1)
    expect[reds] = np.arange(reds.size)
    expect[blacks] = reds.size + np.arange(blacks.size)
    assert np.array_equal(nid, expect)


@pytest.mark.parametrize("omega", [1.0, 1.2, 1.7])
def test_ssor_dense_operator(omega):
    G = gen.vardiff3d(5, seed=2)[0]
    pc = O.Pc(O.Csr.from_gen(G), O.K_SSOR, [omega])
    nid, cp = pc.ssor_order()
    D = dense(G)
    n = G.n
    P = np.zeros((n, n))
    P[nid, np.arange(n)] = 1.0
    B = P @ D @ P.T
    Dg = np.diag(np.diag(B))
    Lw = np.tril(B, -1)
    Uw = np.triu(B, 1)
    Mss = (Dg + omega * Lw) @ np.linalg.inv(Dg) @ (Dg + omega * Uw) / (omega * (2 - omega))
    r = np.random.default_rng(1).standard_normal(n)
    z = P.T @ np.linalg.solve(Mss, P @ r)
    np.testing.assert_allclose(pc.apply(r), z, rtol=1e-10)


def test_ssor_omega_out_of_range():
    G = gen.poisson2d(4)
    for om in (0.0, 2.0, -1.0):
        with pytest.raises(O.OracleError) as e:
            O.Pc(O.Csr.from_gen(G), O.K_SSOR, [om])
        assert e.value.code == 1


def test_greedy_multicolour_fallback():
    # a triangle is not bipartite: greedy colouring with 3 colours
    M = sp.csr_matrix(np.array([[4.0, -1, -1], [-1, 4.0, -1], [-1, -1, 4.0]]))
    pc = O.Pc(ocsr(M), O.K_SSOR, [1.0])
    nid, cp = pc.ssor_order()
    assert len(cp) == 4 and list(nid) == [0, 1, 2]


# ---------------------------------------------------------------- Krylov
def test_pcg_diag_vs_scipy_cg():
    G = gen.poisson2d(24)
    xt, b = gen.rhs(G, 0)
    A = O.Csr.from_gen(G)
    x, it = O.solve_named(A, 0, "jacobi", b, tol=1e-10, max_iter=5000)
    S = G.to_scipy()
    np.testing.assert_allclose(x, spla.spsolve(S.tocsc(), b), rtol=1e-8, atol=1e-9)
    # constant diagonal: Jacobi-PCG iterates equal plain CG's (scaling invariance)
    x2, it2 = O.solve_named(A, 0, "none", b, tol=1e-10, max_iter=5000)
    assert it == it2
    # scipy's CG (same recurrence, same ||r||/||b|| test) needs it+2 matvecs
    cnt = {"k": 0}
    spla.cg(S, b, rtol=1e-10, maxiter=5000, callback=lambda xk: cnt.__setitem__("k", cnt["k"] + 1))
    assert abs(cnt["k"] - (it + 2)) <= 1


def test_pcg_iteration_semantics_and_silent_cap():
    G = gen.poisson2d(16)
    _, b = gen.rhs(G, 1)
    A = O.Csr.from_gen(G)
    x, it = O.solve_named(A, 0, "diag", b, tol=1e-10, max_iter=7)
    assert it == 6  # max_iter - 1, no error (array.cpp:3597)
    x0, it0 = O.solve_named(A, 0, "diag", np.zeros(G.n), tol=1e-10)
    assert it0 == 0 and not x0.any()


def test_pcg_not_spd_raises():
    M = sp.csr_matrix(np.array([[1.0, 0.0], [0.0, -1.0]]))
    with pytest.raises(O.OracleError) as e:
        O.solve_named(ocsr(M), 0, "none", np.array([1.0, 1.0]))
    assert e.value.code == 3


@pytest.mark.parametrize("name", ["ilu0", "ssor", "none", "diag"])
def test_bicgstab_gmres_convdiff(name):
    G = gen.convdiff2d(20)
    _, b = gen.rhs(G, 3)
    A = O.Csr.from_gen(G)
    ref = spla.spsolve(G.to_scipy().tocsc(), b)
    for solver in (1, 2):
        x, it = O.solve_named(A, solver, name, b, tol=1e-10, max_iter=3000, restart=30)
        r = np.linalg.norm(b - G.matvec(x)) / np.linalg.norm(b)
        assert r < 1e-9
        assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-8


def test_gmres_full_space_terminates():
    M = rand_sparse(25, 25, 0.3, 9, diag_shift=3.0)
    b = np.random.default_rng(0).standard_normal(25)
    x, it = O.solve_named(ocsr(M), 2, "none", b, tol=1e-12, restart=25, max_iter=25)
    assert it <= 24
    np.testing.assert_allclose(M @ x, b, atol=1e-10)


def test_alias_tables():
    # per-solver alias sets (array.cpp:3484-3511, 3641-3669, 4215-4238)
    assert O.alias("CG", 0) == O.K_NONE and O.alias("cg", 1) == -1
    assert O.alias("Incomplete_Cholesky", 0) == O.K_IC0 and O.alias("incomplete_cholesky", 2) == -1
    assert O.alias("ICHOL", 2) == O.K_IC0
    for s in (0, 1, 2):
        assert O.alias("Jacobi", s) == O.K_DIAG and O.alias("SOR", s) == O.K_SSOR
        assert O.alias("ilu", s) == O.K_ILU0 and O.alias("amg", s) == O.K_AMG
        assert O.alias("bogus", s) == -1


def test_unknown_preconditioner_raises_only_for_nonzero_b():
    G = gen.poisson2d(4)
    A = O.Csr.from_gen(G)
    x, it = O.solve_named(A, 0, "bogus", np.zeros(G.n))
    assert it == 0
    with pytest.raises(O.OracleError) as e:
        O.solve_named(A, 0, "bogus", np.ones(G.n))
    assert e.value.code == 1


# ---------------------------------------------------------------- AMG
def test_cf_amg_hierarchy_invariants():
    G = gen.poisson2d(40)
    A = O.Csr.from_gen(G)
    pc = O.Pc(A, O.K_AMG, [0.25, 300, 25])
    rows, nnz = pc.amg_levels()
    assert rows[0] == G.n and rows[-1] <= 300 and len(rows) >= 2
    for l in range(len(rows) - 1):
        n, m, rp, ci, v = pc.amg_matrix(l, 0)
        Al = sp.csr_matrix((v, ci, rp), shape=(n, m))
        n2, m2, rp2, ci2, v2 = pc.amg_matrix(l, 1)
        P = sp.csr_matrix((v2, ci2, rp2), shape=(n2, m2))
        n3, m3, rp3, ci3, v3 = pc.amg_matrix(l + 1, 0)
        Ac = sp.csr_matrix((v3, ci3, rp3), shape=(n3, m3))
        np.testing.assert_allclose(Ac.toarray(), (P.T @ Al @ P).toarray(), atol=1e-12)
        # C rows of P are identity rows: every column has exactly one unit row
        Pd = P.toarray()
        unit_rows = [i for i in range(n2) if np.count_nonzero(Pd[i]) == 1 and Pd[i].max() == 1.0]
        assert len(unit_rows) >= m2


def test_cf_amg_interpolation_with_weak_couplings():
    """Anisotropic Laplacian (eps = 0.01 < theta): the y couplings are weak
    and are lumped into the interpolation's diagonal, which the constant
    row-sum check of interior F rows pins (array.cpp:3276-3319)."""
    G = gen.aniso2d(24, 0.01)
    A = O.Csr.from_gen(G)
    pc = O.Pc(A, O.K_AMG, [0.25, 300, 25])
    rows, _ = pc.amg_levels()
    assert len(rows) >= 2
    n, m, rp, ci, v = pc.amg_matrix(0, 1)
    P = sp.csr_matrix((v, ci, rp), shape=(n, m))
    covered = cf_interpolation_checks(G.to_scipy(), P, pc.cf_split(0))
    assert covered > G.n // 4


def test_cf_vcycle_restatement_matches_oracle():
    """The numpy V-cycle the GPU tests use (amg_checks.cf_vcycle), run on the
    oracle's own exported hierarchy, equals the oracle's V-cycle."""
    from amg_checks import cf_vcycle
    G = gen.poisson2d(30)
    pc = O.Pc(O.Csr.from_gen(G), O.K_AMG, [0.25, 300, 25])
    rows, _ = pc.amg_levels()
    mats = []
    for l in range(len(rows)):
        trip = []
        for which in ((0, 1, 2) if l + 1 < len(rows) else (0,)):
            n, m, rp, ci, v = pc.amg_matrix(l, which)
            trip.append(sp.csr_matrix((v, ci, rp), shape=(n, m)))
        mats.append(trip)
    As = [t[0] for t in mats]
    Ps = [t[1] for t in mats[:-1]]
    Rs = [t[2] for t in mats[:-1]]
    r = np.random.default_rng(4).standard_normal(G.n)
    z = pc.apply(r)
    np.testing.assert_allclose(cf_vcycle(As, Ps, Rs, r), z, rtol=1e-11, atol=1e-11 * np.abs(z).max())


def test_amg_vcycle_symmetric_and_solves():
    G = gen.poisson2d(30)
    A = O.Csr.from_gen(G)
    pc = O.Pc(A, O.K_AMG, [0.25, 300, 25])
    rng = np.random.default_rng(0)
    u, v = rng.standard_normal(G.n), rng.standard_normal(G.n)
    assert abs(u @ pc.apply(v) - v @ pc.apply(u)) < 1e-10 * abs(u @ pc.apply(v))
    _, b = gen.rhs(G, 0)
    x, it = O.solve(A, 0, pc, b, tol=1e-10)
    assert np.linalg.norm(b - G.matvec(x)) / np.linalg.norm(b) < 1e-10
    assert it < 40


def _greedy_aggregates(S_lists, n, max_agg):
    agg = -np.ones(n, np.int64)
    na = 0
    for i in range(n):
        if agg[i] != -1:
            continue
        agg[i] = na
        sz = 1
        for j in S_lists[i]:
            if agg[j] == -1 and sz < max_agg:
                agg[j] = na
                sz += 1
        na += 1
    return agg, na


@pytest.mark.parametrize("exact", [0, 1])
def test_saamg_prolongator_and_galerkin(exact):
    G = gen.vardiff3d(8, seed=5)[0]
    A = O.Csr.from_gen(G)
    prm = [0.25, 2 / 3, 1, 1, 20, 4, 50, 2 / 3, exact]
    pc = O.Pc(A, O.K_SAAMG, prm)
    rows, nnz = pc.amg_levels()
    assert rows[0] == G.n and rows[-1] <= 50
    D = dense(G)
    S = []
    for i in range(G.n):
        off = [(j, abs(D[i, j])) for j in range(G.n) if j != i and D[i, j] != 0]
        mx = max(a for _, a in off)
        S.append([j for j, a in off if a >= 0.25 * mx])
    agg, na = _greedy_aggregates(S, G.n, 4)
    assert na == rows[1]
    Pt = np.zeros((G.n, na))
    Pt[np.arange(G.n), agg] = 1.0
    Pexp = (np.eye(G.n) - (2 / 3) * np.diag(1 / np.diag(D)) @ D) @ Pt
    n2, m2, rp2, ci2, v2 = pc.amg_matrix(0, 1)
    P = sp.csr_matrix((v2, ci2, rp2), shape=(n2, m2)).toarray()
    np.testing.assert_allclose(P, Pexp, atol=1e-13)
    n3, m3, rp3, ci3, v3 = pc.amg_matrix(1, 0)
    np.testing.assert_allclose(sp.csr_matrix((v3, ci3, rp3), shape=(n3, m3)).toarray(), P.T @ D @ P, atol=1e-12)


def test_saamg_exact_coarse_pcg():
    G = gen.poisson3d(12)
    A = O.Csr.from_gen(G)
    pc = O.Pc(A, O.K_SAAMG, [0.25, 2 / 3, 1, 1, 20, 4, 50, 2 / 3, 1])
    _, b = gen.rhs(G, 0)
    x, it = O.solve(A, 0, pc, b, tol=1e-10)
    assert np.linalg.norm(b - G.matvec(x)) / np.linalg.norm(b) < 1e-10


# ---------------------------------------------------------------- block path
def test_block_T1_T2_T3_T4():
    N = 12
    S = gen.block_test_scalar(N)
    As = O.Csr.from_gen(S)
    B = O.Bsr.condense(As, 3, N * N, N * N)
    x = gen.drand(np.arange(S.n))
    # T1: condense/matvec round trip < 1e-12 (test_BlockMatrix_mpi.cpp:203)
    assert np.max(np.abs(B.matvec(x) - As.spmv(x))) < 1e-12
    # T2: <u, A v> = <A^T u, v> (:229)
    u, v = gen.drand(np.arange(S.n) + 11), gen.drand(np.arange(S.n) + 29)
    lhs, rhs = u @ B.matvec(v), B.transpose().matvec(u) @ v
    assert abs(lhs - rhs) / abs(lhs) < 1e-12
    # T3: block-Jacobi solves each diagonal block (:255)
    pj = O.Pc.block(B, 0)
    r = gen.drand(np.arange(S.n) + 41)
    z = pj.apply(r)
    rp, ci, vals = B.export()
    blocks = vals.reshape(-1, 3, 3)
    res = 0.0
    for b in range(N * N):
        p = [q for q in range(rp[b], rp[b + 1]) if ci[q] == b][0]
        res = max(res, np.max(np.abs(blocks[p] @ z[3 * b:3 * b + 3] - r[3 * b:3 * b + 3])))
    assert res < 1e-10
    # T4: mini-CG + block-Jacobi true residual < 1e-6 (:277)
    bb = 1.0 + 0.5 * gen.drand(np.arange(S.n) + 7)
    xs, it = B.solve(0, pj, bb, tol=1e-8, max_iter=3000)
    assert np.linalg.norm(bb - B.matvec(xs)) / np.linalg.norm(bb) < 1e-6


def test_block_ilu0_invariant_and_T7():
    N = 10
    S = gen.block_test_scalar(N)
    B = O.Bsr.condense(O.Csr.from_gen(S), 3, N * N, N * N)
    pi = O.Pc.block(B, 1)
    pj = O.Pc.block(B, 0)
    bb = 1.0 + 0.5 * gen.drand(np.arange(S.n) + 7)
    xj, itj = B.solve(0, pj, bb, tol=1e-8, max_iter=3000)
    xi, iti = B.solve(0, pi, bb, tol=1e-8, max_iter=3000)
    assert np.linalg.norm(bb - B.matvec(xi)) / np.linalg.norm(bb) < 1e-6
    assert iti <= itj  # T7 (test_BlockMatrix_mpi.cpp:441)
    # for a 5-point block pattern, block ILU(0) has L U = A on the block pattern:
    # M^{-1} applied to A's columns restricted to the pattern reproduces them
    Dm = dense(S)
    Minv_A = np.column_stack([pi.apply(Dm[:, j]) for j in range(S.n)])
    # (LU)^{-1} A = I + (LU)^{-1}(A - LU); A - LU vanishes on the pattern, so the
    # diagonal blocks of (LU)^{-1}A - I stay small compared with I
    assert np.max(np.abs(np.diag(Minv_A) - 1.0)) < 0.5


def block_lu_on_pattern_error(Gb, L, U):
    """max over the blocks (I, K) of A's pattern of |(L U)_IK - A_IK| / max|A_IK|,
    L unit block lower (strict part stored), U block upper with diagonal."""
    (lp, lc, lv), (up, uc, uv) = L, U
    Ub = {}
    for J in range(Gb.nbr):
        for q in range(up[J], up[J + 1]):
            Ub[(J, int(uc[q]))] = uv[q]
    blocks = Gb.values.reshape(-1, Gb.bs, Gb.bs)
    worst = 0.0
    for I in range(Gb.nbr):
        for p in range(Gb.row_ptr[I], Gb.row_ptr[I + 1]):
            K = int(Gb.col_idx[p])
            acc = Ub[(I, K)].copy() if K >= I else np.zeros((Gb.bs, Gb.bs))
            for q in range(lp[I], lp[I + 1]):
                UJK = Ub.get((int(lc[q]), K))
                if UJK is not None:
                    acc += lv[q] @ UJK
            worst = max(worst, np.abs(acc - blocks[p]).max() / np.abs(blocks[p]).max())
    return worst


@pytest.mark.parametrize("t,seed", [(4, 3), (5, 8)])
def test_block_ilu0_factors_reproduce_A_on_block_pattern(t, seed):
    """Full block IKJ ILU(0) (blockMatrix.hpp:455-623) gives (L U)_IK = A_IK on
    every block of A's pattern. Random non-commuting 3x3 blocks, so a block
    product taken in the wrong order (U^-1 A instead of A U^-1) fails."""
    Gb = gen.bsr7(t, seed)
    B = O.Bsr(3, Gb.nbr, Gb.nbc, Gb.row_ptr, Gb.col_idx, Gb.values)
    L, U = O.Pc.block(B, 1).block_ilu0_factors(3)
    assert block_lu_on_pattern_error(Gb, L, U) < 1e-12


def test_block_ilu0_bs1_equals_scalar_ilu0_on_stencil():
    G = gen.convdiff2d(8)
    A = O.Csr.from_gen(G)
    B = O.Bsr.condense(A, 1, G.n, G.n)
    r = np.random.default_rng(0).standard_normal(G.n)
    np.testing.assert_allclose(O.Pc.block(B, 1).apply(r), O.Pc(A, O.K_ILU0).apply(r), rtol=1e-12)


def test_block_gmres_blockilu():
    N = 8
    S = gen.block_test_scalar(N)
    B = O.Bsr.condense(O.Csr.from_gen(S), 3, N * N, N * N)
    b = gen.drand(np.arange(S.n))
    x, it = B.solve(1, O.Pc.block(B, 1), b, tol=1e-10, max_iter=500, restart=20)
    assert np.linalg.norm(b - B.matvec(x)) / np.linalg.norm(b) < 1e-10


def test_reference_saamg_degenerates_on_lognormal_coefficients():
    """The reference's block SA-AMG with its fixed weights 2/3 stalls near the
    coarse limit and its Galerkin operators blow up on a log-normal
    coefficient field (blockMatrix.hpp:1051-1179 restated); the robust variant
    (weights 4/(3 rho), stall stop) keeps every level SPD and PCG converges."""
    G = gen.vardiff3d(14, seed=3)[0]
    A = O.Csr.from_gen(G)
    ref = O.Pc(A, O.K_SAAMG, [0.25, 2 / 3, 1, 1, 20, 4, 50, 2 / 3, 1])
    rows, _ = ref.amg_levels()
    n, m, rp, ci, v = ref.amg_matrix(len(rows) - 1, 0)
    ev = np.linalg.eigvalsh(sp.csr_matrix((v, ci, rp), shape=(n, m)).toarray())
    assert ev.max() > 1e20 and ev.min() < 0
    rob = O.Pc(A, O.K_SAAMG, [0.25, 0.0, 1, 1, 20, 4, 50, 0.0, 1])
    rows, _ = rob.amg_levels()
    for l in range(len(rows)):
        n, m, rp, ci, v = rob.amg_matrix(l, 0)
        if n <= 3000:
            Ad = sp.csr_matrix((v, ci, rp), shape=(n, m)).toarray()
            assert np.linalg.eigvalsh((Ad + Ad.T) / 2).min() > 0
    _, b = gen.rhs(G, 0)
    x, it = O.solve(A, 0, rob, b, tol=1e-10)
    assert np.linalg.norm(b - G.matvec(x)) / np.linalg.norm(b) < 1e-10
    assert it < 60


def test_unpreconditioned_bicgstab_depends_on_reduction_order():
    """On convection-diffusion with cell Peclet 5 (test_SMatrix_bicgstab.cpp's
    matrix, N=200) the reference's BiCGSTAB without a preconditioner goes
    through near-breakdowns: with one OpenMP thread it converges, while with
    several threads the dot products are summed in an order that changes from
    run to run (the OpenMP reduction combines partial sums as threads finish),
    so the iterates differ and the run may stall or break down. Parity for such
    runs cannot be iterate-by-iterate."""
    N, v = 200, 5.0
    rows, cols, vals = [], [], []
    for i in range(N):
        for j in range(N):
            k = i * N + j
            rows.append(k); cols.append(k); vals.append(4 + 2 * v)
            if j > 0: rows.append(k); cols.append(k - 1); vals.append(-1 - v)
            if j < N - 1: rows.append(k); cols.append(k + 1); vals.append(-1.0)
            if i > 0: rows.append(k); cols.append(k - N); vals.append(-1 - v)
            if i < N - 1: rows.append(k); cols.append(k + N); vals.append(-1.0)
    A = O.Csr.from_coo(N * N, N * N, vals, rows, cols)
    rp, ci, vv = A.export()
    S = sp.csr_matrix((vv, ci, rp), shape=(N * N, N * N))
    b = np.ones(N * N)
    xs = {}
    try:
        for t in (1, 8):
            O.set_threads(t)
            try:
                xs[t], _ = O.solve_named(A, 1, "none", b, 1e-10, 1400)
            except O.OracleError:
                xs[t] = None
    finally:
        O.set_threads(8)
    assert np.linalg.norm(b - S @ xs[1]) / np.linalg.norm(b) < 1e-10
    assert xs[8] is None or not np.array_equal(xs[1], xs[8])

def _strength(G, theta):
    """Strong couplings |a_ij| >= theta sqrt(m_i m_j), m_i = max_{k!=i} |a_ik|."""
    S = sp.csr_matrix((np.abs(G.values), G.col_idx.astype(np.int64), G.row_ptr.astype(np.int64)), shape=(G.n, G.n))
    S.setdiag(0)
    S.eliminate_zeros()
    m = np.asarray(S.max(axis=1).todense()).ravel()
    C = S.tocoo()
    keep = C.data >= theta * np.sqrt(m[C.row] * m[C.col])
    return sp.csr_matrix((np.ones(keep.sum()), (C.row[keep], C.col[keep])), shape=(G.n, G.n))


@pytest.mark.parametrize("case", ["poisson3d", "vardiff3d", "poisson2d"])
def test_mis2_aggregation_properties(case):
    """Robust SA aggregation: roots form a distance-2 maximal independent set
    of the strength graph, each aggregate holds exactly one root, root
    aggregates are numbered in index order, and every member lies within two
    strong hops of its root."""
    G = {"poisson3d": lambda: gen.poisson3d(9), "vardiff3d": lambda: gen.vardiff3d(9, seed=5)[0],
         "poisson2d": lambda: gen.poisson2d(30)}[case]()
    theta = 0.25
    agg, root, na = O.mis2_aggregate(O.Csr.from_gen(G), theta)
    S = _strength(G, theta)
    assert (S != S.T).nnz == 0                      # symmetric strength on SPD input
    I = sp.identity(G.n, format="csr")
    R1 = (I + S).astype(bool).astype(np.int64)
    R2 = (R1 @ R1).astype(bool)                     # within two strong hops
    ridx = np.flatnonzero(root)
    assert R2[ridx][:, ridx].nnz == len(ridx)       # independence: only the diagonal
    assert np.all(np.asarray(R2[:, ridx].sum(axis=1)).ravel() >= 1)   # maximality
    assert np.array_equal(agg[ridx], np.arange(len(ridx)))
    assert agg.min() == 0 and agg.max() == na - 1 and len(np.unique(agg)) == na
    assert na == len(ridx)                          # no leftover singletons
    root_of = np.empty(na, np.int64)
    root_of[agg[ridx]] = ridx
    own = root_of[agg]
    assert np.all(np.asarray(R2[np.arange(G.n), own]).ravel())
    assert G.n / na > (6 if case != "poisson2d" else 3)   # coarsening ratio


def _greedy_colours(M):
    """Greedy colouring of pattern(M + M^T) in row order, written from its
    definition: row i takes the least colour none of its neighbours j < i
    holds."""
    S = (abs(M) + abs(M.T)).tocsr()
    n = S.shape[0]
    col = np.full(n, -1)
    for i in range(n):
        nb = S.indices[S.indptr[i]:S.indptr[i + 1]]
        used = {col[j] for j in nb if j < i}
        c = 0
        while c in used:
            c += 1
        col[i] = c
    return col


def _sgs_vcycle_dense(pc, r, npre, npost, w_fixed):
    """The SGS V-cycle rebuilt from the oracle's exported level operators
    with dense algebra: a colour's rows are updated together (they share no
    entry), the coarsest level is solved by numpy."""
    rows, _ = pc.amg_levels()
    L = len(rows)
    A, P, R = [], [], []
    for l in range(L):
        n, m, rp, ci, v = pc.amg_matrix(l, 0)
        A.append(sp.csr_matrix((v, ci, rp), shape=(n, m)))
        if l + 1 < L:
            n, m, rp, ci, v = pc.amg_matrix(l, 1)
            P.append(sp.csr_matrix((v, ci, rp), shape=(n, m)).toarray())
            n, m, rp, ci, v = pc.amg_matrix(l, 2)
            R.append(sp.csr_matrix((v, ci, rp), shape=(n, m)).toarray())

    def cycle(l, f):
        Ad = A[l].toarray()
        if l == L - 1:
            return np.linalg.solve(Ad, f)
        d = np.diag(Ad)
        col = _greedy_colours(A[l])
        nc = col.max() + 1
        u = np.zeros_like(f)
        if nc <= 8:
            for c in list(range(nc)) * npre:
                k = col == c
                u[k] += w_fixed * (f[k] - Ad[k] @ u) / d[k]
        else:
            rho = np.max(np.abs(Ad / d[:, None]).sum(axis=1))
            for _ in range(npre):
                u += (4 / 3) / rho * (f - Ad @ u) / d
        u += P[l] @ cycle(l + 1, R[l] @ (f - Ad @ u))
        if nc <= 8:
            for c in list(range(nc - 1, -1, -1)) * npost:
                k = col == c
                u[k] += w_fixed * (f[k] - Ad[k] @ u) / d[k]
        else:
            for _ in range(npost):
                u += (4 / 3) / rho * (f - Ad @ u) / d
        return u

    return cycle(0, r)


def test_greedy_colouring_of_grid_is_red_black():
    """On a 3-D 7-point grid in natural order the greedy colouring is the
    parity of i + j + k."""
    N = 6
    G = gen.poisson3d(N)
    col = _greedy_colours(G.to_scipy())
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    assert np.array_equal(col, ((i + j + k) % 2).ravel())


@pytest.mark.parametrize("case", ["poisson3d", "vardiff3d"])
def test_saamg_sgs_vcycle_matches_dense_restatement(case):
    """The oracle's SGS V-cycle (robust SA hierarchy, 2 pre / 2 post sweeps,
    weight 1) equals the V-cycle rebuilt from its exported operators by dense
    algebra, and is a symmetric operator."""
    G = {"poisson3d": lambda: gen.poisson3d(10), "vardiff3d": lambda: gen.vardiff3d(10, seed=2)[0]}[case]()
    A = O.Csr.from_gen(G)
    pc = O.Pc(A, O.K_SAAMG, [0.25, 1.0, 2, 2, 20, 4, 50, 0.0, 1, 1])
    rows, _ = pc.amg_levels()
    assert len(rows) >= 2
    rng = np.random.default_rng(5)
    r = rng.standard_normal(G.n)
    z = pc.apply(r)
    zd = _sgs_vcycle_dense(pc, r, 2, 2, 1.0)
    assert np.linalg.norm(z - zd) < 1e-12 * np.linalg.norm(zd)
    u, v = rng.standard_normal(G.n), rng.standard_normal(G.n)
    a, b = u @ pc.apply(v), v @ pc.apply(u)
    assert abs(a - b) < 1e-11 * abs(a)
    # the smoother option takes effect (the Jacobi V-cycle differs)
    pj = O.Pc(A, O.K_SAAMG, [0.25, 1.0, 2, 2, 20, 4, 50, 0.0, 1, 0])
    assert np.linalg.norm(pj.apply(r) - z) > 1e-6 * np.linalg.norm(z)


# ---------------------------------------------------------------- Schur complement
def test_schur_preconditioner_block_triangular_identity():
    """make_schur_preconditioner (blockMatrix.hpp:625-685) with an exact
    velocity solve (ILU(0) of a tridiagonal A has no fill) and S_pc = D^-1
    (Jacobi of a diagonal matrix), scaled by c: z = M^-1 r for the dense
    M = [[A, B^T], [B, B A^-1 B^T - D / c]], which follows from
    z_p = c S_pc(B A_pc(r_u) - r_p), z_u = A_pc(r_u - B^T z_p). A flipped
    sign, a missing scale or B in place of B^T breaks it."""
    nv, npres, c = 12, 5, 1.7
    A = sp.diags([-np.ones(nv - 1), 2.5 * np.ones(nv), -np.ones(nv - 1)], [-1, 0, 1]).tocsr()
    rng = np.random.default_rng(3)
    Bd = rng.uniform(-1, 1, (npres, nv)) * (rng.random((npres, nv)) < 0.4)
    B, Bt = sp.csr_matrix(Bd), sp.csr_matrix(Bd.T)
    d = rng.uniform(1.0, 3.0, npres)
    D = sp.diags(d).tocsr()
    oc = lambda M: O.Csr(M.shape[0], M.shape[1], M.indptr, M.indices, M.data)
    Ao, Bo, Bto, Do = oc(A), oc(B), oc(Bt), oc(D)
    pa, pd = O.Pc(Ao, O.K_ILU0), O.Pc(Do, O.K_DIAG)
    S = O.Pc.schur(Bo, Bto, pa, pd, c)
    Ad = A.toarray()
    M = np.block([[Ad, Bd.T], [Bd, Bd @ np.linalg.inv(Ad) @ Bd.T - np.diag(d) / c]])
    for seed in range(3):
        r = np.random.default_rng(10 + seed).standard_normal(nv + npres)
        z = S.apply(r)
        assert np.linalg.norm(M @ z - r) < 1e-12 * np.linalg.norm(r) * np.linalg.cond(M)
    # S_pc absent: the identity, z_p = c (B A^-1 r_u - r_p)
    S0 = O.Pc.schur(Bo, Bto, pa, None, c)
    M0 = np.block([[Ad, Bd.T], [Bd, Bd @ np.linalg.inv(Ad) @ Bd.T - np.eye(npres) / c]])
    r = rng.standard_normal(nv + npres)
    assert np.linalg.norm(M0 @ S0.apply(r) - r) < 1e-12 * np.linalg.norm(r) * np.linalg.cond(M0)


def test_schur_gmres_on_stokes3d():
    """The reference test (test_BlockMatrix_schur3d.cpp:134-193): GMRES(100)
    with the Schur preconditioner (ILU(0) velocity block, scaled SA-AMG of the
    pressure Laplacian) solves the 3-D Stokes system. On the periodic grid
    K has the null vector (0, constant pressure), so the solution matches the
    drand one up to a pressure constant: velocity and mean-free pressure
    errors stay below the test's 1e-6."""
    S3 = gen.stokes3d(6)
    K, A, B, Bt, Lp = (O.Csr.from_gen(S3[k]) for k in ("K", "A", "B", "Bt", "Lp"))
    pa = O.Pc(A, O.K_ILU0)
    pl = O.Pc(Lp, O.K_SAAMG, [0.25, 0.0, 1, 1, 20, 4, 50, 0.0, 1])
    S = O.Pc.schur(B, Bt, pa, pl, 6.1)
    xe = gen.drand(np.arange(K.n))
    b = S3["K"].matvec(xe)
    nv = S3["nv"]
    for Sp in (S, O.Pc.schur(B, Bt, pa, None, 1.0)):
        x, it = O.solve(K, 2, Sp, b, tol=1e-10, max_iter=2000, restart=100)
        d = x - xe
        assert np.linalg.norm(b - S3["K"].matvec(x)) < 1e-10 * np.linalg.norm(b)
        assert np.linalg.norm(d[:nv]) < 1e-6
        assert np.linalg.norm(d[nv:] - d[nv:].mean()) < 1e-6



# ---- block SA-AMG (BS = 3) pins ---------------------------------------------
_REF_MODE = [0.25, 2 / 3, 1, 1, 20, 4, 50, 2 / 3, 1]   # the reference's weights, exact coarse solve
_ROBUST = [0.25, 0.0, 1, 1, 20, 4, 50, 0.0, 1]


def _level(pc, l, which):
    n, m, rp, ci, v = pc.amg_matrix(l, which)
    return sp.csr_matrix((v, ci, rp), shape=(n, m))


@pytest.mark.parametrize("params", [_REF_MODE, _ROBUST], ids=["reference", "robust"])
def test_block_amg_bs3_galerkin_on_noncommuting_blocks(params):
    """make_block_amg_preconditioner<double,3> (blockMatrix.hpp:868-1179) on
    random non-commuting 3x3 blocks: every level's restriction is the block
    transpose of its prolongator and the next operator is the Galerkin
    product R A P (scipy on the expanded matrices). A block transposed in
    place, or a block product taken in the wrong order, fails."""
    Gb = gen.bsr7(6, seed=4)
    B = O.Bsr(3, Gb.nbr, Gb.nbc, Gb.row_ptr, Gb.col_idx, Gb.values)
    pc = O.Pc.block_amg(B, params)
    rows, _ = pc.amg_levels()
    assert len(rows) >= 2
    np.testing.assert_allclose(_level(pc, 0, 0).toarray(), Gb.to_scipy().toarray(), rtol=0, atol=0)
    for l in range(len(rows) - 1):
        A, P, R = _level(pc, l, 0), _level(pc, l, 1), _level(pc, l, 2)
        assert abs(R - P.T.tocsr()).max() == 0
        Ac = _level(pc, l + 1, 0)
        G = (R @ A @ P).toarray()
        assert np.abs(Ac.toarray() - G).max() <= 1e-12 * np.abs(G).max()


def test_block_amg_bs3_is_scalar_amg_kron_k():
    """For A = A1 (x) K (block (i, j) = a_ij K, K fixed and non-symmetric) the
    block hierarchy is the scalar one (x) K: Frobenius strength and
    aggregates are the scalar ones, D^-1 A = (D1^-1 A1) (x) I, so P = P1 (x) I
    and A_c = (P1^T A1 P1) (x) K; the V-cycle with block-Jacobi smoothing is
    V1 (x) K^-1, V1 the (pinned) BS = 1 V-cycle. A wrong block inverse or a
    K^-1 applied from the wrong side fails."""
    A1 = gen.poisson3d(6)
    K = np.array([[4.0, 1.0, 0.5], [0.3, 5.0, 1.0], [0.2, 0.4, 6.0]])
    v3 = (A1.values[:, None] * K.ravel()[None, :]).ravel()
    B3 = O.Bsr(3, A1.n, A1.n, A1.row_ptr, A1.col_idx, v3)
    B1 = O.Bsr(1, A1.n, A1.n, A1.row_ptr, A1.col_idx, A1.values)
    p3, p1 = O.Pc.block_amg(B3, _REF_MODE), O.Pc.block_amg(B1, _REF_MODE)
    r3, n3 = p3.amg_levels()
    r1, n1 = p1.amg_levels()
    assert len(r1) >= 2 and np.array_equal(r3, r1) and np.array_equal(n3, n1)
    for l in range(len(r1)):
        Ac3 = _level(p3, l, 0).toarray()
        Ac1 = _level(p1, l, 0).toarray()
        np.testing.assert_allclose(Ac3, np.kron(Ac1, K), rtol=1e-12, atol=1e-12 * np.abs(Ac3).max())
    R = np.random.default_rng(3).standard_normal((A1.n, 3))
    Y = np.stack([p1.apply(R[:, c]) for c in range(3)], axis=1)
    Z = Y @ np.linalg.inv(K).T
    z3 = p3.apply(R.ravel())
    np.testing.assert_allclose(z3, Z.ravel(), rtol=1e-11, atol=1e-11 * np.abs(Z).max())
