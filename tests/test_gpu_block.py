"""GPU parity for block::BlockMatrix: condense, matvec, transpose,
block-Jacobi, block ILU(0), mini_cg and BlockMatrix::gmres, against the
oracle, with the reference's own thresholds (test_BlockMatrix_mpi.cpp T1-T7)."""
import ctypes as C

import numpy as np
import pytest

import femgpu as F
import oracle_api as O
from femgpu import gen

pytestmark = pytest.mark.gpu


class Bsr:
    """Small test-side handle over the fem_bsr_* entry points."""

    def __init__(self, h):
        self.h = h
        bs, nbr, nbc, nz = C.c_int(), C.c_uint32(), C.c_uint32(), C.c_uint64()
        F._check(F.lib().fem_bsr_info(h, C.byref(bs), C.byref(nbr), C.byref(nbc), C.byref(nz)))
        self.bs, self.nbr, self.nbc, self.nnzb = bs.value, nbr.value, nbc.value, nz.value
        self.nrows = self.nbr * self.bs
        self._h = h

    @classmethod
    def condense(cls, A: F.CsrMatrix, bs, nbr, nbc):
        h = C.c_void_p()
        F._check(F.lib().fem_bsr_condense(A._h, bs, nbr, nbc, C.byref(h)))
        return cls(h)

    @classmethod
    def create(cls, bs, nbr, nbc, rp, ci, v):
        rp, ci, v = F._u32(rp), F._u32(ci), F._f64(v)
        h = C.c_void_p()
        F._check(F.lib().fem_bsr_create(bs, nbr, nbc, F._p(rp, C.c_uint32), F._p(ci, C.c_uint32),
                                        F._p(v, C.c_double), C.byref(h)))
        return cls(h)

    def export(self):
        rp = np.empty(self.nbr + 1, np.uint32)
        ci = np.empty(self.nnzb, np.uint32)
        v = np.empty(self.nnzb * self.bs * self.bs)
        F._check(F.lib().fem_bsr_export(self.h, F._p(rp, C.c_uint32), F._p(ci, C.c_uint32), F._p(v, C.c_double)))
        return rp, ci, v

    def transpose(self):
        h = C.c_void_p()
        F._check(F.lib().fem_bsr_transpose(self.h, C.byref(h)))
        return Bsr(h)

    def matvec(self, x):
        x = F._f64(x)
        y = np.empty(self.nrows)
        F._check(F.lib().fem_bsr_matvec(self.h, F._p(x, C.c_double), F._p(y, C.c_double)))
        return y

    def __del__(self):
        try:
            F.lib().fem_bsr_destroy(self.h)
        except Exception:
            pass


def setup(N):
    S = gen.block_test_scalar(N)
    A = F.CsrMatrix.from_gen(S)
    B = Bsr.condense(A, 3, N * N, N * N)
    Bo = O.Bsr.condense(O.Csr.from_gen(S), 3, N * N, N * N)
    return S, A, B, Bo


@pytest.mark.parametrize("N", [16, 128])
def test_T1_T2_condense_matvec_transpose(N):
    S, A, B, Bo = setup(N)
    for a, b in zip(B.export(), Bo.export()):
        assert np.array_equal(a, b)
    x = gen.drand(np.arange(S.n))
    y = B.matvec(x)
    assert np.array_equal(y, Bo.matvec(x))
    assert np.max(np.abs(y - A.spmv(x))) < 1e-12                     # T1
    u, v = gen.drand(np.arange(S.n) + 11), gen.drand(np.arange(S.n) + 29)
    lhs, rhs = u @ B.matvec(v), B.transpose().matvec(u) @ v
    assert abs(lhs - rhs) / abs(lhs) < 1e-12                         # T2
    for a, b in zip(B.transpose().export(), Bo.transpose().export()):
        assert np.array_equal(a, b)


def test_T3_T4_block_jacobi_minicg():
    N = 64
    S, A, B, Bo = setup(N)
    M = F.Preconditioner(B, F.PC_BJACOBI)
    r = gen.drand(np.arange(S.n) + 41)
    z = M.apply(r)
    assert np.array_equal(z, O.Pc.block(Bo, 0).apply(r))
    rp, ci, vals = B.export()
    blocks = vals.reshape(-1, 3, 3)
    res = 0.0
    for b in range(0, N * N, 7):
        p = [q for q in range(rp[b], rp[b + 1]) if ci[q] == b][0]
        res = max(res, np.max(np.abs(blocks[p] @ z[3 * b:3 * b + 3] - r[3 * b:3 * b + 3])))
    assert res < 1e-10                                                # T3
    bb = 1.0 + 0.5 * gen.drand(np.arange(S.n) + 7)
    x, st = F.solve(B, M, bb, solver="minicg", tol=1e-8, max_iter=3000)
    xo, ito = Bo.solve(0, O.Pc.block(Bo, 0), bb, tol=1e-8, max_iter=3000)
    assert np.linalg.norm(bb - B.matvec(x)) / np.linalg.norm(bb) < 1e-6   # T4
    assert abs(st["iterations"] - ito) <= 1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-7


def test_T6_T7_block_ilu0():
    N = 64
    S, A, B, Bo = setup(N)
    Mi = F.Preconditioner(B, F.PC_BILU0)
    Mo = O.Pc.block(Bo, 1)
    r = np.random.default_rng(0).standard_normal(S.n)
    assert np.array_equal(Mi.apply(r), Mo.apply(r))
    bb = 1.0 + 0.5 * gen.drand(np.arange(S.n) + 7)
    xj, sj = F.solve(B, F.Preconditioner(B, F.PC_BJACOBI), bb, solver="minicg", tol=1e-8, max_iter=3000)
    xi, si = F.solve(B, Mi, bb, solver="minicg", tol=1e-8, max_iter=3000)
    assert np.linalg.norm(bb - B.matvec(xi)) / np.linalg.norm(bb) < 1e-6
    assert si["iterations"] <= sj["iterations"]                      # T7


@pytest.mark.parametrize("bs", [1, 3])
def test_condense_random_pattern(bs):
    """BlockMatrix::condense on the device against the oracle (block sizes of
    the reference's instantiations) on a random scalar CSR with unsorted and
    duplicate columns, which the CSR ctor keeps."""
    rng = np.random.default_rng(bs)
    nb = 300
    n = nb * bs
    rows = []
    for i in range(n):
        cols = rng.integers(0, n, rng.integers(0, 12))
        rows.append(cols)
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum([len(cc) for cc in rows])
    ci = np.concatenate(rows).astype(np.uint32)
    v = rng.uniform(-1, 1, ci.size)
    A = F.CsrMatrix(n, n, rp, ci, v)
    B = Bsr.condense(A, bs, nb, nb)
    Bo = O.Bsr.condense(O.Csr(n, n, rp, ci, v), bs, nb, nb)
    for a, b in zip(B.export(), Bo.export()):
        assert np.array_equal(a, b)


def test_block_ilu0_random_bsr_bitexact():
    Bg = gen.bsr7(6, seed=3)
    B = Bsr.create(3, Bg.nbr, Bg.nbc, Bg.row_ptr, Bg.col_idx, Bg.values)
    Bo = O.Bsr(3, Bg.nbr, Bg.nbc, Bg.row_ptr, Bg.col_idx, Bg.values)
    r = np.random.default_rng(1).standard_normal(Bg.nbr * 3)
    assert np.array_equal(F.Preconditioner(B, F.PC_BILU0).apply(r), O.Pc.block(Bo, 1).apply(r))
    assert np.array_equal(B.matvec(r), Bo.matvec(r))


@pytest.mark.parametrize("pc", [F.PC_BILU0, F.PC_BJACOBI, F.PC_NONE])
def test_block_gmres(pc):
    N = 24
    S, A, B, Bo = setup(N)
    b = gen.drand(np.arange(S.n))
    x, st = F.solve(B, F.Preconditioner(B, pc), b, solver="gmres", restart=20, tol=1e-10, max_iter=600)
    assert st["true_relres"] < 1e-10
    kind = {F.PC_BILU0: 1, F.PC_BJACOBI: 0}.get(pc)
    if kind is not None:
        xo, ito = Bo.solve(1, O.Pc.block(Bo, kind), b, tol=1e-10, max_iter=600, restart=20)
        assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8


def test_bs1_block_saamg_matches_scalar_saamg():
    G = gen.poisson3d(10)
    A = F.CsrMatrix.from_gen(G)
    B = Bsr.condense(A, 1, G.n, G.n)
    r = np.random.default_rng(2).standard_normal(G.n)
    z1 = F.Preconditioner(B, F.PC_BSAAMG).apply(r)
    z2 = F.Preconditioner(A, "saamg").apply(r)
    assert np.array_equal(z1, z2)
    # block AMG with the library defaults as a GMRES preconditioner (the
    # reference's schur3d use) on a 3-D Laplacian
    G2 = gen.poisson3d(24)
    A2 = F.CsrMatrix.from_gen(G2)
    B2 = Bsr.condense(A2, 1, G2.n, G2.n)
    _, b = gen.rhs(G2, 0)
    x, st = F.solve(B2, F.Preconditioner(B2, F.PC_BSAAMG), b, solver="gmres", restart=100, max_iter=300)
    assert st["true_relres"] < 1e-10 and st["iterations"] < 30


@pytest.mark.parametrize("mode", ["reference", "robust"])
def test_block_saamg_bs3_matches_oracle(mode):
    """Block SA-AMG on 3x3 blocks (make_block_amg_preconditioner<double,3>,
    blockMatrix.hpp:1051-1179) against the oracle's PcBlockAMG<3> on random
    non-commuting blocks: the same level sizes and block counts, one V-cycle
    within 1e-12, and GMRES solutions within 1e-8 of the oracle's. Reference
    mode: the reference's weights 2/3, greedy aggregates, its coarsest
    elimination sequence; robust mode: the library defaults."""
    Bg = gen.bsr7(10, seed=4)
    B = Bsr.create(3, Bg.nbr, Bg.nbc, Bg.row_ptr, Bg.col_idx, Bg.values)
    Bo = O.Bsr(3, Bg.nbr, Bg.nbc, Bg.row_ptr, Bg.col_idx, Bg.values)
    if mode == "reference":
        M = F.Preconditioner(B, F.PC_BSAAMG, omega=2 / 3, p_omega=2 / 3, coarse_exact=0)
        Mo = O.Pc.block_amg(Bo, [0.25, 2 / 3, 1, 1, 20, 4, 50, 2 / 3, 0])
    else:
        M = F.Preconditioner(B, F.PC_BSAAMG)
        Mo = O.Pc.block_amg(Bo, [0.25, 0.0, 1, 1, 20, 4, 50, 0.0, 1])
    info = M.info()
    rows, nnz = Mo.amg_levels()
    print(f"{mode}: levels {list(rows)} block nnz {list(nnz)}", flush=True)
    assert info["levels"] == len(rows) >= 2
    assert np.array_equal(info["rows"][: len(rows)], rows)
    assert np.array_equal(info["nnz"][: len(rows)], nnz)
    r = np.random.default_rng(7).standard_normal(Bg.nbr * 3)
    zo = Mo.apply(r)
    np.testing.assert_allclose(M.apply(r), zo, rtol=1e-12, atol=1e-12 * np.abs(zo).max())
    b = np.random.default_rng(8).uniform(-1, 1, Bg.nbr * 3)
    x, st = F.solve(B, M, b, solver="gmres", restart=30, tol=1e-10, max_iter=500)
    xo, ito = Bo.solve(1, Mo, b, tol=1e-10, max_iter=500, restart=30)
    print(f"{mode}: gmres iterations gpu {st['iterations']} oracle {ito}", flush=True)
    assert st["true_relres"] < 1e-10
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8
    assert abs(st["iterations"] - ito) <= 2


def test_block_saamg_bs3_fp32_refinement():
    """The BS = 3 hierarchy in float32 inside float64 refinement (PCG)."""
    Bg = gen.bsr7(12, seed=5)
    B = Bsr.create(3, Bg.nbr, Bg.nbc, Bg.row_ptr, Bg.col_idx, Bg.values)
    M = F.Preconditioner(B, F.PC_BSAAMG, fp32=1)
    b = np.random.default_rng(9).uniform(-1, 1, Bg.nbr * 3)
    x, st = F.solve(B, M, b, solver="pcg", precision=F.FP32_IR, max_iter=500)
    A = Bg.to_scipy()
    assert st["converged"] == 1
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10


def test_block_errors():
    # missing diagonal block -> block ILU(0) failure
    B = Bsr.create(2, 2, 2, [0, 1, 2], [1, 0], np.arange(8, dtype=float) + 1)
    with pytest.raises(F.SolverFailure):
        F.Preconditioner(B, F.PC_BILU0)
    with pytest.raises(F.OutOfRange):
        Bsr.create(2, 2, 2, [0, 1, 2], [1, 5], np.ones(8))
