"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the GPU library's results are checked against the
oracle, which is never part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "oracle", "liboracle.so")

K_NONE, K_DIAG, K_IC0, K_ILU0, K_SSOR, K_AMG, K_SAAMG = range(7)
_L = None
P = C.c_void_p


def lib():
    global _L
    if _L is None:
        if not os.path.exists(SO):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = C.CDLL(SO)
        L.or_last_error.restype = C.c_char_p
        L.or_csr_nnz.restype = C.c_uint64
        L.or_csr_rows.restype = C.c_uint32
        L.or_csr_cols.restype = C.c_uint32
        L.or_ic0_nnz.restype = C.c_uint64
        L.or_ssor_export.restype = C.c_uint32
        L.or_bsr_nnzb.restype = C.c_uint64
        for nm in ("or_csr_destroy", "or_pc_destroy", "or_bsr_destroy"):
            getattr(L, nm).argtypes = [P]
        L.or_csr_nnz.argtypes = [P]
        L.or_csr_rows.argtypes = [P]
        L.or_csr_cols.argtypes = [P]
        L.or_ic0_nnz.argtypes = [P]
        L.or_bsr_nnzb.argtypes = [P]
        L.or_ssor_export.argtypes = [P, P, P]
        L.or_dot.argtypes = [C.c_uint64, P, P, C.POINTER(C.c_double)]
        L.or_solve.argtypes = [P, C.c_int, P, P, C.c_double, C.c_int, C.c_int, P, C.POINTER(C.c_int)]
        L.or_solve_named.argtypes = [P, C.c_int, C.c_char_p, P, C.c_double, C.c_int, C.c_int, P, C.POINTER(C.c_int)]
        L.or_bsr_solve.argtypes = [P, C.c_int, P, P, C.c_double, C.c_int, C.c_int, P, C.POINTER(C.c_int)]
        L.or_alias.argtypes = [C.c_char_p, C.c_int]
        L.or_amg_levels.argtypes = [P, P, P, C.c_int]
        L.or_amg_level_export.argtypes = [P, C.c_int, C.c_int, P, P, P, P]
        L.or_mis2_aggregate.argtypes = [P, C.c_double, P, P, C.POINTER(C.c_int)]
        L.or_amg_cf_split.argtypes = [P, C.c_int, P]
        L.or_bsr_pc_create_amg.argtypes = [P, P, P]
        L.or_pc_create_schur.argtypes = [P, P, P, P, C.c_double, P]
        L.or_bilu0_export.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), P, P, P, P, P, P]
        L.or_spai_features.argtypes = [C.c_uint32, P, P, P, P, P, P]
        L.or_mlp_forward.argtypes = [P, C.c_uint32, P, C.c_int, P]
        L.or_spai_assemble.argtypes = [C.c_uint32, P, P, P, P, P, C.c_int, P]
        L.or_spai_loss_grad.argtypes = [C.c_uint32, P, P, P, P, P, P, C.c_int, C.c_int, P, P]
        _L = L
    return _L


class OracleError(Exception):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def _chk(rc):
    if rc:
        raise OracleError(rc, lib().or_last_error().decode())


def _a(x, dt):
    return np.ascontiguousarray(x, dtype=dt)


def _ptr(a):
    return a.ctypes.data_as(P)


def set_threads(t: int):
    lib().or_set_threads(t)


class Csr:
    def __init__(self, n, m, rp, ci, v, handle=None):
        self.h = P()
        if handle is not None:
            self.h = handle
        else:
            self._keep = (_a(rp, np.uint32), _a(ci, np.uint32), _a(v, np.float64))
            _chk(lib().or_csr_create(C.c_uint32(n), C.c_uint32(m), _ptr(self._keep[0]), _ptr(self._keep[1]),
                                     _ptr(self._keep[2]), C.byref(self.h)))
        self.n = lib().or_csr_rows(self.h)
        self.m = lib().or_csr_cols(self.h)

    @classmethod
    def from_gen(cls, A):
        return cls(A.n, A.m, A.row_ptr, A.col_idx, A.values)

    @classmethod
    def from_coo(cls, n, m, v, r, c):
        v, r, c = _a(v, np.float64), _a(r, np.uint32), _a(c, np.uint32)
        h = P()
        _chk(lib().or_csr_from_coo(C.c_uint32(n), C.c_uint32(m), C.c_uint64(v.size), _ptr(v), _ptr(r), _ptr(c),
                                   C.byref(h)))
        return cls(0, 0, None, None, None, handle=h)

    @property
    def nnz(self):
        return lib().or_csr_nnz(self.h)

    def export(self):
        rp = np.empty(self.n + 1, np.uint32)
        ci = np.empty(self.nnz, np.uint32)
        v = np.empty(self.nnz, np.float64)
        lib().or_csr_export(self.h, _ptr(rp), _ptr(ci), _ptr(v))
        return rp, ci, v

    def transpose(self):
        h = P()
        _chk(lib().or_csr_transpose(self.h, C.byref(h)))
        return Csr(0, 0, None, None, None, handle=h)

    def spgemm(self, B):
        h = P()
        _chk(lib().or_csr_spgemm(self.h, B.h, C.byref(h)))
        return Csr(0, 0, None, None, None, handle=h)

    def spmv(self, x):
        x = _a(x, np.float64)
        y = np.empty(self.n)
        _chk(lib().or_spmv(self.h, _ptr(x), _ptr(y)))
        return y

    def __del__(self):
        try:
            lib().or_csr_destroy(self.h)
        except Exception:
            pass


def dot(a, b):
    a, b = _a(a, np.float64), _a(b, np.float64)
    out = C.c_double()
    _chk(lib().or_dot(a.size, _ptr(a), _ptr(b), C.byref(out)))
    return out.value


class Pc:
    def __init__(self, A: Csr, kind: int, params=None, n=None):
        self.h = P()
        self.n = A.n if n is None else n
        if params is None:
            pp = None
        else:
            params = list(params)
            if kind == K_SAAMG and len(params) == 9:
                params.append(0)  # Jacobi smoother
            self._pp = np.ascontiguousarray(params, dtype=np.float64)
            pp = _ptr(self._pp)
        if A is not None:
            _chk(lib().or_pc_create(A.h, kind, pp, C.byref(self.h)))

    @classmethod
    def block(cls, B, kind):
        self = cls.__new__(cls)
        self.h = P()
        self.n = B.nrows
        _chk(lib().or_bsr_pc_create(B.h, kind, C.byref(self.h)))
        return self

    @classmethod
    def schur(cls, B, Bt, A_pc, S_pc=None, scale=1.0):
        """make_schur_preconditioner (blockMatrix.hpp:625-685) over oracle
        CSR matrices B, B^T and preconditioners (S_pc None: identity)."""
        self = cls.__new__(cls)
        self.h = P()
        self.n = B.m + B.n
        self._keep = (B, Bt, A_pc, S_pc)
        _chk(lib().or_pc_create_schur(B.h, Bt.h, A_pc.h, S_pc.h if S_pc is not None else None, scale,
                                      C.byref(self.h)))
        return self

    @classmethod
    def block_amg(cls, B, params):
        """Block SA-AMG with explicit parameters: theta, smoother omega, npre,
        npost, max_levels, max_agg, coarse_max, prolongator omega, exact
        coarse solve (weights <= 0 select the robust variant)."""
        self = cls.__new__(cls)
        self.h = P()
        self.n = B.nrows
        self._pp = np.ascontiguousarray(params, dtype=np.float64)
        _chk(lib().or_bsr_pc_create_amg(B.h, _ptr(self._pp), C.byref(self.h)))
        return self

    def apply(self, r):
        r = _a(r, np.float64)
        z = np.empty(r.size)
        _chk(lib().or_pc_apply(self.h, C.c_uint64(r.size), _ptr(r), _ptr(z)))
        return z

    def ic0_factor(self):
        nz = lib().or_ic0_nnz(self.h)
        rp = np.empty(self.n + 1, np.uint32)
        ci = np.empty(nz, np.uint32)
        v = np.empty(nz)
        lib().or_ic0_export(self.h, _ptr(rp), _ptr(ci), _ptr(v))
        return rp, ci, v

    def ilu0_factors(self):
        ln, un = C.c_uint64(), C.c_uint64()
        lib().or_ilu0_sizes(self.h, C.byref(ln), C.byref(un))
        lp = np.empty(self.n + 1, np.uint32)
        up = np.empty(self.n + 1, np.uint32)
        lc = np.empty(ln.value, np.uint32)
        lv = np.empty(ln.value)
        uc = np.empty(un.value, np.uint32)
        uv = np.empty(un.value)
        lib().or_ilu0_export(self.h, _ptr(lp), _ptr(lc), _ptr(lv), _ptr(up), _ptr(uc), _ptr(uv))
        return (lp, lc, lv), (up, uc, uv)

    def ssor_order(self):
        nid = np.empty(self.n, np.uint32)
        nc = lib().or_ssor_export(self.h, _ptr(nid), None)
        cp = np.empty(nc + 1, np.uint32)
        lib().or_ssor_export(self.h, None, _ptr(cp))
        return nid, cp

    def amg_levels(self, cap=64):
        rows = np.zeros(cap, np.uint32)
        nnz = np.zeros(cap, np.uint64)
        L = lib().or_amg_levels(self.h, _ptr(rows), _ptr(nnz), cap)
        return rows[:L], nnz[:L]

    def amg_matrix(self, level, which):
        dims = np.zeros(3, np.uint32)
        lib().or_amg_level_export(self.h, level, which, None, None, None, _ptr(dims))
        n, m, nz = (int(x) for x in dims)
        rp = np.empty(n + 1, np.uint32)
        ci = np.empty(nz, np.uint32)
        v = np.empty(nz)
        lib().or_amg_level_export(self.h, level, which, _ptr(rp), _ptr(ci), _ptr(v), _ptr(dims))
        return n, m, rp, ci, v

    def cf_split(self, level):
        """C/F split of a C/F AMG level: int8[n], 1 = C, 2 = F."""
        n = int(self.amg_levels()[0][level])
        st = np.empty(n, np.int8)
        if lib().or_amg_cf_split(self.h, level, _ptr(st)) != 0:
            raise OracleError(-1, "no C/F split at this level")
        return st

    def block_ilu0_factors(self, bs):
        """Block ILU(0) factors as ((Lp, Lc, Lv), (Up, Uc, Uv)), values as
        (nnzb, bs, bs) arrays: L strict lower, U upper with the diagonal."""
        ln, un = C.c_uint64(), C.c_uint64()
        if lib().or_bilu0_export(self.h, C.byref(ln), C.byref(un), None, None, None, None, None, None) != 0:
            raise OracleError(-1, "not a block ILU(0) preconditioner")
        nb = self.n // bs
        lp, up = np.empty(nb + 1, np.uint32), np.empty(nb + 1, np.uint32)
        lc, uc = np.empty(ln.value, np.uint32), np.empty(un.value, np.uint32)
        lv, uv = np.empty(ln.value * bs * bs), np.empty(un.value * bs * bs)
        lib().or_bilu0_export(self.h, C.byref(ln), C.byref(un), _ptr(lp), _ptr(lc), _ptr(lv), _ptr(up), _ptr(uc),
                              _ptr(uv))
        return (lp, lc, lv.reshape(-1, bs, bs)), (up, uc, uv.reshape(-1, bs, bs))

    def __del__(self):
        try:
            lib().or_pc_destroy(self.h)
        except Exception:
            pass


def solve(A: Csr, solver: int, pc: Pc, b, tol=1e-10, max_iter=2000, restart=30):
    b = _a(b, np.float64)
    x = np.empty(A.n)
    it = C.c_int()
    _chk(lib().or_solve(A.h, solver, pc.h, _ptr(b), tol, max_iter, restart, _ptr(x), C.byref(it)))
    return x, it.value


def solve_named(A: Csr, solver: int, name: str, b, tol=1e-10, max_iter=2000, restart=30):
    b = _a(b, np.float64)
    x = np.empty(A.n)
    it = C.c_int()
    _chk(lib().or_solve_named(A.h, solver, name.encode(), _ptr(b), tol, max_iter, restart, _ptr(x), C.byref(it)))
    return x, it.value


def mis2_aggregate(A: Csr, theta: float):
    """Robust-mode SA aggregation: (agg[n], root flags[n], number of aggregates)."""
    agg = np.empty(A.n, np.int32)
    st = np.empty(A.n, np.int32)
    na = C.c_int()
    _chk(lib().or_mis2_aggregate(A.h, theta, _ptr(agg), _ptr(st), C.byref(na)))
    return agg, st == 2, na.value


# ---- learned sparse approximate inverse (spai_oracle.cpp) -------------------
def spai_nparam() -> int:
    return lib().or_spai_nparam()


def _grid_csr(G, grid):
    rp, ci, v = _a(G.row_ptr, np.uint32), _a(G.col_idx, np.uint32), _a(G.values, np.float64)
    return rp, ci, v, _a(grid, np.uint32)


def spai_features(G, grid):
    rp, ci, v, g = _grid_csr(G, grid)
    feat = np.empty((G.n, 6))
    dis = np.empty(G.n)
    _chk(lib().or_spai_features(G.n, _ptr(rp), _ptr(ci), _ptr(v), _ptr(g), _ptr(feat), _ptr(dis)))
    return feat, dis


def mlp_forward(params, feat, bf16=False):
    params = _a(params, np.float32)
    feat = _a(feat, np.float64)
    out = np.empty((feat.shape[0], 7))
    _chk(lib().or_mlp_forward(_ptr(params), feat.shape[0], _ptr(feat), int(bf16), _ptr(out)))
    return out


def spai_assemble(G, grid, params, bf16=False):
    rp, ci, v, g = _grid_csr(G, grid)
    params = _a(params, np.float32)
    m = np.empty(len(v))
    _chk(lib().or_spai_assemble(G.n, _ptr(rp), _ptr(ci), _ptr(v), _ptr(g), _ptr(params), int(bf16), _ptr(m)))
    return m


def spai_loss_grad(G, grid, params, V, bf16=False):
    """V: (k, n) probe vectors. Returns (loss, grad[nparam])."""
    rp, ci, v, g = _grid_csr(G, grid)
    params = _a(params, np.float32)
    V = _a(V, np.float64)
    k = V.shape[0]
    loss = np.zeros(1)
    grad = np.empty(spai_nparam())
    _chk(lib().or_spai_loss_grad(G.n, _ptr(rp), _ptr(ci), _ptr(v), _ptr(g), _ptr(params), _ptr(V), k, int(bf16),
                                 _ptr(loss), _ptr(grad)))
    return float(loss[0]), grad


def alias(name: str, solver: int) -> int:
    return lib().or_alias(name.encode(), solver)


class Bsr:
    def __init__(self, bs, nbr, nbc, rp, ci, v, handle=None):
        self.h = P()
        self.bs = bs
        if handle is not None:
            self.h = handle
        else:
            self._keep = (_a(rp, np.uint32), _a(ci, np.uint32), _a(v, np.float64))
            _chk(lib().or_bsr_create(bs, C.c_uint32(nbr), C.c_uint32(nbc), _ptr(self._keep[0]), _ptr(self._keep[1]),
                                     _ptr(self._keep[2]), C.byref(self.h)))
        self.nbr, self.nbc = nbr, nbc
        self.nrows = nbr * bs

    @classmethod
    def condense(cls, A: Csr, bs, nbr, nbc):
        h = P()
        _chk(lib().or_bsr_condense(A.h, bs, C.c_uint32(nbr), C.c_uint32(nbc), C.byref(h)))
        return cls(bs, nbr, nbc, None, None, None, handle=h)

    @property
    def nnzb(self):
        return lib().or_bsr_nnzb(self.h)

    def export(self):
        rp = np.empty(self.nbr + 1, np.uint32)
        ci = np.empty(self.nnzb, np.uint32)
        v = np.empty(self.nnzb * self.bs * self.bs)
        lib().or_bsr_export(self.h, _ptr(rp), _ptr(ci), _ptr(v))
        return rp, ci, v

    def transpose(self):
        h = P()
        _chk(lib().or_bsr_transpose(self.h, C.byref(h)))
        return Bsr(self.bs, self.nbc, self.nbr, None, None, None, handle=h)

    def matvec(self, x):
        x = _a(x, np.float64)
        y = np.empty(self.nrows)
        _chk(lib().or_bsr_matvec(self.h, _ptr(x), _ptr(y)))
        return y

    def solve(self, solver, pc: Pc, b, tol=1e-10, max_iter=2000, restart=30):
        b = _a(b, np.float64)
        x = np.empty(self.nrows)
        it = C.c_int()
        _chk(lib().or_bsr_solve(self.h, solver, pc.h, _ptr(b), tol, max_iter, restart, _ptr(x), C.byref(it)))
        return x, it.value

    def __del__(self):
        try:
            lib().or_bsr_destroy(self.h)
        except Exception:
            pass
