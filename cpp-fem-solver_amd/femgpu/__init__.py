"""Python binding (ctypes) of libfemgpu, the MI355X sparse solve path.

The binding calls the C ABI declared in include/femgpu.h. It never computes
anything itself: when the shared library is missing, or no gfx950 device is
present, every call raises (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfemgpu.so")

# enums (femgpu.h)
PC_NONE, PC_DIAG, PC_IC0, PC_ILU0, PC_SSOR, PC_AMG, PC_SAAMG, PC_LEARNED, PC_BJACOBI, PC_BILU0, PC_BSAAMG, PC_SCHUR, \
    PC_SPAI = range(13)
PCG, BICGSTAB, GMRES, MINICG = 0, 1, 2, 3
FP64, FP32_IR = 0, 1
SMOOTH_JACOBI, SMOOTH_SGS = 0, 1

PC_KINDS = {"none": PC_NONE, "diag": PC_DIAG, "ic0": PC_IC0, "ilu0": PC_ILU0, "ssor": PC_SSOR, "amg": PC_AMG,
            "saamg": PC_SAAMG, "learned": PC_LEARNED, "bjacobi": PC_BJACOBI, "bilu0": PC_BILU0, "bsaamg": PC_BSAAMG}
SOLVERS = {"pcg": PCG, "bicgstab": BICGSTAB, "gmres": GMRES, "minicg": MINICG}


class FemError(Exception):
    code = -1


class InvalidArgument(FemError, ValueError):
    code = 1


class OutOfRange(FemError, IndexError):
    code = 2


class SolverFailure(FemError, RuntimeError):
    code = 3


class DeviceError(FemError, RuntimeError):
    code = 4


_ERRS = {1: InvalidArgument, 2: OutOfRange, 3: SolverFailure, 4: DeviceError}


class PcParams(C.Structure):
    _fields_ = [("omega", C.c_double), ("theta", C.c_double), ("npre", C.c_int), ("npost", C.c_int),
                ("max_levels", C.c_int), ("coarse_size", C.c_int), ("max_agg", C.c_int), ("smoother", C.c_int),
                ("p_omega", C.c_double), ("coarse_exact", C.c_int), ("fp32", C.c_int), ("setup_device", C.c_int)]


class SolveOpts(C.Structure):
    _fields_ = [("solver", C.c_int), ("tol", C.c_double), ("max_iter", C.c_int), ("restart", C.c_int),
                ("precision", C.c_int), ("inner_tol", C.c_double), ("max_refine", C.c_int),
                ("check_every", C.c_int), ("device_ptrs", C.c_int), ("use_graph", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("inner_iterations", C.c_int), ("refinements", C.c_int),
                ("converged", C.c_int), ("true_relres", C.c_double), ("setup_ms", C.c_double),
                ("solve_ms", C.c_double), ("spmv_count", C.c_longlong), ("fallback", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchSystem(C.Structure):
    _fields_ = [("bs", C.c_int), ("nb", C.c_uint32), ("row_ptr", C.POINTER(C.c_uint32)),
                ("col_idx", C.POINTER(C.c_uint32)), ("values", C.POINTER(C.c_double))]


class BatchOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("inner_tol", C.c_double), ("max_refine", C.c_int), ("max_inner", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("refinements", C.c_int), ("converged", C.c_int),
                ("max_relres", C.c_double)]


class BatchSchedule(C.Structure):
    _fields_ = [("block", C.c_int), ("wg_per_cu", C.c_int), ("order", C.c_int), ("split", C.c_int)]


class AdamOpts(C.Structure):
    _fields_ = [("lr", C.c_double), ("beta1", C.c_double), ("beta2", C.c_double), ("eps", C.c_double)]


HOST_CB = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_uint64, C.c_void_p)

_lib = None

P_u32 = C.POINTER(C.c_uint32)
P_f64 = C.POINTER(C.c_double)
P_f32 = C.POINTER(C.c_float)
P_vp = C.c_void_p

_SIGS = {
    "fem_last_error": (C.c_char_p, []),
    "fem_device_info": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int)]),
    "fem_pc_params_default": (None, [C.c_int, C.POINTER(PcParams)]),
    "fem_solve_opts_default": (None, [C.c_int, C.POINTER(SolveOpts)]),
    "fem_csr_create": (C.c_int, [C.c_uint32, C.c_uint32, P_u32, P_u32, P_f64, C.POINTER(P_vp)]),
    "fem_csr_create_coo": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint64, P_f64, P_u32, P_u32, C.POINTER(P_vp)]),
    "fem_csr_info": (C.c_int, [P_vp, P_u32, P_u32, C.POINTER(C.c_uint64)]),
    "fem_csr_export": (C.c_int, [P_vp, P_u32, P_u32, P_f64]),
    "fem_csr_transpose": (C.c_int, [P_vp, C.POINTER(P_vp)]),
    "fem_csr_multiply": (C.c_int, [P_vp, P_vp, C.POINTER(P_vp)]),
    "fem_csr_destroy": (C.c_int, [P_vp]),
    "fem_csr_set_row_lanes": (C.c_int, [P_vp, C.c_int]),
    "fem_csr_pack_columns": (C.c_int, [P_vp, C.c_int, C.POINTER(C.c_int)]),
    "fem_spmv": (C.c_int, [P_vp, P_f64, P_f64]),
    "fem_spmv_device": (C.c_int, [P_vp, P_vp, P_vp, C.c_int]),
    "fem_dot": (C.c_int, [C.c_uint64, P_f64, P_f64, P_f64]),
    "fem_dot_device_noalloc": (C.c_int, [C.c_uint64, P_vp, P_vp, P_vp, P_vp]),
    "fem_precond_create": (C.c_int, [P_vp, C.c_int, C.POINTER(PcParams), C.POINTER(P_vp)]),
    "fem_precond_apply": (C.c_int, [P_vp, P_f64, P_f64]),
    "fem_precond_create_schur": (C.c_int, [P_vp, P_vp, P_vp, P_vp, C.c_double, C.POINTER(P_vp)]),
    "fem_precond_destroy": (C.c_int, [P_vp]),
    "fem_precond_info": (C.c_int, [P_vp, C.POINTER(C.c_int), P_u32, C.POINTER(C.c_uint64), C.c_int, P_f64]),
    "fem_precond_bytes": (C.c_int, [P_vp, C.c_int, P_f64]),
    "fem_bench_precond": (C.c_int, [P_vp, C.c_int, C.c_int, P_f64]),
    "fem_bench_stream": (C.c_int, [C.c_uint64, C.c_int, C.c_int, P_f64]),
    "fem_amg_level_export": (C.c_int, [P_vp, C.c_int, C.c_int, P_u32, P_u32, P_u32, P_f64]),
    "fem_amg_split": (C.c_int, [P_vp, C.c_int, C.POINTER(C.c_uint8)]),
    "fem_solve": (C.c_int, [P_vp, P_vp, P_f64, P_f64, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "fem_pcg": (C.c_int, [P_vp, P_vp, P_f64, C.c_double, C.c_int, P_f64, C.POINTER(C.c_int)]),
    "fem_bicgstab": (C.c_int, [P_vp, P_vp, P_f64, C.c_double, C.c_int, P_f64, C.POINTER(C.c_int)]),
    "fem_gmres": (C.c_int, [P_vp, P_vp, P_f64, C.c_int, C.c_double, C.c_int, P_f64, C.POINTER(C.c_int)]),
    "fem_gmres_callback": (C.c_int, [P_vp, HOST_CB, P_vp, P_f64, C.c_int, C.c_double, C.c_int, P_f64,
                                     C.POINTER(C.c_int)]),
    "fem_solver_create": (C.c_int, [P_vp, P_vp, C.POINTER(SolveOpts), C.POINTER(P_vp)]),
    "fem_solver_solve": (C.c_int, [P_vp, P_f64, P_f64, C.POINTER(SolveStats)]),
    "fem_solver_destroy": (C.c_int, [P_vp]),
    "fem_bsr_create": (C.c_int, [C.c_int, C.c_uint32, C.c_uint32, P_u32, P_u32, P_f64, C.POINTER(P_vp)]),
    "fem_bsr_condense": (C.c_int, [P_vp, C.c_int, C.c_uint32, C.c_uint32, C.POINTER(P_vp)]),
    "fem_bsr_info": (C.c_int, [P_vp, C.POINTER(C.c_int), P_u32, P_u32, C.POINTER(C.c_uint64)]),
    "fem_bsr_export": (C.c_int, [P_vp, P_u32, P_u32, P_f64]),
    "fem_bsr_transpose": (C.c_int, [P_vp, C.POINTER(P_vp)]),
    "fem_bsr_matvec": (C.c_int, [P_vp, P_f64, P_f64]),
    "fem_bsr_matvec_device": (C.c_int, [P_vp, P_vp, P_vp, C.c_int]),
    "fem_bsr_precond_create": (C.c_int, [P_vp, C.c_int, C.POINTER(PcParams), C.POINTER(P_vp)]),
    "fem_bsr_solve": (C.c_int, [P_vp, P_vp, P_f64, P_f64, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "fem_bsr_destroy": (C.c_int, [P_vp]),
    "fem_bench_spmv": (C.c_int, [P_vp, C.c_int, C.c_int, P_f64]),
    "fem_bench_spmv_cold": (C.c_int, [P_vp, C.c_int, C.c_int, P_f64]),
    "fem_mlp_create": (C.c_int, [P_f32, C.c_uint64, C.POINTER(P_vp)]),
    "fem_mlp_destroy": (C.c_int, [P_vp]),
    "fem_mlp_get_params": (C.c_int, [P_vp, P_f32]),
    "fem_mlp_set_params": (C.c_int, [P_vp, P_f32]),
    "fem_mlp_forward": (C.c_int, [P_vp, P_f32, C.c_uint32, P_f32]),
    "fem_spai_features": (C.c_int, [P_vp, P_u32, P_f32]),
    "fem_spai_loss_grad": (C.c_int, [P_vp, P_vp, P_u32, P_f64, C.c_int, C.c_int, P_f64, P_f32]),
    "fem_spai_train_step": (C.c_int, [P_vp, P_vp, P_u32, P_f64, C.c_int, C.c_int, C.POINTER(AdamOpts), P_f64]),
    "fem_precond_create_learned": (C.c_int, [P_vp, P_u32, P_vp, C.c_int, C.c_int, C.POINTER(P_vp)]),
    "fem_precond_learned_reference": (C.c_int, [P_vp, C.c_int]),
    "fem_precond_create_spai": (C.c_int, [P_vp, P_vp, C.c_int, C.POINTER(P_vp)]),
    "fem_batch_create": (C.c_int, [C.c_int, C.POINTER(BatchSystem), C.c_int, C.POINTER(P_vp)]),
    "fem_batch_set_rhs": (C.c_int, [P_vp, P_f64]),
    "fem_batch_solve": (C.c_int, [P_vp, C.POINTER(BatchOpts), C.POINTER(BatchStats), P_f64]),
    "fem_batch_get_solution": (C.c_int, [P_vp, P_f64]),
    "fem_batch_tune": (C.c_int, [P_vp, C.POINTER(BatchOpts), C.c_int, C.POINTER(BatchSchedule), P_f64, C.c_int]),
    "fem_batch_set_schedule": (C.c_int, [P_vp, C.POINTER(BatchSchedule)]),
    "fem_batch_destroy": (C.c_int, [P_vp]),
}


def header_symbols() -> list[str]:
    """Function names declared in include/femgpu.h."""
    import re
    hdr = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "femgpu.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(fem_[a-z0-9_]+)\s*\(", txt)))


def lib():
    """Load libfemgpu.so (built by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DeviceError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        msg = lib().fem_last_error().decode("utf-8", "replace")
        raise _ERRS.get(rc, FemError)(msg)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def bench_stream(nbytes: int, cold: bool = False, reps: int = 10) -> float:
    """Average time (ms) of a streaming read of nbytes (fem_bench_stream):
    the device's read-bandwidth ceiling, warm or after a cache flush."""
    ms = C.c_double()
    _check(lib().fem_bench_stream(int(nbytes), int(bool(cold)), int(reps), C.byref(ms)))
    return ms.value


def device_info() -> tuple[str, int]:
    buf = C.create_string_buffer(256)
    cu = C.c_int(0)
    _check(lib().fem_device_info(buf, 256, C.byref(cu)))
    return buf.value.decode(), cu.value


class CsrMatrix:
    """Device CSR matrix handle (sparse::SMatrix)."""

    def __init__(self, rows, cols, row_ptr, col_idx, values, _handle=None):
        self._h = C.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            rp, ci, v = _u32(row_ptr), _u32(col_idx), _f64(values)
            _check(lib().fem_csr_create(rows, cols, _p(rp, C.c_uint32), _p(ci, C.c_uint32), _p(v, C.c_double),
                                        C.byref(self._h)))
        r, c, z = C.c_uint32(), C.c_uint32(), C.c_uint64()
        _check(lib().fem_csr_info(self._h, C.byref(r), C.byref(c), C.byref(z)))
        self.rows, self.cols, self.nnz = r.value, c.value, z.value

    @classmethod
    def from_gen(cls, A):
        return cls(A.n, A.m, A.row_ptr, A.col_idx, A.values)

    @classmethod
    def from_coo(cls, rows, cols, values, rowlist, collist):
        v, r, c = _f64(values), _u32(rowlist), _u32(collist)
        h = C.c_void_p()
        _check(lib().fem_csr_create_coo(rows, cols, v.size, _p(v, C.c_double), _p(r, C.c_uint32),
                                        _p(c, C.c_uint32), C.byref(h)))
        return cls(rows, cols, None, None, None, _handle=h)

    def export(self):
        rp = np.empty(self.rows + 1, np.uint32)
        ci = np.empty(self.nnz, np.uint32)
        v = np.empty(self.nnz, np.float64)
        _check(lib().fem_csr_export(self._h, _p(rp, C.c_uint32), _p(ci, C.c_uint32), _p(v, C.c_double)))
        return rp, ci, v

    def transpose(self):
        h = C.c_void_p()
        _check(lib().fem_csr_transpose(self._h, C.byref(h)))
        return CsrMatrix(0, 0, None, None, None, _handle=h)

    def __matmul__(self, other):
        if isinstance(other, CsrMatrix):
            h = C.c_void_p()
            _check(lib().fem_csr_multiply(self._h, other._h, C.byref(h)))
            return CsrMatrix(0, 0, None, None, None, _handle=h)
        return self.spmv(other)

    def spmv(self, x):
        x = _f64(x)
        if x.size != self.cols:
            raise InvalidArgument("spmv: vector length differs from the matrix columns")
        y = np.empty(self.rows, np.float64)
        _check(lib().fem_spmv(self._h, _p(x, C.c_double), _p(y, C.c_double)))
        return y

    def set_row_lanes(self, lanes: int):
        """SpMV variant: 0 = CSR-stream (bit-exact with SMatrix::spmv),
        4/8/16/32/64 = that many lanes per row."""
        _check(lib().fem_csr_set_row_lanes(self._h, int(lanes)))

    def pack_columns(self, on=-1) -> bool:
        """16-bit packed column offsets for the CSR-stream SpMV: on = 1 builds
        them (where every 64-entry chunk spans < 65536 columns), 0 drops
        them, -1 only asks. Returns whether the packed form is in use."""
        p = C.c_int()
        _check(lib().fem_csr_pack_columns(self._h, int(on), C.byref(p)))
        return bool(p.value)

    def bench_spmv(self, fp32=False, reps=100, cold=False) -> float:
        """Average SpMV time (ms): back-to-back launches, or each launch
        after evicting the caches (cold)."""
        ms = C.c_double()
        f = lib().fem_bench_spmv_cold if cold else lib().fem_bench_spmv
        _check(f(self._h, int(fp32), reps, C.byref(ms)))
        return ms.value

    def close(self):
        if self._h:
            lib().fem_csr_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BsrMatrix:
    """Device block-CSR matrix handle (block::BlockMatrix<double, BS>)."""

    def __init__(self, bs, nbr, nbc, row_ptr, col_idx, values, _handle=None):
        self._h = C.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            rp, ci, v = _u32(row_ptr), _u32(col_idx), _f64(values)
            _check(lib().fem_bsr_create(int(bs), int(nbr), int(nbc), _p(rp, C.c_uint32), _p(ci, C.c_uint32),
                                        _p(v, C.c_double), C.byref(self._h)))
        bs_, nbr_, nbc_, nz = C.c_int(), C.c_uint32(), C.c_uint32(), C.c_uint64()
        _check(lib().fem_bsr_info(self._h, C.byref(bs_), C.byref(nbr_), C.byref(nbc_), C.byref(nz)))
        self.bs, self.nbr, self.nbc, self.nnzb = bs_.value, nbr_.value, nbc_.value, nz.value
        self.nrows = self.nbr * self.bs

    @classmethod
    def condense(cls, A, bs, nbr, nbc):
        """BlockMatrix::condense of a CSR matrix (on the device)."""
        h = C.c_void_p()
        _check(lib().fem_bsr_condense(A._h, int(bs), int(nbr), int(nbc), C.byref(h)))
        return cls(0, 0, 0, None, None, None, _handle=h)

    def matvec(self, x):
        x = _f64(x)
        y = np.empty(self.nrows, np.float64)
        _check(lib().fem_bsr_matvec(self._h, _p(x, C.c_double), _p(y, C.c_double)))
        return y

    def matvec_device(self, x_ptr: int, y_ptr: int, fp32=False):
        """y = A x on device buffers (synchronises the matrix's stream)."""
        _check(lib().fem_bsr_matvec_device(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), int(fp32)))

    def close(self):
        if self._h:
            lib().fem_bsr_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pc_params(kind: int, **kw) -> PcParams:
    p = PcParams()
    lib().fem_pc_params_default(kind, C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise InvalidArgument(f"unknown preconditioner parameter {k}")
        setattr(p, k, v)
    return p


class Preconditioner:
    def __init__(self, A, kind, **params):
        self.kind = PC_KINDS[kind] if isinstance(kind, str) else kind
        self._h = C.c_void_p()
        # the preconditioner runs on its matrix's stream: keep the matrix alive
        self._A = A
        self.n = A.rows if isinstance(A, CsrMatrix) else A.nrows
        p = pc_params(self.kind, **params)
        if isinstance(A, CsrMatrix):
            _check(lib().fem_precond_create(A._h, self.kind, C.byref(p), C.byref(self._h)))
        else:
            _check(lib().fem_bsr_precond_create(A._h, self.kind, C.byref(p), C.byref(self._h)))

    def apply(self, r):
        r = _f64(r)
        z = np.empty(self.n, np.float64)
        _check(lib().fem_precond_apply(self._h, _p(r, C.c_double), _p(z, C.c_double)))
        return z

    def info(self, cap=64):
        lv = C.c_int()
        rows = np.zeros(cap, np.uint32)
        nnz = np.zeros(cap, np.uint64)
        ms = C.c_double()
        _check(lib().fem_precond_info(self._h, C.byref(lv), _p(rows, C.c_uint32),
                                      nnz.ctypes.data_as(C.POINTER(C.c_uint64)), cap, C.byref(ms)))
        return {"levels": lv.value, "rows": rows, "nnz": nnz, "setup_ms": ms.value}

    def amg_matrix(self, level, which):
        """AMG level matrix (which: 0 operator, 1 prolongator, 2 restriction)
        as a scipy CSR matrix (fem_amg_level_export)."""
        import scipy.sparse as sp
        dims = np.zeros(3, np.uint32)
        _check(lib().fem_amg_level_export(self._h, level, which, _p(dims, C.c_uint32), None, None, None))
        n, m, nz = (int(x) for x in dims)
        rp = np.empty(n + 1, np.uint32)
        ci = np.empty(max(nz, 1), np.uint32)
        v = np.empty(max(nz, 1), np.float64)
        _check(lib().fem_amg_level_export(self._h, level, which, _p(dims, C.c_uint32), _p(rp, C.c_uint32),
                                          _p(ci, C.c_uint32), _p(v, C.c_double)))
        return sp.csr_matrix((v[:nz], ci[:nz], rp), shape=(n, m))

    def amg_split(self, level):
        """C/F split of a coarsened C/F AMG level: int8[n], 1 = C, 2 = F."""
        n = int(self.info()["rows"][level])
        st = np.empty(n, np.uint8)
        _check(lib().fem_amg_split(self._h, level, st.ctypes.data_as(C.POINTER(C.c_uint8))))
        return st.astype(np.int8)

    def bench_apply(self, fp32=False, reps=20) -> float:
        """Average device time (ms) of back-to-back applications
        (fem_bench_precond)."""
        ms = C.c_double()
        _check(lib().fem_bench_precond(self._h, int(fp32), reps, C.byref(ms)))
        return ms.value

    def apply_bytes(self, fp32=False):
        """Algorithmic bytes of one application (fem_precond_bytes); 0 where
        the kind is not tallied."""
        out = C.c_double()
        _check(lib().fem_precond_bytes(self._h, 1 if fp32 else 0, C.byref(out)))
        return out.value

    def close(self):
        if self._h:
            lib().fem_precond_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SchurPreconditioner(Preconditioner):
    """make_schur_preconditioner (fem_precond_create_schur) for the saddle-point
    matrix [[A, B^T], [B, 0]]: B (n_pres x n_vel) and Bt as CsrMatrix, A_pc and
    S_pc (None: identity) as Preconditioner; S_pc's output is scaled by
    `scale`. Composed on the device."""

    def __init__(self, B, Bt, A_pc, S_pc=None, scale=1.0):
        self.kind = PC_SCHUR
        self._h = C.c_void_p()
        self.n = B.rows + B.cols
        self._A = (B, Bt, A_pc, S_pc)
        _check(lib().fem_precond_create_schur(B._h, Bt._h, A_pc._h, S_pc._h if S_pc is not None else None, scale,
                                              C.byref(self._h)))


class LearnedPreconditioner(Preconditioner):
    """Learned factorized SPAI M = G^T G (fem_precond_create_learned): two
    sparse products per apply; a solve stops after `budget` iterations (0: no
    explicit limit) or after the fallback's recorded iteration count, and is
    then redone with the fallback preconditioner."""

    def __init__(self, A, grid, net, fallback="ic0", budget=0):
        self.kind = PC_LEARNED
        self._h = C.c_void_p()
        self.n = A.rows
        self._A = (A, net)
        fb = PC_KINDS[fallback] if isinstance(fallback, str) else fallback
        g = _u32(grid)
        _check(lib().fem_precond_create_learned(A._h, _p(g, C.c_uint32), net._h, fb, int(budget), C.byref(self._h)))

    def reference(self, iterations: int):
        """The fallback's iteration count on this matrix (fem_precond_learned_reference)."""
        _check(lib().fem_precond_learned_reference(self._h, int(iterations)))


class SpaiPreconditioner(Preconditioner):
    """A given sparse approximate inverse (fem_precond_create_spai): z = M r,
    or with factorized=True, Mat is G and z = G^T (G r)."""

    def __init__(self, A, Mat, factorized=False):
        self.kind = PC_SPAI
        self._h = C.c_void_p()
        self.n = A.rows
        self._A = (A, Mat)
        _check(lib().fem_precond_create_spai(A._h, Mat._h, 1 if factorized else 0, C.byref(self._h)))


def solve_opts(solver="pcg", tol=1e-10, max_iter=2000, restart=30, precision=FP64, inner_tol=None,
               max_refine=None, check_every=0, use_graph=True, device_ptrs=False) -> SolveOpts:
    o = SolveOpts()
    sk = SOLVERS[solver] if isinstance(solver, str) else solver
    lib().fem_solve_opts_default(sk, C.byref(o))
    o.tol, o.max_iter, o.restart, o.precision = tol, max_iter, restart, precision
    if inner_tol is not None:
        o.inner_tol = inner_tol
    if max_refine is not None:
        o.max_refine = max_refine
    o.check_every = check_every
    o.use_graph = int(use_graph)
    o.device_ptrs = int(device_ptrs)
    return o


def solve(A, M, b, **kw):
    """Solve A x = b; returns (x, stats dict)."""
    b = _f64(b)
    x = np.empty_like(b)
    o = solve_opts(**kw)
    st = SolveStats()
    mh = M._h if M is not None else None
    if isinstance(A, CsrMatrix):
        _check(lib().fem_solve(A._h, mh, _p(b, C.c_double), _p(x, C.c_double), C.byref(o), C.byref(st)))
    else:
        _check(lib().fem_bsr_solve(A._h, mh, _p(b, C.c_double), _p(x, C.c_double), C.byref(o), C.byref(st)))
    return x, st.as_dict()


class Solver:
    """Persistent solver (workspace and captured graphs reused)."""

    def __init__(self, A, M, **kw):
        self.A, self.M = A, M
        self.n = A.rows
        self._h = C.c_void_p()
        o = solve_opts(**kw)
        _check(lib().fem_solver_create(A._h, M._h if M is not None else None, C.byref(o), C.byref(self._h)))

    def solve(self, b):
        b = _f64(b)
        x = np.empty_like(b)
        st = SolveStats()
        _check(lib().fem_solver_solve(self._h, _p(b, C.c_double), _p(x, C.c_double), C.byref(st)))
        return x, st.as_dict()

    def solve_device(self, b_ptr: int, x_ptr: int):
        st = SolveStats()
        _check(lib().fem_solver_solve(self._h, C.cast(b_ptr, P_f64), C.cast(x_ptr, P_f64), C.byref(st)))
        return st.as_dict()

    def close(self):
        if self._h:
            lib().fem_solver_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gmres_callback(A, fn, b, restart=30, tol=1e-10, max_iter=2000):
    """GMRES with a host-side preconditioner fn(r) -> z (SMatrix::gmres with
    a std::function preconditioner)."""
    n = A.rows

    def cb(r, z, nn, user):
        rr = np.ctypeslib.as_array(r, shape=(nn,))
        zz = np.ctypeslib.as_array(z, shape=(nn,))
        zz[:] = fn(rr.copy())

    cfn = HOST_CB(cb)
    b = _f64(b)
    x = np.empty(n)
    it = C.c_int()
    _check(lib().fem_gmres_callback(A._h, cfn, None, _p(b, C.c_double), restart, tol, max_iter,
                                    _p(x, C.c_double), C.byref(it)))
    return x, it.value


def dot(a, b) -> float:
    a, b = _f64(a), _f64(b)
    if a.size != b.size:
        raise InvalidArgument("dot: vector lengths differ")
    out = C.c_double()
    _check(lib().fem_dot(a.size, _p(a, C.c_double), _p(b, C.c_double), C.byref(out)))
    return out.value


# ---------------------------------------------------------------------------
# Device buffers (HIP runtime through ctypes) so callers can keep inputs
# resident in HBM and use the device-pointer entry points.
# ---------------------------------------------------------------------------
_hip = None


def _hiprt():
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


def device_synchronize():
    """Wait for all work on the device (hipDeviceSynchronize)."""
    rc = _hiprt().hipDeviceSynchronize()
    if rc != 0:
        raise DeviceError(f"hipDeviceSynchronize failed ({rc})")


class DeviceArray:
    """float64 buffer in device memory."""

    def __init__(self, n: int):
        self.n = int(n)
        self.ptr = C.c_void_p()
        rc = _hiprt().hipMalloc(C.byref(self.ptr), max(self.n, 1) * 8)
        if rc != 0:
            raise DeviceError(f"hipMalloc failed ({rc})")

    @classmethod
    def from_host(cls, a):
        a = _f64(a)
        d = cls(a.size)
        d.upload(a)
        return d

    def upload(self, a):
        a = _f64(a)
        rc = _hiprt().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), a.size * 8, 1)
        if rc != 0:
            raise DeviceError(f"hipMemcpy H2D failed ({rc})")

    def download(self):
        out = np.empty(self.n)
        rc = _hiprt().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.n * 8, 2)
        if rc != 0:
            raise DeviceError(f"hipMemcpy D2H failed ({rc})")
        return out

    @property
    def address(self) -> int:
        return int(self.ptr.value or 0)

    def __del__(self):
        try:
            if self.ptr:
                _hiprt().hipFree(self.ptr)
        except Exception:
            pass


# ---- learned sparse approximate inverse --------------------------------------
MLP_NPARAM = 5063


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Mlp:
    """The learned SPAI network (6 -> 64 -> 64 -> 7) with its Adam state on
    the device (fem_mlp_*)."""

    def __init__(self, params=None, seed=0):
        self._h = C.c_void_p()
        if params is None:
            _check(lib().fem_mlp_create(None, seed, C.byref(self._h)))
        else:
            p = _f32(params)
            if p.size != MLP_NPARAM:
                raise InvalidArgument(f"expected {MLP_NPARAM} parameters")
            _check(lib().fem_mlp_create(_p(p, C.c_float), seed, C.byref(self._h)))

    @property
    def params(self):
        p = np.empty(MLP_NPARAM, np.float32)
        _check(lib().fem_mlp_get_params(self._h, _p(p, C.c_float)))
        return p

    @params.setter
    def params(self, value):
        p = _f32(value)
        _check(lib().fem_mlp_set_params(self._h, _p(p, C.c_float)))

    def forward(self, feat):
        f = _f32(feat)
        n = f.shape[0]
        out = np.empty((n, 7), np.float32)
        _check(lib().fem_mlp_forward(self._h, _p(f, C.c_float), n, _p(out, C.c_float)))
        return out

    def loss_grad(self, A, grid, V):
        V = _f64(np.atleast_2d(V))
        g = np.empty(MLP_NPARAM, np.float32)
        L = C.c_double()
        gr = _u32(grid)
        _check(lib().fem_spai_loss_grad(self._h, A._h, _p(gr, C.c_uint32), _p(V, C.c_double), V.shape[0], 0,
                                        C.byref(L), _p(g, C.c_float)))
        return L.value, g

    def train_step(self, A, grid, V, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, device_k=None):
        """One Adam step. V: (k, n) host array, or a device pointer with
        device_k = k."""
        o = AdamOpts(lr, beta1, beta2, eps)
        L = C.c_double()
        gr = _u32(grid)
        if device_k is None:
            V = _f64(np.atleast_2d(V))
            _check(lib().fem_spai_train_step(self._h, A._h, _p(gr, C.c_uint32), _p(V, C.c_double), V.shape[0], 0,
                                             C.byref(o), C.byref(L)))
        else:
            _check(lib().fem_spai_train_step(self._h, A._h, _p(gr, C.c_uint32), C.cast(V, P_f64), int(device_k), 1,
                                             C.byref(o), C.byref(L)))
        return L.value

    def close(self):
        if self._h:
            lib().fem_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spai_features(A, grid):
    f = np.empty((A.rows, 6), np.float32)
    g = _u32(grid)
    _check(lib().fem_spai_features(A._h, _p(g, C.c_uint32), _p(f, C.c_float)))
    return f


# ---- batched solves ------------------------------------------------------------
class Batch:
    """Independent SPD systems solved in one persistent launch (fem_batch_*).
    systems: gen.CSR (block size 1) or gen.BSR (block size 3) matrices."""

    def __init__(self, systems, nrhs=4):
        keep = []
        arr = (BatchSystem * len(systems))()
        self.sizes = []
        for k, A in enumerate(systems):
            if hasattr(A, "bs"):
                bs, nb, rp, ci, v = A.bs, A.nbr, A.row_ptr, A.col_idx, A.values
            else:
                bs, nb, rp, ci, v = 1, A.n, A.row_ptr, A.col_idx, A.values
            rp, ci, v = _u32(rp), _u32(ci), _f64(v)
            keep += [rp, ci, v]
            arr[k] = BatchSystem(bs, nb, _p(rp, C.c_uint32), _p(ci, C.c_uint32), _p(v, C.c_double))
            self.sizes.append(nb * bs)
        self.nrhs = nrhs
        self._h = C.c_void_p()
        _check(lib().fem_batch_create(len(systems), arr, nrhs, C.byref(self._h)))

    def set_rhs(self, Bs):
        """Bs: one (nrhs, n) array per system."""
        flat = _f64(np.concatenate([np.asarray(B, np.float64).reshape(self.nrhs, -1).ravel() for B in Bs]))
        _check(lib().fem_batch_set_rhs(self._h, _p(flat, C.c_double)))

    def solve(self, tol=1e-10, inner_tol=1e-4, max_refine=40, max_inner=20000):
        st = (BatchStats * len(self.sizes))()
        ms = C.c_double()
        o = BatchOpts(tol, inner_tol, max_refine, max_inner)
        _check(lib().fem_batch_solve(self._h, C.byref(o), st, C.byref(ms)))
        stats = [{k: getattr(x, k) for k, _ in BatchStats._fields_} for x in st]
        return stats, ms.value

    def solution(self):
        X = np.empty(sum(self.sizes) * self.nrhs)
        _check(lib().fem_batch_get_solution(self._h, _p(X, C.c_double)))
        out, o = [], 0
        for n in self.sizes:
            out.append(X[o:o + n * self.nrhs].reshape(self.nrhs, n))
            o += n * self.nrhs
        return out

    def tune(self, divisor=8, tol=1e-10, inner_tol=1e-4, max_refine=40, max_inner=20000):
        """Profile-guided schedule search (fem_batch_tune); returns the best
        schedule and the trial times (ms): 18 capped runs, then full runs of
        the fastest of each workgroup size, unsplit and split."""
        best = BatchSchedule()
        t = np.full(32, np.nan)
        o = BatchOpts(tol, inner_tol, max_refine, max_inner)
        _check(lib().fem_batch_tune(self._h, C.byref(o), divisor, C.byref(best), _p(t, C.c_double), 32))
        return ({"block": best.block, "wg_per_cu": best.wg_per_cu, "order": best.order, "split": best.split},
                t[~np.isnan(t)])

    def set_schedule(self, block=256, wg_per_cu=0, order=0, split=0):
        """split: the `split` systems of largest estimated work are queued as
        one work item per right-hand side."""
        s = BatchSchedule(block, wg_per_cu, order, split)
        _check(lib().fem_batch_set_schedule(self._h, C.byref(s)))

    def close(self):
        if self._h:
            lib().fem_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
