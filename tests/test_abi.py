"""The C-ABI library loads without a GPU and exports every symbol that
include/femgpu.h declares (no compute calls here)."""
import ctypes
import os

import femgpu


def test_library_built():
    assert os.path.exists(femgpu.LIB_PATH), "run __graft_entry__.build() first"


def test_exports_every_header_symbol():
    L = ctypes.CDLL(femgpu.LIB_PATH)
    names = femgpu.header_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    # every declared entry point has a ctypes signature in the binding
    assert set(femgpu.header_symbols()) <= set(femgpu._SIGS), set(femgpu.header_symbols()) - set(femgpu._SIGS)


def test_defaults_without_device():
    p = femgpu.pc_params(femgpu.PC_SSOR)
    assert abs(p.omega - 1.2) < 1e-15 and p.theta == 0.25
    o = femgpu.solve_opts("gmres")
    assert o.tol == 1e-10 and o.max_iter == 2000 and o.restart == 30
