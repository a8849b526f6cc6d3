"""ctypes wrapper of the C restatement (oracle/fdtd_oracle.c), built as
oracle/_build/liboracle.so (FP64) and oracle/_build/liboracle_f32.so (FP32
mode).

TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).  Same
arguments and results as ``fdtd_oracle.forward`` and
``fdtd_oracle.adjoint_reference``; ``dtype=np.float32`` runs the FP32 mode,
bitwise the numpy oracle's float32 mode.  Coefficient set-up stays in the
numpy oracle.  Build with ``make -C oracle``.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from . import fdtd_oracle as O

LIB_PATHS = {np.dtype(np.float64): Path(__file__).resolve().parent / "_build" / "liboracle.so",
             np.dtype(np.float32): Path(__file__).resolve().parent / "_build" / "liboracle_f32.so"}
LIB_PATH = LIB_PATHS[np.dtype(np.float64)]
_libs: dict = {}

_I = ctypes.POINTER(ctypes.c_int64)
_G = ctypes.POINTER(ctypes.c_double)


def lib(dtype=np.float64):
    dt = np.dtype(dtype)
    if dt not in _libs:
        path = LIB_PATHS[dt]
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built (make -C oracle)")
        L = ctypes.CDLL(str(path))
        R = ctypes.POINTER(ctypes.c_double if dt == np.float64 else ctypes.c_float)
        common = [ctypes.c_int] * 4 + [ctypes.c_double] + [R] * 7
        L.oracle_forward.argtypes = common + [ctypes.c_int, _I, R, ctypes.c_int, _I,
                                              ctypes.c_int, R, R]
        L.oracle_adjoint.argtypes = common + [ctypes.c_int, _I, ctypes.c_int, R, R, _G, _G]
        L.oracle_forward.restype = ctypes.c_int
        L.oracle_adjoint.restype = ctypes.c_int
        L.oracle_threads.restype = ctypes.c_int
        _libs[dt] = L
    return _libs[dt]


def available(dtype=np.float64) -> bool:
    return LIB_PATHS[np.dtype(dtype)].exists()


def threads() -> int:
    return int(lib().oracle_threads())


def _a(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a):
    ct = ctypes.c_double if a.dtype == np.float64 else ctypes.c_float
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _coeff_args(c: O.Coeffs, pml: dict, P: int, dtype):
    planes = [_a(getattr(c, k), dtype) for k in ("uE0", "uE1", "uE4", "uH0", "uH1", "uH4")]
    if P:
        table = _a(np.concatenate([np.asarray(pml[k], dtype) for k in
                                   ("ERA", "ERB", "ERE", "ERF", "HRA", "HRB", "HRE", "HRF")]),
                   dtype)
    else:
        table = np.zeros(1, dtype)
    return planes, table


def forward(c: O.Coeffs, pml: dict, P: int, dx: float, src, scale, rx, nt: int, store=True,
            dtype=np.float64):
    """Forward (oracle/fdtd_oracle.py ``forward``) in ``dtype``: traces
    (S, nt, nrx) and snapshots (nt, S, NI, NJ) or None."""
    planes, table = _coeff_args(c, pml, P, dtype)
    src = np.ascontiguousarray(src, dtype=np.int64)
    rx = np.ascontiguousarray(rx, dtype=np.int64)
    S, nsrc, nrx = src.shape[0], src.shape[1], rx.shape[1]
    NI, NJ = c.uE0.shape[0] - 1, c.uE0.shape[1] - 1
    sc = _a(scale, dtype)
    traces = np.zeros((S, nt, nrx), dtype)
    snaps = np.zeros((nt, S, NI, NJ), dtype) if store else None
    rc = lib(dtype).oracle_forward(S, NI, NJ, P, float(dx), *[_p(a) for a in planes], _p(table),
                                   nsrc, src.ctypes.data_as(_I), _p(sc), nrx,
                                   rx.ctypes.data_as(_I), nt, _p(traces),
                                   _p(snaps) if store else None)
    if rc:
        raise MemoryError("oracle_forward could not allocate its state")
    return traces, snaps


def adjoint_reference(cadj: O.Coeffs, pml: dict, P: int, dx: float, rx, resid, snaps,
                      dtype=np.float64):
    """Reference-mode adjoint (``fdtd_oracle.adjoint_reference``) in
    ``dtype`` with float64 sums: unnormalised (g_eps, g_sig)."""
    planes, table = _coeff_args(cadj, pml, P, dtype)
    rx = np.ascontiguousarray(rx, dtype=np.int64)
    res = _a(resid, dtype)
    sn = _a(snaps, dtype)
    S, nt, nrx = res.shape
    NI, NJ = cadj.uE0.shape[0] - 1, cadj.uE0.shape[1] - 1
    ge = np.zeros((NI, NJ))
    gs = np.zeros((NI, NJ))
    rc = lib(dtype).oracle_adjoint(S, NI, NJ, P, float(dx), *[_p(a) for a in planes], _p(table),
                                   nrx, rx.ctypes.data_as(_I), nt, _p(res), _p(sn),
                                   ge.ctypes.data_as(_G), gs.ctypes.data_as(_G))
    if rc:
        raise MemoryError("oracle_adjoint could not allocate its state")
    return ge, gs
