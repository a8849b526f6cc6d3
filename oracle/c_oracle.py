"""ctypes wrapper of oracle/_build/liboracle.so (oracle/fdtd_oracle.c).

TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).  Same
arguments and results as ``fdtd_oracle.forward`` and
``fdtd_oracle.adjoint_reference`` in float64; coefficient set-up stays in
the numpy oracle.  Build with ``make -C oracle``.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from . import fdtd_oracle as O

LIB_PATH = Path(__file__).resolve().parent / "_build" / "liboracle.so"
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int64)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is not built (make -C oracle)")
        _lib = ctypes.CDLL(str(LIB_PATH))
        common = [ctypes.c_int] * 4 + [ctypes.c_double] + [_D] * 7
        _lib.oracle_forward.argtypes = common + [ctypes.c_int, _I, _D, ctypes.c_int, _I,
                                                 ctypes.c_int, _D, _D]
        _lib.oracle_adjoint.argtypes = common + [ctypes.c_int, _I, ctypes.c_int, _D, _D, _D, _D]
        _lib.oracle_forward.restype = ctypes.c_int
        _lib.oracle_adjoint.restype = ctypes.c_int
        _lib.oracle_threads.restype = ctypes.c_int
    return _lib


def available() -> bool:
    return LIB_PATH.exists()


def threads() -> int:
    return int(lib().oracle_threads())


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return a.ctypes.data_as(_D)


def _coeff_args(c: O.Coeffs, pml: dict, P: int):
    planes = [_d(getattr(c, k)) for k in ("uE0", "uE1", "uE4", "uH0", "uH1", "uH4")]
    if P:
        table = _d(np.concatenate([np.asarray(pml[k], np.float64) for k in
                                   ("ERA", "ERB", "ERE", "ERF", "HRA", "HRB", "HRE", "HRF")]))
    else:
        table = np.zeros(1)
    return planes, table


def forward(c: O.Coeffs, pml: dict, P: int, dx: float, src, scale, rx, nt: int, store=True):
    """FP64 forward (oracle/fdtd_oracle.py ``forward``): traces (S, nt, nrx)
    and snapshots (nt, S, NI, NJ) or None."""
    planes, table = _coeff_args(c, pml, P)
    src = np.ascontiguousarray(src, dtype=np.int64)
    rx = np.ascontiguousarray(rx, dtype=np.int64)
    S, nsrc, nrx = src.shape[0], src.shape[1], rx.shape[1]
    NI, NJ = c.uE0.shape[0] - 1, c.uE0.shape[1] - 1
    sc = _d(scale)
    traces = np.zeros((S, nt, nrx))
    snaps = np.zeros((nt, S, NI, NJ)) if store else None
    rc = lib().oracle_forward(S, NI, NJ, P, float(dx), *[_p(a) for a in planes], _p(table),
                              nsrc, src.ctypes.data_as(_I), _p(sc), nrx, rx.ctypes.data_as(_I),
                              nt, _p(traces), _p(snaps) if store else None)
    if rc:
        raise MemoryError("oracle_forward could not allocate its state")
    return traces, snaps


def adjoint_reference(cadj: O.Coeffs, pml: dict, P: int, dx: float, rx, resid, snaps):
    """FP64 reference-mode adjoint (``fdtd_oracle.adjoint_reference``):
    unnormalised (g_eps, g_sig)."""
    planes, table = _coeff_args(cadj, pml, P)
    rx = np.ascontiguousarray(rx, dtype=np.int64)
    res = _d(resid)
    sn = _d(snaps)
    S, nt, nrx = res.shape
    NI, NJ = cadj.uE0.shape[0] - 1, cadj.uE0.shape[1] - 1
    ge = np.zeros((NI, NJ))
    gs = np.zeros((NI, NJ))
    rc = lib().oracle_adjoint(S, NI, NJ, P, float(dx), *[_p(a) for a in planes], _p(table),
                              nrx, rx.ctypes.data_as(_I), nt, _p(res), _p(sn), _p(ge), _p(gs))
    if rc:
        raise MemoryError("oracle_adjoint could not allocate its state")
    return ge, gs
