"""ctypes binding of libfastgpr.so (the C ABI declared in include/fastgpr.h).

Replaces the reference's six ``ctypes.cdll.LoadLibrary`` handles and their
argtypes tables (src/__init__.py:30-221) with one library.  Unlike the
reference (src/__init__.py:14-18) nothing is compiled on import: the shared
object must already be built (``fast-gpr-fwi_amd/csrc/build.py``), and a
missing library is an ImportError -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

# torch first: it loads its bundled HIP/HSA runtime, and libfastgpr.so must
# bind to that same runtime instance (loading /opt/rocm's copy first would put
# two HSA runtimes in the process and the second one sees no device).
import torch  # noqa: F401

_LIB_PATH = Path(os.environ.get("FASTGPR_LIB", Path(__file__).with_name("libfastgpr.so")))

ABI_VERSION = 13
EXEC_EAGER = 0
EXEC_GRAPH = 1
GRAD_REFERENCE = 0
GRAD_EXACT = 1
(KERNEL_FWD_H, KERNEL_FWD_E, KERNEL_ADJ_INJECT, KERNEL_ADJ_E, KERNEL_ADJ_H,
 KERNEL_TB_FWD, KERNEL_TB_ADJ) = range(7)
PROFILE_GRAPH = 0x100      # or'd into a kernel id: replay the timed launches from a graph


class Problem(ctypes.Structure):
    """Mirror of ``fgpr_problem`` (include/fastgpr.h)."""

    _fields_ = [
        ("NI", ctypes.c_int32), ("NJ", ctypes.c_int32), ("P", ctypes.c_int32),
        ("S", ctypes.c_int32), ("nt", ctypes.c_int32), ("nsrc", ctypes.c_int32),
        ("nrx", ctypes.c_int32), ("pitch", ctypes.c_int32),
        ("dx", ctypes.c_float), ("dt", ctypes.c_float),
        ("coef", ctypes.c_void_p), ("pml", ctypes.c_void_p), ("src", ctypes.c_void_p),
        ("srcscale", ctypes.c_void_p), ("rx", ctypes.c_void_p),
        ("Ez", ctypes.c_void_p), ("Hx", ctypes.c_void_p), ("Hy", ctypes.c_void_p),
        ("psi", ctypes.c_void_p), ("traces", ctypes.c_void_p), ("snaps", ctypes.c_void_p),
        ("resid", ctypes.c_void_p), ("grad_eps", ctypes.c_void_p), ("grad_sig", ctypes.c_void_p),
        ("ctr", ctypes.c_void_p),
        ("Ez2", ctypes.c_void_p), ("Hx2", ctypes.c_void_p), ("Hy2", ctypes.c_void_p),
        ("psi2", ctypes.c_void_p),
        ("tb_ri", ctypes.c_int32), ("tb_nt", ctypes.c_int32), ("tb_k", ctypes.c_int32),
        ("tb_ti", ctypes.c_int32), ("tb_tj", ctypes.c_int32), ("tb_list_cap", ctypes.c_int32),
        ("tb_rx_list", ctypes.c_void_p), ("tb_inj_list", ctypes.c_void_p),
        ("gshot", ctypes.c_void_p),
        ("scratch", ctypes.c_void_p),
        ("snap_base", ctypes.c_int32), ("nsnap", ctypes.c_int32), ("snap_mod", ctypes.c_int32),
        ("snap_tiled", ctypes.c_int32),
    ]


class TBItem(ctypes.Structure):
    """Mirror of ``fgpr_tb_item``."""

    _fields_ = [("prob", ctypes.POINTER(Problem)), ("n0", ctypes.c_int32),
                ("nsteps", ctypes.c_int32), ("parity", ctypes.c_int32), ("pad", ctypes.c_int32)]


# name -> (restype, argtypes); every symbol include/fastgpr.h declares.
SIGNATURES = {
    "fgpr_abi_version": (ctypes.c_int, []),
    "fgpr_last_error": (ctypes.c_char_p, []),
    "fgpr_problem_size": (ctypes.c_size_t, []),
    "fgpr_pitch": (ctypes.c_int, [ctypes.c_int]),
    "fgpr_psi_floats": (ctypes.c_int64, [ctypes.c_int] * 4),
    "fgpr_validate": (ctypes.c_int, [ctypes.POINTER(Problem)]),
    "fgpr_tb_tiles": (ctypes.c_int, [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_int)] * 4),
    "fgpr_cell_coeffs": (ctypes.c_int, [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_int,
                                         ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p]),
    "fgpr_forward": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "fgpr_adjoint": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "fgpr_normalise": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double,
                                      ctypes.c_void_p, ctypes.c_void_p]),
    "fgpr_profile_kernel": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_void_p]),
    "fgpr_scale": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double,
                                  ctypes.c_void_p, ctypes.c_void_p]),
    "fgpr_tb_table_bytes": (ctypes.c_int64, [ctypes.c_int]),
    "fgpr_tb_pack": (ctypes.c_int, [ctypes.POINTER(TBItem), ctypes.c_int, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int)]),
    "fgpr_tb_launch_table": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_void_p]),
    "fgpr_tb_settle": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.c_int, ctypes.c_void_p]),
    "fgpr_reduce_shots": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.c_void_p]),
    "fgpr_release_graphs": (ctypes.c_int, []),
    "fgpr_graph_count": (ctypes.c_int, []),
    "fgpr_set_graph_limit": (ctypes.c_int, [ctypes.c_int]),
    "fgpr_tv_workspace_bytes": (ctypes.c_int64, [ctypes.c_int, ctypes.c_int]),
    "fgpr_row_sums": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_void_p]),
    "fgpr_snap_slot_floats": (ctypes.c_int64, [ctypes.POINTER(Problem)]),
    "fgpr_tb_snap_floats": (ctypes.c_int64, [ctypes.c_int] * 5),
    "fgpr_tb_map": (ctypes.c_int64, [ctypes.POINTER(Problem), ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "fgpr_fastdiv_mismatches": (ctypes.c_int64, [ctypes.c_int, ctypes.c_int64, ctypes.c_int64]),
    "fgpr_snap_pack": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_void_p]),
    "fgpr_fir": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "fgpr_tv_static": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_void_p,
                                      ctypes.c_void_p]),
    "fgpr_tv": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                               ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_void_p,
                               ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
}


def _load():
    if not _LIB_PATH.is_file():
        raise ImportError(
            f"libfastgpr.so not found at {_LIB_PATH}; build it with "
            "`python fast-gpr-fwi_amd/csrc/build.py` (hipcc, gfx950). "
            "There is no CPU fallback.")
    lib = ctypes.CDLL(str(_LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.fgpr_abi_version() != ABI_VERSION:
        raise ImportError(f"libfastgpr ABI {lib.fgpr_abi_version()} != expected {ABI_VERSION}")
    if lib.fgpr_problem_size() != ctypes.sizeof(Problem):
        raise ImportError("fgpr_problem layout mismatch between libfastgpr and the ctypes mirror")
    return lib


lib = _load()
LIB_PATH = _LIB_PATH


def check(rc: int, what: str = "") -> None:
    """Raise RuntimeError carrying fgpr_last_error() for a non-zero status."""
    if rc != 0:
        msg = lib.fgpr_last_error().decode(errors="replace")
        raise RuntimeError(f"fastgpr{(' ' + what) if what else ''} failed ({rc}): {msg}")


def tb_tiles(NI: int, NJ: int, ri: int, nt: int, k: int):
    """(TI, TJ, tiles_i, tiles_j) of a time-blocked variant."""
    out = [ctypes.c_int() for _ in range(4)]
    check(lib.fgpr_tb_tiles(NI, NJ, ri, nt, k, *[ctypes.byref(o) for o in out]), "tb_tiles")
    return tuple(o.value for o in out)


def pitch(NJ: int) -> int:
    return lib.fgpr_pitch(NJ)


def psi_floats(NI: int, NJ: int, P: int, S: int) -> int:
    return lib.fgpr_psi_floats(NI, NJ, P, S)


def tb_snap_floats(NI: int, NJ: int, tb) -> int:
    """Floats per shot of one snapshot slot in the time-blocked layout of
    variant ``tb`` = (region rows, threads, steps per launch)."""
    n = lib.fgpr_tb_snap_floats(NI, NJ, *tb)
    if n < 0:
        raise ValueError(f"time-blocking variant {tuple(tb)} is not compiled in")
    return int(n)
