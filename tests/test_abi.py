"""The C-ABI library loads and exports every symbol include/fastgpr.h
declares; host-side argument validation (no GPU work is launched)."""
import ctypes
import re
from pathlib import Path

import pytest

HEADER = Path(__file__).resolve().parents[1] / "include" / "fastgpr.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fgpr_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "fgpr_forward" in names and "fgpr_adjoint" in names and len(names) >= 9


def test_library_exports_every_declared_symbol():
    from fastgpr import _native
    for name in declared_functions():
        assert hasattr(_native.lib, name), name
        assert name in _native.SIGNATURES, f"{name} has no ctypes signature"


def test_abi_version_and_layout_helpers():
    from fastgpr import _native
    assert _native.lib.fgpr_abi_version() == _native.ABI_VERSION
    for nj in (1, 2, 63, 64, 100, 200, 1000):
        p = _native.pitch(nj)
        assert p >= nj + 1 and p % 64 == 0
    assert _native.psi_floats(200, 100, 20, 3) == 4 * 3 * 20 * 100 + 4 * 3 * 200 * 20


def _problem(**kw):
    from fastgpr import _native
    p = _native.Problem()
    p.NI, p.NJ, p.P, p.S, p.nt, p.nsrc, p.nrx = 50, 40, 5, 2, 10, 1, 3
    p.pitch = _native.pitch(40)
    p.dx, p.dt = 0.05, 1e-10
    fake = 0x1000   # never dereferenced: fgpr_validate only inspects the descriptor
    p.coef = p.pml = p.Ez = p.Hx = p.Hy = p.psi = fake
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("field,value,msg", [
    ("pitch", 41, "pitch"),
    ("P", 25, "overlap"),
    ("S", 0, "bad S"),
    ("coef", None, "null"),
    ("dx", 0.0, "dx"),
])
def test_validate_rejects_bad_descriptors(field, value, msg):
    from fastgpr import _native
    p = _problem(**{field: value})
    rc = _native.lib.fgpr_validate(ctypes.byref(p))
    assert rc == -1
    assert msg in _native.lib.fgpr_last_error().decode()


def test_validate_accepts_consistent_descriptor():
    from fastgpr import _native
    assert _native.lib.fgpr_validate(ctypes.byref(_problem())) == 0


def test_check_raises_with_message():
    from fastgpr import _native
    p = _problem(P=30)
    with pytest.raises(RuntimeError, match="overlap"):
        _native.check(_native.lib.fgpr_validate(ctypes.byref(p)), "validate")
