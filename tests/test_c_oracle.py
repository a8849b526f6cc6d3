"""The C/OpenMP restatement (oracle/fdtd_oracle.c) against the numpy oracle
on identical inputs.  FP64: forward traces, snapshots and reference-mode
gradients agree to rounding (different summation grouping only).  FP32 mode:
the same operations in the same order in IEEE float, so traces, snapshots and
gradient sums are bitwise the numpy float32 mode's."""
import numpy as np
import pytest

from oracle import c_oracle as CO
from oracle import fdtd_oracle as O

pytestmark = pytest.mark.skipif(not (CO.available() and CO.available(np.float32)),
                                reason="oracle/_build/liboracle*.so not built")


def _problem(NI=40, NJ=30, P=6, S=3, nt=120, nrx=5, seed=3):
    rng = np.random.default_rng(seed)
    er = 4.0 + 6.0 * rng.random((NI, NJ))
    se = 1e-3 + 1e-2 * rng.random((NI, NJ))
    se[NI // 2, NJ // 2] = 500.0                          # a PEC cell
    mr = 1.0 + 0.2 * rng.random((NI, NJ))
    dx = 0.05
    dt = 0.9 * O.cfl_dt(dx)
    src = np.stack([[(rng.integers(P + 1, NI - P - 1), rng.integers(P + 1, NJ - P - 1)),
                     (rng.integers(1, NI - 1), rng.integers(1, NJ - 1))] for _ in range(S)])
    rx = rng.integers(1, NJ - 1, size=(S, nrx, 2))
    rx[..., 0] = rng.integers(1, NI - 1, size=(S, nrx))
    rx[:, -1] = rx[:, 0]                                   # colliding receivers
    w = O.ricker(nt, dt, 2e8, 1.0, (nt - 1.5) * dt)
    wav = np.stack([w, 0.3 * w])
    return O.Problem(er=er, se=se, mr=mr, dx=dx, dt=dt, nt=nt, P=P, src=src, rx=rx, wavelet=wav)


@pytest.mark.parametrize("P", [0, 6])
def test_c_forward_matches_numpy(P):
    pb = _problem(P=P)
    c, pml = O.setup(pb, np.float64)
    scale = O.source_scale(pb.wavelet, pb.dx, np.float64)
    tr_n, sn_n = O.forward(c, pml, pb.P, pb.dx, pb.src, scale, pb.rx, pb.nt, np.float64, True)
    tr_c, sn_c = CO.forward(c, pml, pb.P, pb.dx, pb.src, scale, pb.rx, pb.nt, True)
    assert np.abs(tr_n).max() > 0
    np.testing.assert_allclose(tr_c, tr_n, rtol=0, atol=1e-12 * np.abs(tr_n).max())
    np.testing.assert_allclose(sn_c, sn_n, rtol=0, atol=1e-12 * np.abs(sn_n).max())


def test_c_adjoint_matches_numpy():
    pb = _problem()
    c, pml = O.setup(pb, np.float64)
    scale = O.source_scale(pb.wavelet, pb.dx, np.float64)
    _, snaps = O.forward(c, pml, pb.P, pb.dx, pb.src, scale, pb.rx, pb.nt, np.float64, True)
    resid = np.random.default_rng(9).standard_normal((pb.src.shape[0], pb.nt, pb.rx.shape[1]))
    cadj, _ = O.setup(pb, np.float64, adjoint=True)
    ge_n, gs_n = O.adjoint_reference(cadj, pml, pb.P, pb.dx, pb.rx, resid, snaps)
    ge_c, gs_c = CO.adjoint_reference(cadj, pml, pb.P, pb.dx, pb.rx, resid, snaps)
    for a, b in ((ge_c, ge_n), (gs_c, gs_n)):
        assert np.abs(b).max() > 0
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-11 * np.abs(b).max())


@pytest.mark.parametrize("P", [0, 6])
def test_c_fp32_mode_is_bitwise_numpy_fp32(P):
    pb = _problem(P=P)
    f = np.float32
    c, pml = O.setup(pb, f)
    scale = O.source_scale(pb.wavelet, pb.dx, f)
    tr_n, sn_n = O.forward(c, pml, pb.P, pb.dx, pb.src, scale, pb.rx, pb.nt, f, True)
    tr_c, sn_c = CO.forward(c, pml, pb.P, pb.dx, pb.src, scale, pb.rx, pb.nt, True, dtype=f)
    assert tr_c.dtype == f and np.abs(tr_n).max() > 0
    assert np.array_equal(tr_c, tr_n) and np.array_equal(sn_c, sn_n)
    resid = np.random.default_rng(9).standard_normal((pb.src.shape[0], pb.nt, pb.rx.shape[1]))
    cadj, _ = O.setup(pb, f, adjoint=True)
    ge_n, gs_n = O.adjoint_reference(cadj, pml, pb.P, pb.dx, pb.rx, resid, sn_n, f)
    ge_c, gs_c = CO.adjoint_reference(cadj, pml, pb.P, pb.dx, pb.rx, resid, sn_n, dtype=f)
    assert np.abs(ge_n).max() > 0
    assert np.array_equal(ge_c, ge_n) and np.array_equal(gs_c, gs_n)


def test_c_oracle_reports_threads():
    assert CO.threads() >= 1


def test_sanitized_oracle_stays_in_bounds():
    """AddressSanitizer + UndefinedBehaviorSanitizer build of the C
    restatement (both precisions) over small grids with sources and
    receivers on corners and the padding row / column: no report.  A planted
    receiver past the fields is caught (the check has teeth)."""
    import shutil
    import subprocess
    from pathlib import Path
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("gcc / make not available")
    here = Path(__file__).resolve().parents[1] / "oracle"
    subprocess.run(["make", "-s", "-C", str(here), "sanitize"], check=True, timeout=300)
    for exe in ("sanitize_check", "sanitize_check_f32"):
        r = subprocess.run([str(here / "_build" / exe)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0 and "Sanitizer" not in r.stderr, r.stderr[-2000:]
    r = subprocess.run([str(here / "_build" / "sanitize_check"), "oob"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode != 0 and "AddressSanitizer: heap-buffer-overflow" in r.stderr
