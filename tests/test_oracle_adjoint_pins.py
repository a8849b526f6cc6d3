"""Pins for the oracle's reference-mode adjoint (``adjoint_reference``,
compute.py:1048-1180, back.cu:8-37) that do not restate its loop.

1. Lossless medium, no CPML.  There the forward and adjoint coefficient sets
   coincide (uE0 = 1 for both, uc.cu:34-41 and :80-86), and with uE0 = 1
   injecting before the E update equals injecting after it.  Write one
   forward step as H' = H - A C E, E' = E + B C^T H' (A = uH1, B = uE1 on the
   cells each update touches).  The transpose recursion of the exact adjoint
   (mu for E, nu for H), run backwards, becomes the same H-then-E stepping in
   the variables lambda = B mu, eta = -A nu.  So the reference's lambda after
   step n is -B mu^n for the residual divided by B at each receiver.  Its
   epsilon gradient sum_n (E^{n+1} - E^n) lambda^n then equals
   dt / (eps0 dx) times the exact epsilon gradient
   -(eps0/dt) sum_k uE4 (s[k] - s[k-1]) mu^{k+1} evaluated on snapshots
   shifted one step (s[k] = E^{k+2}; a wavelet that is zero at step 0 makes
   E^1 = 0).  Its sigma gradient sum_n E^{n+1} lambda^n is the same with s
   the running sum of the shifted snapshots.  ``adjoint_exact`` is itself
   pinned by dot-product and finite-difference tests, here also at P = 0 and
   sigma = 0.
2. Loss and CPML.  Injection before the E update of a cell equals injection
   of uE0 times the residual after it, because the E update and its CPML
   correction are affine in the cell's own Ez with slope uE0.  The backward
   loop is therefore the forward loop (H, E, source) run with the adjoint
   coefficients, a source at every receiver emitting uE0/uE4 times the
   time-reversed residual: lambda after step n equals that run's Ez after
   step nt-1-n.

Both fail when the oracle's phases are swapped or its residual index is off
by one (checked below by patching the oracle on the CPU)."""
import numpy as np
import pytest

from oracle import fdtd_oracle as O


def _lossless_case(seed=11, NI=26, NJ=22, nt=90, S=2, nrx=4):
    rng = np.random.default_rng(seed)
    er = rng.uniform(2.0, 9.0, (NI, NJ))
    mr = rng.uniform(1.0, 1.5, (NI, NJ))
    se = np.zeros((NI, NJ))
    dx = 0.05
    dt = 0.9 * O.cfl_dt(dx)
    src = np.stack([[(rng.integers(3, NI - 3), rng.integers(3, NJ - 3))] for _ in range(S)])
    rx = np.stack([rng.integers(1, NI, size=nrx), rng.integers(1, NJ, size=nrx)], -1)
    rx = np.repeat(rx[None], S, 0)
    rx[:, -1] = rx[:, 0]                                    # colliding receivers
    w = O.ricker(nt, dt, 4e8)[None, :].copy()
    w[:, 0] = 0.0                                           # E^1 = 0
    pb = O.Problem(er, se, mr, dx, dt, nt, 0, src, rx, w)
    _, snaps = O.run_forward(pb, np.float64, store=True)
    resid = rng.standard_normal((S, nt, nrx))
    return pb, snaps, resid


def _identity_sides(pb, snaps, resid):
    """(reference-mode (ge, gs), the same predicted from the exact adjoint)."""
    assert np.all(snaps[0] == 0)
    cadj, _ = O.setup(pb, np.float64, adjoint=True)
    ge_r, gs_r = O.adjoint_reference(cadj, {}, 0, pb.dx, pb.rx, resid, snaps)
    c, _ = O.setup(pb, np.float64)
    B = c.uE1[pb.rx[..., 0], pb.rx[..., 1]]                 # (S, nrx)
    resid_b = resid / B[:, None, :]
    shifted = np.zeros_like(snaps)
    shifted[:-1] = snaps[1:]
    scale = pb.dt / (O.EPS0_KERNEL * pb.dx)
    ge_x, _ = O.adjoint_exact(c, {}, 0, pb.dx, pb.dt, pb.rx, resid_b, shifted)
    gs_x, _ = O.adjoint_exact(c, {}, 0, pb.dx, pb.dt, pb.rx, resid_b, np.cumsum(shifted, 0))
    return (ge_r, gs_r), (scale * ge_x, scale * gs_x)


def test_reference_adjoint_equals_shifted_exact_adjoint_lossless():
    pb, snaps, resid = _lossless_case()
    (ge_r, gs_r), (ge_p, gs_p) = _identity_sides(pb, snaps, resid)
    for a, b in ((ge_r, ge_p), (gs_r, gs_p)):
        assert np.abs(b).max() > 0
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-10 * np.abs(b).max())


def test_exact_adjoint_dot_product_lossless_no_cpml():
    """``adjoint_exact`` in the regime the identity uses (P = 0, sigma = 0)."""
    pb, snaps, resid = _lossless_case(seed=5, nt=70)
    c, _ = O.setup(pb, np.float64)
    ge, _ = O.adjoint_exact(c, {}, 0, pb.dx, pb.dt, pb.rx, resid, snaps)
    rng = np.random.default_rng(6)
    d = rng.standard_normal(pb.er.shape)
    h = 1e-4 * pb.er.mean()

    def tr(er):
        q = O.Problem(er, pb.se, pb.mr, pb.dx, pb.dt, pb.nt, 0, pb.src, pb.rx, pb.wavelet)
        return O.run_forward(q, np.float64, store=False)[0]
    fd = np.sum((tr(pb.er + h * d) - tr(pb.er - h * d)) / (2 * h) * resid)
    ad = np.sum(d * ge)
    assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)


def _lossy_cpml_case(seed=21, NI=30, NJ=24, P=5, nt=70, S=2, nrx=3):
    rng = np.random.default_rng(seed)
    er = rng.uniform(3.0, 10.0, (NI, NJ))
    se = rng.uniform(2e-3, 2e-2, (NI, NJ))
    mr = rng.uniform(1.0, 1.3, (NI, NJ))
    dx = 0.05
    dt = 0.9 * O.cfl_dt(dx)
    src = np.stack([[(rng.integers(P + 2, NI - P - 2), rng.integers(P + 2, NJ - P - 2))]
                    for _ in range(S)])
    # receivers anywhere on the model grid, CPML slabs included
    rx = np.stack([rng.integers(1, NI, size=(S, nrx)), rng.integers(1, NJ, size=(S, nrx))], -1)
    rx[:, -1] = rx[:, 0]
    w = O.ricker(nt, dt, 4e8)[None, :]
    pb = O.Problem(er, se, mr, dx, dt, nt, P, src, rx, w)
    _, snaps = O.run_forward(pb, np.float64, store=True)
    resid = rng.standard_normal((S, nt, rx.shape[1]))
    return pb, snaps, resid


def _lambda_sides(pb, snaps, resid):
    """lambda after each backward step, and the time-reversed forward runs."""
    cadj, pml = O.setup(pb, np.float64, adjoint=True)
    S, nt, nrx = resid.shape
    NI, NJ = pb.er.shape
    lam = np.zeros((nt, S, NI, NJ))

    def hook(n, st):
        lam[n] = st.Ez[:, :NI, :NJ]
    O.adjoint_reference(cadj, pml, pb.P, pb.dx, pb.rx, resid, snaps, hook=hook)
    fwd = np.zeros_like(lam)
    for s in range(S):
        ri, rj = pb.rx[s, :, 0], pb.rx[s, :, 1]
        ratio = cadj.uE0[ri, rj] / cadj.uE4[ri, rj]
        amp = np.zeros((nrx, nt))
        amp[:, :nt - 1] = ratio[:, None] * resid[s, ::-1][:nt - 1].T   # step m <- n = nt-1-m
        _, sn = O.forward(cadj, pml, pb.P, pb.dx, pb.rx[s:s + 1], amp, pb.rx[s:s + 1], nt,
                          np.float64, True)
        fwd[1:, s] = sn[nt - 2::-1, 0]                       # lambda(n) = run step nt-1-n
    return lam[1:], fwd[1:]


def test_reference_adjoint_is_time_reversed_forward_with_loss_and_cpml():
    pb, snaps, resid = _lossy_cpml_case()
    lam, fwd = _lambda_sides(pb, snaps, resid)
    assert np.abs(fwd).max() > 0
    np.testing.assert_allclose(lam, fwd, rtol=0, atol=1e-11 * np.abs(fwd).max())


# --- the pins reject plausible mistakes, planted in the oracle ----------------
def _planted(mistake):
    """``adjoint_reference`` with one plausible mistake; the forward runs the
    pins compare against stay untouched."""
    real = O.adjoint_reference

    def bad(c, pml, P, dx, rx, r, sn, *a, **k):
        if mistake == "residual_index":                      # resid[n-1] injected at n
            return real(c, pml, P, dx, rx, np.roll(r, 1, axis=1), sn, *a, **k)
        e, h = O._e_phase, O._h_phase                        # H before E
        O._e_phase, O._h_phase = h, e
        try:
            return real(c, pml, P, dx, rx, r, sn, *a, **k)
        finally:
            O._e_phase, O._h_phase = e, h
    return bad


@pytest.mark.parametrize("mistake", ["swapped_phases", "residual_index"])
def test_pins_fail_on_planted_mistakes(monkeypatch, mistake):
    pb, snaps, resid = _lossless_case(nt=60)
    pb2, snaps2, resid2 = _lossy_cpml_case(nt=50)
    monkeypatch.setattr(O, "adjoint_reference", _planted(mistake))
    (ge_r, _), (ge_p, _) = _identity_sides(pb, snaps, resid)
    assert np.abs(ge_r - ge_p).max() > 1e-3 * np.abs(ge_p).max()
    lam, fwd = _lambda_sides(pb2, snaps2, resid2)
    assert np.abs(lam - fwd).max() > 1e-3 * np.abs(fwd).max()
