"""numpy restatement of the reference's 2-D TMz FDTD forward / adjoint path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product
package never imports it.

Two arithmetic modes share one code path:

* ``dtype=np.float64`` -- the FP64 transcription of the update equations
  (the accuracy reference for FP32 drift and for gradients).
* ``dtype=np.float32`` -- the same equations evaluated in IEEE float32 in the
  reference's operation order (no fused multiply-adds; numpy never fuses),
  including the one ``double`` promotion in the coefficient kernel.  This is
  the stand-in for "the reference FP32 result": the CUDA reference cannot run
  in this container.

Arrays follow the reference's padded layout with the dead ``k`` plane
dropped: every field is ``(S, NI+1, NJ+1)`` where ``NI = G.nx`` (depth, first
axis) and ``NJ = G.ny`` (lateral, second axis); compute.py:499-518.

Reference citations (paths relative to the reference repository):

* constants ........ src/__init__.py:3-4,21; compute.py:5-6; src/lib/uc.cu:8-9
* dt / nt .......... compute.py:752-769, 803
* Ricker ........... compute.py:334-357, 383-400
* CPML profile ..... compute.py:80-123 (CFS), 166-222 (HORIPML, forward branch),
                     267-313 (slab geometry, sigma_max averaging over the x0 row)
* cell coefficients  src/lib/uc.cu:11-45 (forward), 60-91 (adjoint)
* forward step ..... compute.py:574-712; fields_updates_gpu.cu:8-78;
                     pml_updates_h.cu:13-427; pml_updates_e.cu:12-402;
                     sourcereceiver.cu:11-87
* adjoint step ..... compute.py:1048-1180; back.cu:8-37

Pinning of each function (what checks it, in ``tests/test_oracle.py``):

* ``cfl_dt``, ``n_iterations``, ``cell_coeffs``, ``cpml_profile``, ``ricker``:
  closed forms and hand-computed known answers (P=2 profile vectors, Ricker
  peak/symmetry, coefficient identities, PEC switch).
* ``forward``: lossless-box energy bound, mirror symmetry, source/receiver
  reciprocity, CPML reflection against a 4x larger domain, FP32-vs-FP64 drift,
  linearity in the source, and agreement with the independent C restatement.
* ``adjoint_reference`` (``tests/test_oracle_adjoint_pins.py``): in a
  lossless medium without CPML the reference's backward loop is the exact
  transpose of the forward with the adjoint fields scaled by uE1, so both of
  its gradients equal dt/(eps0 dx) times ``adjoint_exact``'s epsilon gradient
  for a residual divided by uE1 at the receivers and snapshots shifted one
  step (their running sum for sigma); with loss and CPML its lambda fields
  equal a forward run with the adjoint coefficients driven at the receivers
  by the time-reversed residual times uE0/uE4.  Swapped phases or an
  off-by-one residual index fail both.  Also linearity, unit norm, zero
  rows/columns and agreement with the C restatement (``tests/test_oracle.py``).
* ``adjoint_exact``: the dot-product test <J dm, r> = <dm, J^T r> and central
  finite differences of the forward map.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# --- constants -----------------------------------------------------------
# Host-side (Python) constants come from scipy.constants (CODATA 2022 in the
# scipy the reference pins); the coefficient kernel hard-codes CODATA 2018
# float literals (uc.cu:8-9).  Both are kept exactly where each is used.
EPS0 = 8.8541878188e-12          # scipy.constants.epsilon_0, compute.py:6
MU0 = 1.25663706127e-06          # scipy.constants.mu_0, __init__.py:4
C0 = 299792458.0                 # scipy.constants.c, compute.py:5
Z0 = math.sqrt(MU0 / EPS0)       # __init__.py:21
EPS0_KERNEL = 8.8541878128e-12   # uc.cu:8
MU0_KERNEL = 1.25663706212e-06   # uc.cu:9


def cfl_dt(dx: float) -> float:
    """2-D Courant limit with dy = dx (compute.py:754)."""
    return 1.0 / (C0 * ((1.0 / dx) * (1.0 / dx) + (1.0 / dx) * (1.0 / dx)) ** 0.5)


def resolve_dt(dx: float, dt: float | None) -> float:
    """dt rule of compute.py:752-765: CFL when absent, clamped to CFL."""
    lim = cfl_dt(dx)
    if dt is None or dt > lim:
        return lim
    return float(dt)


def n_iterations(time_window: float, dt: float) -> int:
    """nt = ceil(tw / dt) + 1 with the division done in float32 (compute.py:803)."""
    q = np.float32(time_window) / np.float32(dt)
    return int(np.ceil(q)) + 1


# --- source wavelet --------------------------------------------------------
def ricker(nt: int, dt: float, freq: float, amp: float = 1.0,
           t_stop: float | None = None, dtype=np.float64) -> np.ndarray:
    """Ricker samples w[n] at t = n*dt (compute.py:351-357, 383-400).

    w(t) = -(2 z (2 z tau^2 - 1) exp(-z tau^2)) / (2 z) * amp with
    tau = t - sqrt(2)/f and z = pi^2 f^2.  Samples outside [0, t_stop] are 0.
    """
    f = dtype
    out = np.zeros(nt, dtype=f)
    fr = f(freq)
    chi = f(np.sqrt(f(2.0))) / fr
    zeta = f(np.pi ** 2) * (fr * fr)
    two_z = f(2.0) * zeta
    normalise = f(1.0) / two_z
    stop = np.float32(t_stop) if t_stop is not None else None
    for n in range(nt):
        t = dt * n
        if t < 0 or (stop is not None and np.float32(t) > stop):
            continue
        delay = f(t) - chi
        d2 = delay * delay
        v = -(two_z * (two_z * d2 - f(1.0)) * f(np.exp(-zeta * d2))) * normalise
        out[n] = v * f(amp)
    return out


def source_scale(w: np.ndarray, dx: float, dtype=np.float64) -> np.ndarray:
    """Per-step injection amount w*dl/(dx*dy*dz) with dl = dy = dz = dx
    (sourcereceiver.cu:73-75, compute.py:571)."""
    f = dtype
    d = f(dx)
    return (w.astype(f) * d) / ((d * d) * d)


# --- CPML ------------------------------------------------------------------
def sigma_max(dx: float, er_mean: float, mr_mean: float) -> float:
    """sigma_max = 0.8 (m+1) / (z0 d sqrt(er mr)), m = 4 (compute.py:80-84)."""
    return (0.8 * 5) / (Z0 * dx * math.sqrt(er_mean * mr_mean))


def cpml_profile(P: int, dx: float, dt: float, er_row0: np.ndarray,
                 mr_row0: np.ndarray, dtype=np.float64) -> dict:
    """HORIPML recursive-convolution coefficients for one slab, length P.

    sigma_E[q] = smax (q/P)^4, sigma_H[q] = smax ((q+1/2)/P)^4, kappa = 1,
    alpha = 0 (compute.py:86-123 with the profiles set at :73-77).
    R_A = R_B = 2 e0 / (2 e0 + dt s), R_E = (2 e0 - dt s)/(2 e0 + dt s),
    R_F = 2 s dt / (2 e0 + dt s) (compute.py:211-222).  smax uses the mean
    of er/mr along the x0 slab's first row for every slab, because the four
    slabs share one CFS object that caches sigma_max (compute.py:178, 278-283).
    """
    f = dtype
    er_mean = float(np.sum(np.asarray(er_row0, dtype=np.float64))) / len(er_row0)
    mr_mean = float(np.sum(np.asarray(mr_row0, dtype=np.float64))) / len(mr_row0)
    smax = f(sigma_max(float(np.float32(dx)) if dtype == np.float32 else dx,
                       er_mean, mr_mean))
    q = np.arange(P, dtype=np.float64)
    sE = f(smax) * ((q / P) ** 4).astype(f)
    sH = f(smax) * (((q + 0.5) / P) ** 4).astype(f)
    two_e0 = f(2 * EPS0)
    fdt = f(dt)
    out = {}
    for tag, s in (("E", sE), ("H", sH)):
        den = two_e0 + fdt * s
        ra = (two_e0 + fdt * f(0.0)) / den
        out[tag + "RA"] = ra.astype(f)
        out[tag + "RB"] = (two_e0 / den).astype(f)
        out[tag + "RE"] = ((two_e0 - fdt * s) / den).astype(f)
        out[tag + "RF"] = ((f(2.0) * s * fdt) / den).astype(f)
    out["smax"] = float(smax)
    return out


# --- per-cell coefficients --------------------------------------------------
@dataclass
class Coeffs:
    """Padded (NI+1, NJ+1) coefficient planes; zero at i = NI or j = NJ."""
    uE0: np.ndarray
    uE1: np.ndarray
    uE4: np.ndarray
    uH0: np.ndarray
    uH1: np.ndarray
    uH4: np.ndarray


def cell_coeffs(er, se, mr, dt: float, dx: float, adjoint: bool = False,
                dtype=np.float64) -> Coeffs:
    """Update coefficients of uc.cu:11-45 (forward) / 60-91 (adjoint).

    HA = mu0 mr / dt, uH0 = 1, uH1 = 1/(dx HA), uH4 = 1/HA.
    sigma > 100 (PEC): uE0 = uE1 = uE4 = 0.  Otherwise
    EA = e0 er/dt + se/2, EB = e0 er/dt - se/2, uE0 = EB/EA (forward) or
    2 e0 er / (2 e0 er + se dt) (adjoint), uE1 = 1/(dx EA), uE4 = 1/EA.
    In float32 mode ``0.5*se`` is a double product added in double and
    rounded once, as the C source ``(e0*er/dt) + 0.5*se`` does.
    """
    er = np.asarray(er, dtype=np.float64)
    se = np.asarray(se, dtype=np.float64)
    mr = np.asarray(mr, dtype=np.float64)
    NI, NJ = er.shape
    f = dtype
    e0 = f(EPS0_KERNEL)
    m0 = f(MU0_KERNEL)
    fdt = f(dt)
    fdx = f(dx)
    erf, sef, mrf = er.astype(f), se.astype(f), mr.astype(f)
    HA = (m0 * mrf) / fdt
    inv_dx = f(1.0) / fdx
    uH1 = (inv_dx * f(1.0)) / HA
    uH4 = f(1.0) / HA
    t = (e0 * erf) / fdt
    if dtype == np.float32:
        half_se = 0.5 * sef.astype(np.float64)
        EA = (t.astype(np.float64) + half_se).astype(np.float32)
        EB = (t.astype(np.float64) - half_se).astype(np.float32)
    else:
        EA = t + 0.5 * sef
        EB = t - 0.5 * sef
    if adjoint:
        two_e0_er = (f(2.0) * e0) * erf
        uE0 = two_e0_er / (two_e0_er + sef * fdt)
    else:
        uE0 = EB / EA
    uE1 = (inv_dx * f(1.0)) / EA
    uE4 = f(1.0) / EA
    pec = sef > 100
    for a in (uE0, uE1, uE4):
        a[pec] = 0

    def pad(a):
        p = np.zeros((NI + 1, NJ + 1), dtype=f)
        p[:NI, :NJ] = a
        return p

    return Coeffs(pad(uE0), pad(uE1), pad(uE4), pad(np.ones_like(uH1)), pad(uH1), pad(uH4))


# --- time stepping ---------------------------------------------------------
def _slab_rows(kind: str, P: int, n: int) -> np.ndarray:
    """Field index touched by ψ entry q in a slab (pml_updates_h.cu:39,155,276,391;
    pml_updates_e.cu:41,132 and the y analogues).  H slabs sit one cell lower
    than E slabs on the minus side."""
    q = np.arange(P)
    if kind == "H-":
        return P - 1 - q
    if kind == "E-":
        return P - q
    return n - P + q          # plus side, both E and H


class _State:
    def __init__(self, S, NI, NJ, P, f):
        shp = (S, NI + 1, NJ + 1)
        self.Ez = np.zeros(shp, f)
        self.Hx = np.zeros(shp, f)
        self.Hy = np.zeros(shp, f)
        # ψ per slab: x slabs (S, P, NJ) indexed [s, q, j];
        # y slabs (S, NI, P) indexed [s, i, q].
        self.psi = {}
        for fld in ("E", "H"):
            for side in ("x0", "xmax"):
                self.psi[fld + side] = np.zeros((S, P, NJ), f)
            for side in ("y0", "ymax"):
                self.psi[fld + side] = np.zeros((S, NI, P), f)


def _h_phase(st: _State, c: Coeffs, pml: dict, P: int, dx, f):
    Ez, Hx, Hy = st.Ez, st.Hx, st.Hy
    S, NI1, NJ1 = Ez.shape
    NI, NJ = NI1 - 1, NJ1 - 1
    one = f(1.0)
    fdx = f(dx)
    # main update, fields_updates_gpu.cu:59-77.  Neighbour reads past the
    # padded edge are multiplied by a zero coefficient there; read as 0.
    EzJ = np.zeros_like(Ez)
    EzJ[:, :, :NJ] = Ez[:, :, 1:]
    EzI = np.zeros_like(Ez)
    EzI[:, :NI, :] = Ez[:, 1:, :]
    Hx[:, 1:, :] = c.uH0[1:, :] * Hx[:, 1:, :] - c.uH1[1:, :] * (EzJ[:, 1:, :] - Ez[:, 1:, :])
    Hy[:, :, 1:] = c.uH0[:, 1:] * Hy[:, :, 1:] + c.uH1[:, 1:] * (EzI[:, :, 1:] - Ez[:, :, 1:])
    if P == 0:
        return
    RA01 = pml["HRA"] - one
    RB, RE, RF = pml["HRB"], pml["HRE"], pml["HRF"]
    for side in ("x0", "y0", "xmax", "ymax"):   # slab order of compute.py:614
        psi = st.psi["H" + side]
        if side[0] == "x":
            rows = _slab_rows("H-" if side == "x0" else "H+", P, NI)
            d = (Ez[:, rows + 1, :NJ] - Ez[:, rows, :NJ]) / fdx
            a, b, e_, g = RA01[:, None], RB[:, None], RE[:, None], RF[:, None]
            Hy[:, rows, :NJ] = Hy[:, rows, :NJ] + c.uH4[rows, :NJ] * (a * d + b * psi)
        else:
            cols = _slab_rows("H-" if side == "y0" else "H+", P, NJ)
            d = (Ez[:, :NI, cols + 1] - Ez[:, :NI, cols]) / fdx
            a, b, e_, g = RA01[None, :], RB[None, :], RE[None, :], RF[None, :]
            Hx[:, :NI, cols] = Hx[:, :NI, cols] - c.uH4[:NI, cols] * (a * d + b * psi)
        st.psi["H" + side] = e_ * psi - g * d


def _e_phase(st: _State, c: Coeffs, pml: dict, P: int, dx, f):
    Ez, Hx, Hy = st.Ez, st.Hx, st.Hy
    S, NI1, NJ1 = Ez.shape
    NI, NJ = NI1 - 1, NJ1 - 1
    one = f(1.0)
    fdx = f(dx)
    # main update, fields_updates_gpu.cu:21-32
    Ez[:, 1:, 1:] = (c.uE0[1:, 1:] * Ez[:, 1:, 1:]
                     + c.uE1[1:, 1:] * (Hy[:, 1:, 1:] - Hy[:, :-1, 1:])) \
        - c.uE1[1:, 1:] * (Hx[:, 1:, 1:] - Hx[:, 1:, :-1])
    if P == 0:
        return
    RA01 = pml["ERA"] - one
    RB, RE, RF = pml["ERB"], pml["ERE"], pml["ERF"]
    for side in ("x0", "y0", "xmax", "ymax"):   # slab order of compute.py:659
        psi = st.psi["E" + side]
        if side[0] == "x":
            rows = _slab_rows("E-" if side == "x0" else "E+", P, NI)
            d = (Hy[:, rows, :NJ] - Hy[:, rows - 1, :NJ]) / fdx
            a, b, e_, g = RA01[:, None], RB[:, None], RE[:, None], RF[:, None]
            Ez[:, rows, :NJ] = Ez[:, rows, :NJ] + c.uE4[rows, :NJ] * (a * d + b * psi)
        else:
            cols = _slab_rows("E-" if side == "y0" else "E+", P, NJ)
            d = (Hx[:, :NI, cols] - Hx[:, :NI, cols - 1]) / fdx
            a, b, e_, g = RA01[None, :], RB[None, :], RE[None, :], RF[None, :]
            Ez[:, :NI, cols] = Ez[:, :NI, cols] - c.uE4[:NI, cols] * (a * d + b * psi)
        st.psi["E" + side] = e_ * psi - g * d


def forward(c: Coeffs, pml: dict, P: int, dx: float, src: np.ndarray,
            scale: np.ndarray, rx: np.ndarray, nt: int, dtype=np.float64,
            store: bool = True, hook=None):
    """Forward time loop (compute.py:574-712).

    src: int (S, nsrc, 2) cell indices (i, j); scale: (nsrc, nt) injection
    amounts (only row 0 is non-zero in the reference, compute.py:563-564);
    rx: int (S, nrx, 2).  Per step n: sample Ez at the receivers (before the
    update, sourcereceiver.cu:28-34), H update + H-CPML, E update + E-CPML,
    source Ez -= uE4*scale[n], store Ez[:, :NI, :NJ].
    Returns traces (S, nt, nrx) and snapshots (nt, S, NI, NJ) or None.
    ``hook(n, state)``, if given, sees the fields after each step.
    """
    f = dtype
    S = src.shape[0]
    NI, NJ = c.uE0.shape[0] - 1, c.uE0.shape[1] - 1
    st = _State(S, NI, NJ, P, f)
    nrx = rx.shape[1]
    traces = np.zeros((S, nt, nrx), f)
    snaps = np.zeros((nt, S, NI, NJ), f) if store else None
    sidx = np.arange(S)[:, None]
    sc = scale.astype(f)
    for n in range(nt):
        traces[:, n, :] = st.Ez[sidx, rx[:, :, 0], rx[:, :, 1]]
        _h_phase(st, c, pml, P, dx, f)
        _e_phase(st, c, pml, P, dx, f)
        for k in range(src.shape[1]):
            i, j = src[:, k, 0], src[:, k, 1]
            st.Ez[np.arange(S), i, j] = st.Ez[np.arange(S), i, j] - c.uE4[i, j] * sc[k, n]
        if store:
            snaps[n] = st.Ez[:, :NI, :NJ]
        if hook is not None:
            hook(n, st)
    return traces, snaps


def adjoint_reference(cadj: Coeffs, pml: dict, P: int, dx: float,
                      rx: np.ndarray, resid: np.ndarray, snaps: np.ndarray,
                      dtype=np.float64, hook=None):
    """Reference-mode adjoint (compute.py:1048-1172).

    lambda fields start at zero with the adjoint coefficients (uc.cu:60-91).
    For n = nt-1 .. 1: Ez[rx] -= resid[:, n, :] (back.cu:30-34, no scaling);
    E update + E-CPML; H update + H-CPML; then
    g_eps += sum_s (snap[n] - snap[n-1]) * lamEz and
    g_sig += sum_s snap[n] * lamEz.  Products are formed in ``dtype`` and
    summed in float64 (the reference sums in float32; the build accumulates
    in float64).  Returns the unnormalised (g_eps, g_sig) with g_sig still
    missing its constant dt factor (it cancels under normalisation).
    ``hook(n, state)``, if given, sees the lambda fields after step n.
    """
    f = dtype
    S, nt, nrx = resid.shape
    NI, NJ = cadj.uE0.shape[0] - 1, cadj.uE0.shape[1] - 1
    st = _State(S, NI, NJ, P, f)
    ge = np.zeros((NI, NJ), np.float64)
    gs = np.zeros((NI, NJ), np.float64)
    s_idx = np.repeat(np.arange(S), nrx)
    ri = rx[:, :, 0].reshape(-1)
    rj = rx[:, :, 1].reshape(-1)
    res = resid.astype(f)
    for n in range(nt - 1, 0, -1):
        np.subtract.at(st.Ez, (s_idx, ri, rj), res[:, n, :].reshape(-1))
        _e_phase(st, cadj, pml, P, dx, f)
        _h_phase(st, cadj, pml, P, dx, f)
        lam = st.Ez[:, :NI, :NJ]
        ge += ((snaps[n] - snaps[n - 1]) * lam).astype(np.float64).sum(axis=0)
        gs += (snaps[n] * lam).astype(np.float64).sum(axis=0)
        if hook is not None:
            hook(n, st)
    return ge, gs


def _slab_rows_e(side, P, n):
    return _slab_rows("E-" if side in ("x0", "y0") else "E+", P, n)


def _slab_rows_h(side, P, n):
    return _slab_rows("H-" if side in ("x0", "y0") else "H+", P, n)


def adjoint_exact(c: Coeffs, pml: dict, P: int, dx: float, dt: float, rx: np.ndarray,
                  resid: np.ndarray, snaps: np.ndarray):
    """Exact discrete adjoint (transpose) of ``forward`` in float64.

    Returns (dL/d er, dL/d se) for dL/dtraces = ``resid`` through the
    per-cell coefficients of ``cell_coeffs``, using
        dE^{n+1}/d er = -(e0/dt) uE4 (E^{n+1} - E^n),
        dE^{n+1}/d se = -uE4 (E^{n+1} + E^n)/2   (main-update cells),
    which hold because every term of the E update other than uE0 E^n is
    uE4 (or uE1 = uE4/dx) times something independent of the cell's own
    er / se.  The CPML sigma_max dependence on the mean of er over row 0 is
    not included (finite-difference checks keep row 0 fixed).
    """
    f = np.float64
    S, nt, nrx = resid.shape
    NI, NJ = c.uE0.shape[0] - 1, c.uE0.shape[1] - 1
    cc = Coeffs(*(np.asarray(a, f) for a in (c.uE0, c.uE1, c.uE4, c.uH0, c.uH1, c.uH4)))
    fdx = f(dx)
    e0dt = f(EPS0_KERNEL) / f(dt)
    pm = {k: np.asarray(v, f) for k, v in pml.items() if k != "smax"} if P else {}
    shp = (S, NI + 1, NJ + 1)
    lE = np.zeros(shp)
    lHx = np.zeros(shp)
    lHy = np.zeros(shp)
    mE = {sd: np.zeros((S, P, NJ) if sd[0] == "x" else (S, NI, P)) for sd in ("x0", "xmax", "y0", "ymax")}
    mH = {sd: np.zeros((S, P, NJ) if sd[0] == "x" else (S, NI, P)) for sd in ("x0", "xmax", "y0", "ymax")}
    ge = np.zeros((NI, NJ))
    gs = np.zeros((NI, NJ))
    ii, jj = np.meshgrid(np.arange(NI + 1), np.arange(NJ + 1), indexing="ij")
    mainE = (ii >= 1) & (jj >= 1)
    coefE = np.where(mainE, cc.uE0, 1.0)
    hxmain = ii >= 1
    hymain = jj >= 1
    s_idx = np.repeat(np.arange(S), nrx)
    ri = rx[:, :, 0].reshape(-1)
    rj = rx[:, :, 1].reshape(-1)
    res = np.asarray(resid, f)

    def up(X):      # X[i+1]
        Y = np.zeros_like(X); Y[:, :-1] = X[:, 1:]; return Y

    def down(X):    # X[i-1]
        Y = np.zeros_like(X); Y[:, 1:] = X[:, :-1]; return Y

    def left(X):    # X[j+1]
        Y = np.zeros_like(X); Y[:, :, :-1] = X[:, :, 1:]; return Y

    def right(X):   # X[j-1]
        Y = np.zeros_like(X); Y[:, :, 1:] = X[:, :, :-1]; return Y

    for n in range(nt - 1, -1, -1):
        e = lE
        En1 = np.asarray(snaps[n], f)
        En = np.asarray(snaps[n - 1], f) if n >= 1 else np.zeros_like(En1)
        eC = e[:, :NI, :NJ]
        u4 = cc.uE4[:NI, :NJ]
        mC = mainE[:NI, :NJ]
        ge += np.sum(-e0dt * u4 * eC * (En1 - En), axis=0)
        gs += np.sum(np.where(mC, -u4 * eC * (En1 + En) / 2, -0.5 * u4 * eC * (En1 - En)), axis=0)
        # transpose of the E half-step
        U = np.where(mainE, cc.uE1 * e, 0.0)
        Lx = np.zeros(shp)
        Ly = np.zeros(shp)
        if P:
            RA1, RB, RE, RF = pm["ERA"] - 1, pm["ERB"], pm["ERE"], pm["ERF"]
            for sd in ("x0", "xmax"):
                rows = _slab_rows_e(sd, P, NI)
                ec = e[:, rows, :NJ]
                u = cc.uE4[rows, :NJ]
                m = mE[sd]
                Lx[:, rows, :NJ] += u * RA1[:, None] * ec - RF[:, None] * m
                mE[sd] = u * RB[:, None] * ec + RE[:, None] * m
            for sd in ("y0", "ymax"):
                cols = _slab_rows_e(sd, P, NJ)
                ec = e[:, :NI, cols]
                u = cc.uE4[:NI, cols]
                m = mE[sd]
                Ly[:, :NI, cols] += -u * RA1[None, :] * ec - RF[None, :] * m
                mE[sd] = -u * RB[None, :] * ec + RE[None, :] * m
        lHy = lHy + U - up(U) + (Lx - up(Lx)) / fdx
        lHx = lHx - U + left(U) + (Ly - left(Ly)) / fdx
        lHx[:, NI, :] = 0
        lHx[:, :, NJ] = 0
        lHy[:, NI, :] = 0
        lHy[:, :, NJ] = 0
        # Hx on row 0 / Hy on column 0 (and Ez on both lines) form the closed,
        # identically zero boundary subsystems of the forward map: their
        # adjoints never reach the interior and carry zero gradient weight,
        # and that subsystem is unstable when run backwards, so it is held at 0.
        lHx[:, 0, :] = 0
        lHy[:, :, 0] = 0
        # transpose of the H half-step
        Vx = np.where(hxmain, cc.uH1 * lHx, 0.0)
        Vy = np.where(hymain, cc.uH1 * lHy, 0.0)
        Mx = np.zeros(shp)
        My = np.zeros(shp)
        if P:
            RA1, RB, RE, RF = pm["HRA"] - 1, pm["HRB"], pm["HRE"], pm["HRF"]
            for sd in ("x0", "xmax"):
                rows = _slab_rows_h(sd, P, NI)
                hc = lHy[:, rows, :NJ]
                u = cc.uH4[rows, :NJ]
                m = mH[sd]
                Mx[:, rows, :NJ] += u * RA1[:, None] * hc - RF[:, None] * m
                mH[sd] = u * RB[:, None] * hc + RE[:, None] * m
            for sd in ("y0", "ymax"):
                cols = _slab_rows_h(sd, P, NJ)
                hc = lHx[:, :NI, cols]
                u = cc.uH4[:NI, cols]
                m = mH[sd]
                My[:, :NI, cols] += -u * RA1[None, :] * hc - RF[None, :] * m
                mH[sd] = -u * RB[None, :] * hc + RE[None, :] * m
        lE = (coefE * e + Vx - right(Vx) - Vy + down(Vy)
              + (down(Mx) - Mx) / fdx + (right(My) - My) / fdx)
        lE[:, 0, :] = 0
        lE[:, :, 0] = 0
        np.add.at(lE, (s_idx, ri, rj), res[:, n, :].reshape(-1))
    return ge, gs


def normalise(g: np.ndarray) -> np.ndarray:
    """g / ||g||_2 (compute.py:1174-1180)."""
    return g / np.sqrt(np.sum(np.asarray(g, np.float64) ** 2))


# --- problem set-up and runs ------------------------------------------------------
@dataclass
class Problem:
    er: np.ndarray          # (NI, NJ) relative permittivity
    se: np.ndarray          # (NI, NJ) conductivity S/m
    mr: np.ndarray          # (NI, NJ) relative permeability
    dx: float
    dt: float
    nt: int
    P: int
    src: np.ndarray         # int (S, nsrc, 2)
    rx: np.ndarray          # int (S, nrx, 2)
    wavelet: np.ndarray     # (nsrc, nt) raw waveform samples (row 0 emits)


def setup(pb: Problem, dtype=np.float64, adjoint: bool = False):
    er = np.maximum(np.asarray(pb.er, np.float64), 1.0)     # compute.py:39
    se = np.maximum(np.asarray(pb.se, np.float64), 0.0)     # compute.py:38
    c = cell_coeffs(er, se, pb.mr, pb.dt, pb.dx, adjoint=adjoint, dtype=dtype)
    pml = cpml_profile(pb.P, pb.dx, pb.dt, er[0, :], np.asarray(pb.mr)[0, :], dtype) if pb.P else {}
    return c, pml


def run_forward(pb: Problem, dtype=np.float64, store=True):
    c, pml = setup(pb, dtype)
    scale = source_scale(pb.wavelet, pb.dx, dtype)
    return forward(c, pml, pb.P, pb.dx, pb.src, scale, pb.rx, pb.nt, dtype, store)


def run_gradient_exact(pb: Problem, resid: np.ndarray, snaps: np.ndarray):
    """Exact (unnormalised) dL/d er, dL/d se of the FP64 forward map."""
    c, pml = setup(pb, np.float64)
    return adjoint_exact(c, pml, pb.P, pb.dx, pb.dt, pb.rx, resid, snaps)


def run_gradient(pb: Problem, resid: np.ndarray, snaps: np.ndarray, dtype=np.float64):
    """Reference-mode normalised gradients (g_eps, g_sig) and the raw sums."""
    cadj, pml = setup(pb, dtype, adjoint=True)
    ge, gs = adjoint_reference(cadj, pml, pb.P, pb.dx, pb.rx, resid, snaps, dtype)
    return normalise(ge), normalise(gs), ge, gs
