"""Seeded synthetic workloads used by the benchmark and the tests.

Each builder returns a ``Workload``: true and initial models, grid, time
stepping and acquisition as integer cell indices.  No data files are read.

* ``c1_plain``      200 x 100 grid, dx 0.05 m, 20-cell CPML, dt 0.9 CFL,
                    1500 steps, Ricker 100 MHz, 1 shot at (100, 25),
                    40 receivers (61 + 2r, 75); two-blob eps, one-blob sigma.
* ``c2_crosshole``  400 x 200 grid, dx 0.025 m, 3000 steps, Ricker 200 MHz,
                    13 sources down column 30, 80 receivers down column 170,
                    8 seeded layers + 3 inclusions.
* ``c4_memory``     2000 x 1000 grid, dx 0.01 m, 8000 steps, 4 shots,
                    200 receivers, Ricker 400 MHz (checkpointed gradients).
* ``layered_blob``  generic layered + blob model of any size (uneven batches,
                    serving requests).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .grid import cfl_dt, n_iterations


@dataclass
class Workload:
    name: str
    er: np.ndarray          # (NI, NJ) true relative permittivity
    se: np.ndarray          # (NI, NJ) true conductivity [S/m]
    mr: np.ndarray          # (NI, NJ) relative permeability
    er0: np.ndarray         # initial (smoothed) model
    se0: np.ndarray
    dx: float
    dt: float
    nt: int
    P: int
    freq: float
    src: np.ndarray         # int32 (S, nsrc, 3)
    rx: np.ndarray          # int32 (S, nrx, 3)

    @property
    def NI(self):
        return self.er.shape[0]

    @property
    def NJ(self):
        return self.er.shape[1]

    @property
    def S(self):
        return self.src.shape[0]

    @property
    def time_window(self) -> float:
        """A time window whose float32 step count (compute.py:803) is nt."""
        return window_for(self.nt, self.dt)

    def cell_updates_per_gradient(self) -> int:
        return 2 * self.S * self.NI * self.NJ * self.nt


def window_for(nt: int, dt: float) -> float:
    tw = (nt - 1.5) * dt
    assert n_iterations(tw, dt) == nt, (nt, dt)
    return tw


def _blob(NI, NJ, ci, cj, radius):
    ii, jj = np.meshgrid(np.arange(NI), np.arange(NJ), indexing="ij")
    r2 = (ii - ci) ** 2 + (jj - cj) ** 2
    return np.exp(-r2 / (2.0 * radius ** 2))


def _smooth(a, sigma):
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(a, sigma, mode="nearest")


def _cells(points):
    """list of (i, j) per shot -> int32 (S, n, 3) with k = 0."""
    arr = np.asarray(points, dtype=np.int32)
    return np.concatenate([arr, np.zeros(arr.shape[:-1] + (1,), np.int32)], axis=-1)


def c1_plain() -> Workload:
    NI, NJ, P = 200, 100, 20
    rng = np.random.default_rng(0)
    lo_i, hi_i, lo_j, hi_j = P + 10, NI - P - 10, P + 10, NJ - P - 10
    centres = [(rng.integers(lo_i, hi_i), rng.integers(lo_j, hi_j)) for _ in range(2)]
    b1 = _blob(NI, NJ, *centres[0], 8.0)
    b2 = _blob(NI, NJ, *centres[1], 8.0)
    er = 6.0 + 3.0 * np.maximum(b1, b2)
    se = 2e-3 + 3e-3 * b1
    mr = np.ones((NI, NJ))
    dx = 0.05
    dt = 0.9 * float(cfl_dt(dx))
    src = _cells([[(100, 25)]])
    rx = _cells([[(61 + 2 * r, 75) for r in range(40)]])
    return Workload("c1_plain_200x100", er.astype(np.float32), se.astype(np.float32),
                    mr.astype(np.float32), _smooth(er, 10).astype(np.float32),
                    _smooth(se, 10).astype(np.float32), dx, dt, 1500, P, 1e8, src, rx)


def layered_blob(NI, NJ, seed, nlayers=8, eps=(4.0, 12.0), sig=(1e-3, 1e-2), nblobs=3,
                 P=20):
    """Horizontal layers (uniform eps, log-uniform sigma) plus Gaussian
    inclusions, all drawn from ``default_rng(seed)``."""
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.choice(np.arange(1, NI), size=nlayers - 1, replace=False))
    bounds = np.concatenate([[0], cuts, [NI]])
    er = np.empty((NI, NJ))
    se = np.empty((NI, NJ))
    for a, b in zip(bounds[:-1], bounds[1:]):
        er[a:b] = rng.uniform(*eps)
        se[a:b] = np.exp(rng.uniform(np.log(sig[0]), np.log(sig[1])))
    for _ in range(nblobs):
        ci = rng.integers(P + 5, NI - P - 5)
        cj = rng.integers(P + 5, NJ - P - 5)
        rad = rng.uniform(4, max(5.0, min(NI, NJ) / 12))
        b = _blob(NI, NJ, ci, cj, rad)
        er = er + rng.uniform(-2, 3) * b
        se = se * (1 + rng.uniform(-0.5, 1.0) * b)
    er = np.clip(er, eps[0] - 1, eps[1] + 3)
    se = np.clip(se, 0.5 * sig[0], 2 * sig[1])
    return er, se


def c2_crosshole() -> Workload:
    NI, NJ, P = 400, 200, 20
    er, se = layered_blob(NI, NJ, seed=1)
    dx = 0.025
    dt = 0.9 * float(cfl_dt(dx))
    si = np.round(np.linspace(40, 360, 13)).astype(int)
    src = _cells([[(int(i), 30)] for i in si])
    rx = _cells([[(40 + 4 * r, 170) for r in range(80)] for _ in si])
    mr = np.ones((NI, NJ))
    return Workload("c2_crosshole_400x200_13shots", er.astype(np.float32), se.astype(np.float32),
                    mr.astype(np.float32), _smooth(er, 10).astype(np.float32),
                    _smooth(se, 10).astype(np.float32), dx, dt, 3000, P, 2e8, src, rx)


def crosshole(NI: int, NJ: int, S: int, nt: int, nrx: int = 80, P: int = 20) -> Workload:
    """The cross-hole geometry scaled to any grid: S sources down a column at
    15 % of the width, nrx receivers down a column at 85 % (both at least 5
    cells inside the CPML), over the middle 80 % of the depth (again at
    least 5 cells inside the CPML); the 8-layer
    model of seed 1, dx 0.025 m, Ricker 200 MHz.  The autotuner envelope's
    grids (``bench.py --envelope``)."""
    er, se = layered_blob(NI, NJ, seed=1, P=P)
    dx = 0.025
    dt = 0.9 * float(cfl_dt(dx))
    lo, hi = max(0.1 * NI, P + 5), min(0.9 * NI, NI - P - 6)
    si = np.round(np.linspace(lo, hi, S)).astype(int)
    ri = [int(lo + r * (hi - lo) / nrx) for r in range(nrx)]
    sj, rj = max(int(0.15 * NJ), P + 5), min(int(0.85 * NJ), NJ - P - 6)
    src = _cells([[(int(i), sj)] for i in si])
    rx = _cells([[(i, rj) for i in ri] for _ in si])
    return Workload(f"crosshole_{NI}x{NJ}_{S}shots", er.astype(np.float32), se.astype(np.float32),
                    np.ones((NI, NJ), np.float32), _smooth(er, 10).astype(np.float32),
                    _smooth(se, 10).astype(np.float32), dx, dt, nt, P, 2e8, src, rx)


def small_case(NI=60, NJ=44, P=8, S=3, nt=300, nrx=6, seed=5, dx=0.05, freq=2e8) -> Workload:
    """A few-second test problem: heterogeneous model, one PEC cell, shots
    with different sources, receivers including a colliding pair."""
    er, se = layered_blob(NI, NJ, seed=seed, nlayers=4, nblobs=2, P=P)
    se[NI // 2, NJ // 3] = 500.0                    # PEC cell (uc.cu:28-32)
    rng = np.random.default_rng(seed + 100)
    mr = (1.0 + 0.3 * rng.random((NI, NJ))).astype(np.float32)
    dt = 0.9 * float(cfl_dt(dx))
    src = _cells([[(int(rng.integers(P + 2, NI - P - 2)), int(rng.integers(P + 2, NJ - P - 2)))]
                  for _ in range(S)])
    rx_pts = []
    for s in range(S):
        pts = [(int(rng.integers(1, NI - 1)), int(rng.integers(1, NJ - 1))) for _ in range(nrx)]
        pts[-1] = pts[0]                            # two receivers on one cell
        rx_pts.append(pts)
    rx = _cells(rx_pts)
    return Workload(f"small_{NI}x{NJ}", er.astype(np.float32), se.astype(np.float32), mr,
                    er.astype(np.float32), se.astype(np.float32), dx, dt, nt, P, freq, src, rx)


def c4_memory(nt: int = 8000) -> Workload:
    """2000 x 1000 grid at dx 0.01 m, 8000 steps, 4 shots, 200 receivers,
    Ricker 400 MHz (>= 20 cells per shortest wavelength at eps <= 12).
    Storing every snapshot would take ~64 GB per shot: the gradient uses
    checkpointing."""
    NI, NJ, P = 2000, 1000, 20
    er, se = layered_blob(NI, NJ, seed=3, nlayers=10, nblobs=6)
    dx = 0.01
    dt = 0.9 * float(cfl_dt(dx))
    src = _cells([[(i, 100)] for i in (400, 800, 1200, 1600)])
    rx = _cells([[(100 + 9 * r, 900) for r in range(200)] for _ in range(4)])
    mr = np.ones((NI, NJ))
    return Workload("c4_memory_2000x1000_4shots", er.astype(np.float32), se.astype(np.float32),
                    mr.astype(np.float32), _smooth(er, 20).astype(np.float32),
                    _smooth(se, 20).astype(np.float32), dx, dt, nt, P, 4e8, src, rx)


def crop(wl: Workload, i0: int, j0: int, NI: int, NJ: int, nt: int, shots=None,
         name: str | None = None, receivers=None) -> Workload:
    """A sub-grid of a workload (same dx, dt, CPML thickness and frequency),
    keeping the sources that fall inside, re-indexed.  ``receivers`` (crop
    cell indices, shared by every shot) replaces the receivers; otherwise the
    ones inside the crop are kept."""
    sl = (slice(i0, i0 + NI), slice(j0, j0 + NJ))
    pick = range(wl.S) if shots is None else shots
    src, rx = [], []
    for s in pick:
        ss = [(i - i0, j - j0) for i, j, _ in wl.src[s] if i0 <= i < i0 + NI and j0 <= j < j0 + NJ]
        rr = ([tuple(p) for p in receivers] if receivers is not None else
              [(i - i0, j - j0) for i, j, _ in wl.rx[s] if i0 <= i < i0 + NI and j0 <= j < j0 + NJ])
        if ss and rr:
            src.append(ss)
            rx.append(rr)
    if not rx:
        raise ValueError("no shot has both a source and a receiver inside the crop")
    n = min(len(r) for r in rx)
    rx = [r[:n] for r in rx]
    return Workload(name or f"{wl.name}_crop{NI}x{NJ}", wl.er[sl].copy(), wl.se[sl].copy(),
                    wl.mr[sl].copy(), wl.er0[sl].copy(), wl.se0[sl].copy(), wl.dx, wl.dt, nt,
                    wl.P, wl.freq, _cells(src), _cells(rx))


def c3_uneven(seed: int = 2) -> list:
    """64 shots over 4 seeded models of different sizes and step counts:
    300x150/2000, 500x250/3000, 800x400/4500, 1200x600/6000 steps, dx 0.025 m,
    Ricker 200 MHz, 20-cell CPML; 8-24 sources and 50-150 receivers per model,
    layered + blob eps in [3, 15], sigma in [0.5, 20] mS/m."""
    rng = np.random.default_rng(seed)
    dims = [(300, 150, 2000), (500, 250, 3000), (800, 400, 4500), (1200, 600, 6000)]
    while True:
        shots = rng.integers(8, 25, size=4)
        if shots.sum() == 64:
            break
    out = []
    for m, ((NI, NJ, nt), S) in enumerate(zip(dims, shots)):
        er, se = layered_blob(NI, NJ, seed=seed * 10 + m, eps=(3.0, 12.0), sig=(5e-4, 2e-2),
                              nblobs=4)
        er = np.clip(er, 3.0, 15.0)
        se = np.clip(se, 5e-4, 2e-2)
        nrx = int(rng.integers(50, 151))
        dx = 0.025
        dt = 0.9 * float(cfl_dt(dx))
        sj = int(rng.integers(25, NJ // 4))
        rj = int(rng.integers(3 * NJ // 4, NJ - 25))
        si = np.round(np.linspace(30, NI - 31, int(S))).astype(int)
        ri = np.round(np.linspace(30, NI - 31, nrx)).astype(int)
        src = _cells([[(int(i), sj)] for i in si])
        rx = _cells([[(int(i), rj) for i in ri] for _ in si])
        out.append(Workload(f"c3_model{m}_{NI}x{NJ}_{int(S)}shots", er.astype(np.float32),
                            se.astype(np.float32), np.ones((NI, NJ), np.float32),
                            _smooth(er, 10).astype(np.float32), _smooth(se, 10).astype(np.float32),
                            dx, dt, nt, 20, 2e8, src, rx))
    return out
