"""Seeded synthetic workloads used by the benchmark and the tests.

Each builder returns a ``Workload``: true and initial models, grid, time
stepping and acquisition as integer cell indices.  No data files are read.

* ``c1_plain``      200 x 100 grid, dx 0.05 m, 20-cell CPML, dt 0.9 CFL,
                    1500 steps, Ricker 100 MHz, 1 shot at (100, 25),
                    40 receivers (61 + 2r, 75); two-blob eps, one-blob sigma.
* ``c2_crosshole``  400 x 200 grid, dx 0.025 m, 3000 steps, Ricker 200 MHz,
                    13 sources down column 30, 80 receivers down column 170,
                    8 seeded layers + 3 inclusions.
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
