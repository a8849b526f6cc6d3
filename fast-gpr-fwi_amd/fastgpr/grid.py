"""Host-side set-up of one solve: dt, step count, acquisition geometry,
Ricker wavelet and CPML profile.

These are small O(nt + P + shots) computations.  They are evaluated with
torch on the CPU in float32, the same dtypes and operation order as the
reference's torch code, so the values handed to the device match it:

* dt rule ............ compute.py:752-769
* geometry ........... compute.py:774-796
* nt ................. compute.py:798-803
* Ricker ............. compute.py:334-357, 383-400, 810-827
* CPML profile ....... compute.py:54-123, 166-222, 267-313, 828-830
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

# scipy.constants values (compute.py:5-6, src/__init__.py:3-4,21).
C0 = 299792458.0
EPS0 = 8.8541878188e-12
MU0 = 1.25663706127e-06
Z0 = (MU0 / EPS0) ** 0.5


def _cpu(x):
    return x.detach().to("cpu") if isinstance(x, torch.Tensor) else x


def _to_dev(t: torch.Tensor, dev) -> torch.Tensor:
    """Host tensor -> ``dev`` without waiting for queued device work (a
    pageable copy would): through pinned memory, asynchronously."""
    if t.device.type == "cpu" and torch.device(dev).type == "cuda":
        return t.contiguous().pin_memory().to(dev, non_blocking=True)
    return t.to(dev)


def cfl_dt(dx):
    """2-D TMz Courant limit, dy = dx (compute.py:754).  Keeps dx's type:
    a float32 tensor dx gives a float32 tensor dt, as in the reference."""
    dx = _cpu(dx)
    return 1 / (C0 * ((((1 / dx) * (1 / dx)) + ((1 / dx) * (1 / dx))) ** 0.5))


def resolve_dt(dx, dt):
    """compute.py:752-765: CFL when dt is None; a larger dt is clamped."""
    lim = cfl_dt(dx)
    if dt is None:
        return lim
    if dt > lim:
        print(lim)
        print(dt)
        print("dt is too large")
        return lim
    return dt


def n_iterations(time_windows, dt) -> int:
    """nt = ceil(tw / dt) + 1 evaluated on a float32 time-window tensor
    (compute.py:798-803)."""
    tw = time_windows if isinstance(time_windows, torch.Tensor) else torch.tensor(time_windows)
    return int(torch.ceil(_cpu(tw) / _cpu(dt))) + 1


def ricker(nt: int, dt, amp, freq, t_stop) -> torch.Tensor:
    """Ricker samples at t = n*dt for n < nt, zero outside [0, t_stop].

    Vectorised form of compute.py:351-357 with calculate_value's ricker branch
    (compute.py:383-400): chi = sqrt(2)/f, zeta = pi^2 f^2,
    w = -(2 zeta (2 zeta tau^2 - 1) exp(-zeta tau^2)) / (2 zeta) * amp.
    """
    f = torch.tensor(freq, dtype=torch.float)
    a = torch.tensor(amp, dtype=torch.float)
    chi = torch.sqrt(torch.tensor(2.0)) / f
    zeta = (torch.pi ** 2) * (f ** 2)
    dtc = _cpu(dt)
    if isinstance(dtc, torch.Tensor):
        t = dtc.float() * torch.arange(nt, dtype=torch.float)
    else:
        t = (torch.arange(nt, dtype=torch.float64) * dtc).float()
    stop = _cpu(t_stop) if isinstance(t_stop, torch.Tensor) else torch.tensor(t_stop, dtype=torch.float)
    inside = (t >= 0) & (t <= stop)
    delay = t - chi
    normalise = 1 / (2 * zeta)
    v = -(2 * zeta * (2 * zeta * delay ** 2 - 1) * torch.exp(-zeta * delay ** 2)) * normalise
    v = v * a
    return torch.where(inside, v, torch.zeros_like(v))


def pml_profile(P: int, dx, dt, er, mr) -> torch.Tensor:
    """(8, P) float32: ERA, ERB, ERE, ERF, HRA, HRB, HRE, HRF.

    sigma_max from the mean of er / mr over the x0 slab's first row
    (compute.py:278-283), quartic sigma, kappa = 1, alpha = 0, forward HORIPML
    coefficients (compute.py:211-222).  All four slabs share one CFS object,
    so one set of vectors serves x0, y0, xmax and ymax (compute.py:178, 829).
    """
    n = er.shape[1]
    if er.is_cuda:
        # on the device (no copy of the model to the host, no stall of the
        # launch queue); the sums are sequential FP32 like the loop below
        from . import _native
        er, mr = er.detach().float(), mr.detach().float().to(er.device)
        if mr.stride(1) != er.stride(1):
            mr = mr.contiguous()
            er = er.contiguous()
        sums = torch.empty(2, dtype=torch.float32, device=er.device)
        _native.check(_native.lib.fgpr_row_sums(
            er[0, :, 0].data_ptr(), mr[0, :, 0].data_ptr(), n, er.stride(1), sums.data_ptr(),
            torch.cuda.current_stream(er.device).cuda_stream), "row sums")
        sumer, summr = sums[0], sums[1]
        dev = er.device
    else:
        er0 = _cpu(er)[0, :, 0]
        mr0 = _cpu(mr)[0, :, 0]
        sumer = 0
        summr = 0
        for j in range(er0.shape[0]):       # sequential sum, as compute.py:278-281
            sumer = sumer + er0[j]
            summr = summr + mr0[j]
        dev = torch.device("cpu")
    er_mean = sumer / n
    mr_mean = summr / n
    d = _cpu(dx)
    if isinstance(d, torch.Tensor):
        d = _to_dev(d, dev)
    smax = (0.8 * (4 + 1)) / (Z0 * d * torch.sqrt(er_mean * mr_mean))
    L = P + 1
    tmp = (torch.linspace(0, (L - 1) + 0.5, steps=2 * L, device=dev) / (L - 1)) ** 4
    Esig = (tmp[0:-1:2] * (smax - 0) + 0)[:-1]
    Hsig = (tmp[1::2] * (smax - 0) + 0)[:-1]
    kap = torch.ones(P, device=dev)
    alp = torch.zeros(P, device=dev)
    dtc = _cpu(dt)
    if isinstance(dtc, torch.Tensor):
        dtc = _to_dev(dtc, dev)
    rows = []
    for sig in (Esig, Hsig):
        tmpd = (2 * EPS0 * kap) + dtc * (alp * kap + sig)
        rows.append((2 * EPS0 + dtc * alp) / tmpd)
        rows.append((2 * EPS0 * kap) / tmpd)
        rows.append(((2 * EPS0 * kap) - dtc * (alp * kap + sig)) / tmpd)
        rows.append((2 * sig * dtc) / (kap * tmpd))
    return torch.stack([r.float() for r in rows]).contiguous()


def _positions(n, ts, location, step, dx, device):
    """compute.py:775-781 / 786-792 evaluated on the CPU in float32."""
    location = _cpu(location)
    step = _cpu(step)
    pos = torch.zeros((ts, n, 3))
    for c in range(3):
        pos[:, :, c] = (torch.arange(n) * step[c] + location[c]).repeat(ts, 1)
    pos = pos + torch.arange(ts).view(-1, 1, 1) * step
    return torch.round(pos / _cpu(dx)).int().contiguous()


def geometry(nsrc, nrx, ts, source_location, source_step, receiver_location,
             receiver_step, dx, Fixedreceiver=False, customsource=None,
             customreceiver=None):
    """(src, rx) int32 (ts, n, 3) cell indices (compute.py:774-796)."""
    if customsource is None:
        src = _positions(nsrc, ts, source_location, source_step, dx, "cpu")
    else:
        src = _cpu(customsource).int()
    if customreceiver is None:
        rx = _positions(nrx, ts, receiver_location, receiver_step, dx, "cpu")
        if Fixedreceiver:
            rx[:, :, :] = rx[0, :, :].unsqueeze(0)
    else:
        rx = _cpu(customreceiver).int()
    return src, rx


def device_cells(idx: torch.Tensor, NI: int, NJ: int, what: str) -> torch.Tensor:
    """(S, n, 3) int cell indices -> (S, n, 2) int32 (i, j) for the device.

    Entries on the dead k = 1 plane of the reference's padded arrays (k != 0)
    can never emit or record (their coefficients are zero, uc.cu:19), so they
    become dead markers (i = -1).  Indices outside the padded grid raise.
    """
    idx = idx.to(torch.int64)
    i, j, k = idx[..., 0], idx[..., 1], idx[..., 2]
    bad = (i < 0) | (i > NI) | (j < 0) | (j > NJ) | (k < 0)
    if bool(bad.any()):
        raise ValueError(f"{what} cell index outside the grid [0,{NI}]x[0,{NJ}]")
    out = torch.stack([torch.where(k == 0, i, torch.full_like(i, -1)), j], dim=-1)
    return out.to(torch.int32).contiguous()


@dataclass
class FDTDGrid:
    """Solve description shared by the forward and adjoint passes (the
    reference's ``G`` / ``H``, compute.py:442-518, 945-1017).  Holds no
    device state; the solver allocates its buffers per call."""
    device: torch.device
    NI: int
    NJ: int
    S: int
    nsrc: int
    nrx: int
    dx: object
    dt: object
    iterations: int
    pmlthick: int
    source: torch.Tensor            # int32 (S, nsrc, 3), as G.source
    receiver: torch.Tensor          # int32 (S, nrx, 3), as G.receiver
    srcwaves: torch.Tensor          # float32 (nsrc, nt): row 0 = Ricker, others 0
    pml: torch.Tensor | None        # float32 (8, P)
    tv: bool = True
    exec_mode: str = "graph"
    graph_chunk: int = 32
    time_block: tuple | None = None   # (region rows, threads, steps per launch) or None
    options: dict = field(default_factory=dict)

    @property
    def dx32(self) -> float:
        return float(np.float32(float(self.dx)))

    @property
    def dt32(self) -> float:
        return float(np.float32(float(self.dt)))
