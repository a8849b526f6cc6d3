"""Device side of one solve: buffers, coefficient set-up and the native
forward / adjoint sweeps (the reference's ``Forward`` compute.py:542-720 and
``Backward`` compute.py:1020-1180).

All device memory is allocated here as torch tensors (so the caching
allocator and any HBM budget see it) and handed to libfastgpr by pointer.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

from . import _native
from . import trace as _trace
from ._native import check, lib

_MODES = {"eager": _native.EXEC_EAGER, "graph": _native.EXEC_GRAPH}


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def to_device(t: torch.Tensor, device) -> torch.Tensor:
    """Host tensor -> device through pinned memory, asynchronously: a copy
    from pageable memory first waits for all queued device work, which would
    idle the GPU while the host prepares the next sweep."""
    if t.device.type != "cpu":
        return t.to(device)
    return t.contiguous().pin_memory().to(device, non_blocking=True)


def device_empty(shape, dtype, device) -> torch.Tensor:
    """torch.empty that, when the caching allocator's free blocks are too
    fragmented for a large buffer, releases them and tries once more."""
    try:
        return torch.empty(shape, dtype=dtype, device=device)
    except torch.OutOfMemoryError:
        torch.cuda.empty_cache()
        return torch.empty(shape, dtype=dtype, device=device)


def require_hip(device) -> torch.device:
    device = torch.device(device)
    if device.type != "cuda":
        raise RuntimeError("fastgpr runs on a HIP (ROCm) device only; got device "
                           f"{device}. There is no CPU path.")
    if not torch.cuda.is_available():
        raise RuntimeError("fastgpr needs a ROCm GPU; torch.cuda.is_available() is False")
    return device


def cell_coeffs(er, se, mr, dt32: float, dx32: float, adjoint: bool, device,
                out=None) -> torch.Tensor:
    """5 x (NI+1) x pitch coefficient planes (Ucget / Ucgeta, uc.cu:48, 94),
    into ``out`` when given.  Contiguous float32 device inputs are read in
    place (nothing is allocated)."""
    NI, NJ = er.shape[0], er.shape[1]
    pitch = _native.pitch(NJ)
    e, s, m = (to_device(x.detach().reshape(NI, NJ).to(torch.float32), device).contiguous()
               for x in (er, se, mr))
    coef = (torch.empty((5, NI + 1, pitch), dtype=torch.float32, device=device)
            if out is None else out)
    check(lib.fgpr_cell_coeffs(_ptr(e), _ptr(s), _ptr(m), NI, NJ, dt32, dx32,
                               1 if adjoint else 0, _ptr(coef), _stream(device)), "cell_coeffs")
    return coef


TB_MAX_LIST = 128   # receivers one time-blocked region can hold (entries per wave list)
NO_ENTRY = -1       # empty slot of a receiver block (0xFFFFFFFF)


def _blocks(keys: np.ndarray, ents: np.ndarray, nkeys: int, cap: int) -> np.ndarray:
    """Entries grouped per key into blocks of ``cap`` slots, in the given
    (receiver) order, the rest of each block NO_ENTRY."""
    out = np.full((max(nkeys, 1), cap), NO_ENTRY, np.int64)
    if len(keys):
        order = np.argsort(keys, kind="stable")
        k = keys[order]
        start = np.searchsorted(k, k, side="left")
        out[k, np.arange(len(k)) - start] = ents[order]
    return out.astype(np.uint32).view(np.int32).reshape(-1)


_TILE_CACHE: dict = {}
_TILE_CACHE_MAX = 16


def tile_tables(rx: np.ndarray, NI: int, NJ: int, tb) -> dict:
    """``_tile_tables`` of the same receivers, grid and variant computed once
    (a job's forward and adjoint batches, repeated geometries): the last
    few results are kept; callers copy the arrays and never write them."""
    rx = np.ascontiguousarray(rx, np.int64)
    key = (rx.shape, rx.tobytes(), int(NI), int(NJ), tuple(tb))
    hit = _TILE_CACHE.pop(key, None)
    if hit is None:
        hit = _tile_tables(rx, NI, NJ, tb)
    _TILE_CACHE[key] = hit                       # most recent last
    while len(_TILE_CACHE) > _TILE_CACHE_MAX:
        _TILE_CACHE.pop(next(iter(_TILE_CACHE)))
    return hit


def _tile_tables(rx: np.ndarray, NI: int, NJ: int, tb) -> dict:
    """Receiver blocks of the time-blocked kernels (host, numpy).

    rx: (S, nrx, 2) int cell indices, i < 0 = dead.  Forward: each live
    receiver is sampled by the one tile whose centre holds it.  Adjoint: each
    receiver on an updated cell (i < NI, j < NJ) is injected by every tile
    whose region (centre + halo) holds it.  Per (shot, tile) a block of
    ``cap`` (64 or 128) entries li << 26 | lj << 20 | r, the receiver's cell
    relative to the region's first row and column and its index, in receiver
    order (the order the adjoint applies receivers sharing a cell in)."""
    ri, nt_, k = tb
    TI, TJ, ti, tj = _native.tb_tiles(NI, NJ, ri, nt_, k)
    S, nrx = rx.shape[0], rx.shape[1]
    if nrx >= 1 << 20:
        raise ValueError("the time-blocked path holds at most 2^20 - 1 receivers per shot")
    ntiles = ti * tj
    # every (shot, receiver) in shot-major, receiver order
    sh = np.repeat(np.arange(S, dtype=np.int64), nrx)
    r = np.tile(np.arange(nrx, dtype=np.int64), S)
    i = np.asarray(rx[..., 0], np.int64).reshape(-1)
    j = np.asarray(rx[..., 1], np.int64).reshape(-1)
    live = i >= 0
    sh, r, i, j = sh[live], r[live], i[live], j[live]
    a, b = i // TI, j // TJ
    fk = sh * ntiles + a * tj + b
    fv = ((i - a * TI + k) << 26) | ((j - b * TJ + k) << 20) | r
    # adjoint: every region (centre + halo) holding an updated receiver cell;
    # a region spans fewer than ri // TI + 2 tile rows and 64 // TJ + 2 columns
    upd = (i < NI) & (j < NJ)
    su, ru, iu, ju = sh[upd], r[upd], i[upd], j[upd]
    a_lo = np.maximum(0, (iu + k - ri) // TI)
    a_hi = np.minimum(ti - 1, (iu + k) // TI)
    b_lo = np.maximum(0, (ju + k - 64) // TJ)
    b_hi = np.minimum(tj - 1, (ju + k) // TJ)
    keys, rs, vals = [], [], []
    for da in range(ri // TI + 2):
        aa = a_lo + da
        oka = (aa <= a_hi) & (aa * TI - k <= iu) & (iu < aa * TI - k + ri)
        for db in range(64 // TJ + 2):
            bb = b_lo + db
            ok = oka & (bb <= b_hi) & (bb * TJ - k <= ju) & (ju < bb * TJ - k + 64)
            keys.append((su * ntiles + aa * tj + bb)[ok])
            rs.append(ru[ok])
            vals.append((((iu - aa * TI + k) << 26) | ((ju - bb * TJ + k) << 20) | ru)[ok])
    ak = np.concatenate(keys) if keys else np.zeros(0, np.int64)
    rr = np.concatenate(rs) if rs else np.zeros(0, np.int64)
    av = np.concatenate(vals) if vals else np.zeros(0, np.int64)
    order = np.lexsort((rr, ak))          # per region, in receiver order
    ak, av = ak[order], av[order]
    most = max(int(np.bincount(fk, minlength=1).max()) if len(fk) else 0,
               int(np.bincount(ak, minlength=1).max()) if len(ak) else 0)
    if most > TB_MAX_LIST:
        raise ValueError(f"time-blocked variant {tuple(tb)}: {most} receivers in one tile region "
                         f"(at most {TB_MAX_LIST}); use smaller regions or the per-step kernels")
    cap = 64 if most <= 64 else 128
    return dict(TI=TI, TJ=TJ, ti=ti, tj=tj, cap=cap,
                fwd=_blocks(fk, fv, S * ntiles, cap),
                adj=_blocks(ak, av, S * ntiles, cap))


class Batch:
    """Device buffers of one batch of S shots on one grid.

    ``tb`` = (region rows, threads, steps per launch) selects the
    time-blocked kernels; None selects the per-step kernels.  ``shots`` =
    (lo, hi) restricts the batch to shots lo..hi-1 of ``G`` (a shot group);
    ``traces`` may then be the group's slice of a whole-batch tensor.

    ``alloc(shape, dtype)`` supplies the large buffers (fields, CPML memory,
    traces, snapshots) uninitialised -- a caller carving them from its own
    device arena then clears them with ``reset_state`` and ``traces.zero_()``
    in stream order; by default they are zeroed torch tensors."""

    def __init__(self, G, store: bool = True, tb=None, tiled=None, shots=None, traces=None,
                 alloc=None):
        dev = G.device
        self.G = G
        self.alloc = alloc
        zeros = ((lambda shape, dt: torch.zeros(shape, dtype=dt, device=dev)) if alloc is None
                 else alloc)
        self.lo, self.hi = shots if shots is not None else (0, G.S)
        self.NI, self.NJ, self.S, self.nt = G.NI, G.NJ, self.hi - self.lo, G.iterations
        self.P = G.pmlthick
        self.pitch = _native.pitch(self.NJ)
        self.fields = zeros((3, self.S, self.NI + 1, self.pitch), torch.float32)
        npsi = max(_native.psi_floats(self.NI, self.NJ, self.P, self.S), 1)
        self.psi = zeros((npsi,), torch.float32)
        self.pml = (to_device(G.pml, dev).contiguous() if self.P > 0
                    else torch.zeros(8, dtype=torch.float32, device=dev))
        from .grid import device_cells
        if G.source.shape[0] < self.hi or G.receiver.shape[0] < self.hi:
            raise ValueError(f"source/receiver arrays have fewer rows than the {self.hi} shots")
        self.src = to_device(device_cells(G.source[self.lo:self.hi], self.NI, self.NJ, "source"), dev)
        rx_host = device_cells(G.receiver[self.lo:self.hi], self.NI, self.NJ, "receiver")
        self.rx = to_device(rx_host, dev)
        self.nsrc = self.src.shape[1]
        self.nrx = self.rx.shape[1]
        self.ctr = torch.zeros(1, dtype=torch.int32, device=dev)
        self.traces = (traces if traces is not None else
                       zeros((self.S, self.nt, self.nrx), torch.float32))
        self.tb = tuple(tb) if tb else None
        # Snapshots in the time-blocked layout when both sweeps run the same
        # time-blocked variant (the exact adjoint reads the row-major one).
        if tiled is None:
            tiled = bool(self.tb) and G.options.get("gradient", "reference") != "exact"
        self.tiled = bool(tiled) and bool(self.tb)
        self.gshot = None
        if self.tb:
            self.fields2 = zeros(tuple(self.fields.shape), torch.float32)
            self.psi2 = zeros(tuple(self.psi.shape), torch.float32)
            t = tile_tables(rx_host.numpy(), self.NI, self.NJ, self.tb)
            self.tiles = (t["ti"], t["tj"])
            as_dev = lambda a: to_device(torch.from_numpy(np.ascontiguousarray(a)), dev)
            self.tb_tables = [as_dev(t["fwd"]), as_dev(t["adj"])]
            self.tb_list_cap = t["cap"]
        empty = ((lambda shape, dt: device_empty(shape, dt, dev)) if alloc is None
                 else alloc)
        self.snaps = empty((self.nt,) + self.snap_slot_shape(), torch.float32) if store else None

    def snap_slot_shape(self) -> tuple:
        """Shape of one snapshot slot: (S, NI, pitch) row-major, or
        (S, floats per shot) in the time-blocked layout."""
        if self.tiled:
            return (self.S, _native.tb_snap_floats(self.NI, self.NJ, self.tb))
        return (self.S, self.NI, self.pitch)

    def seed_slot(self, p, slot: int, ring: torch.Tensor, ez: torch.Tensor):
        """Write snapshot slot ``slot`` of ``ring`` (the buffer ``p.snaps``
        points to) from a row-major Ez state (S, NI+1, pitch)."""
        if self.tiled:
            check(lib.fgpr_snap_pack(ctypes.byref(p), slot, _ptr(ez), _stream(ez.device)), "snap_pack")
        else:
            ring[slot].copy_(ez[:, :self.NI, :])

    def reset_state(self):
        self.fields.zero_()
        self.psi.zero_()
        if self.tb:
            self.fields2.zero_()
            self.psi2.zero_()

    def gshot_buffer(self, zero: bool = True):
        """Per-shot FP64 partial gradients of the time-blocked adjoint,
        (S, 2, NI, NJ), zeroed unless ``zero`` is false."""
        if self.gshot is None:
            shape = (self.S, 2, self.NI, self.NJ)
            self.gshot = (torch.empty(shape, dtype=torch.float64, device=self.fields.device)
                          if self.alloc is None else self.alloc(shape, torch.float64))
        if zero:
            self.gshot.zero_()
        return self.gshot

    def problem(self, coef, srcscale=None, resid=None, grad_eps=None, grad_sig=None):
        G = self.G
        p = _native.Problem()
        p.NI, p.NJ, p.P, p.S, p.nt = self.NI, self.NJ, self.P, self.S, self.nt
        p.nsrc, p.nrx, p.pitch = self.nsrc, self.nrx, self.pitch
        p.dx, p.dt = G.dx32, G.dt32
        p.coef = coef.data_ptr()
        p.pml = self.pml.data_ptr()
        p.src = self.src.data_ptr()
        p.srcscale = srcscale.data_ptr() if srcscale is not None else None
        p.rx = self.rx.data_ptr()
        p.Ez = self.fields[0].data_ptr()
        p.Hx = self.fields[1].data_ptr()
        p.Hy = self.fields[2].data_ptr()
        p.psi = self.psi.data_ptr()
        p.traces = self.traces.data_ptr()
        p.snaps = self.snaps.data_ptr() if self.snaps is not None else None
        p.snap_base = 0
        p.snap_mod = 0
        p.nsnap = self.snaps.shape[0] if self.snaps is not None else 0
        p.snap_tiled = 1 if self.tiled else 0
        p.resid = resid.data_ptr() if resid is not None else None
        p.grad_eps = grad_eps.data_ptr() if grad_eps is not None else None
        p.grad_sig = grad_sig.data_ptr() if grad_sig is not None else None
        p.ctr = self.ctr.data_ptr()
        if self.tb:
            p.Ez2 = self.fields2[0].data_ptr()
            p.Hx2 = self.fields2[1].data_ptr()
            p.Hy2 = self.fields2[2].data_ptr()
            p.psi2 = self.psi2.data_ptr()
            p.tb_ri, p.tb_nt, p.tb_k = self.tb
            p.tb_ti, p.tb_tj = self.tiles
            p.tb_rx_list, p.tb_inj_list = [t.data_ptr() for t in self.tb_tables]
            p.tb_list_cap = self.tb_list_cap
            p.gshot = self.gshot.data_ptr() if self.gshot is not None else None
        return p


class ShotGroups:
    """One batch split into shot groups whose sweeps run on separate HIP
    streams.  A time-blocked launch of one group finishes its last, partial
    round of workgroups on part of the CUs; the other group's launches run on
    the rest, so the tails of the launches overlap instead of idling CUs.
    Each group keeps its own state; traces and the per-shot partial
    gradients are slices of whole-batch tensors, and the gradient is reduced
    once over all shots in shot order, as for one batch."""

    def __init__(self, G, store: bool, tb, n: int):
        dev = G.device
        self.G, self.S, self.nt, self.NI, self.NJ, self.tb = G, G.S, G.iterations, G.NI, G.NJ, tb
        cuts = [round(k * G.S / n) for k in range(n + 1)]
        self.nrx = int(G.receiver.shape[1])
        self.traces = torch.zeros((self.S, self.nt, self.nrx), dtype=torch.float32, device=dev)
        self.groups = [Batch(G, store=store, tb=tb, shots=(cuts[k], cuts[k + 1]),
                             traces=self.traces[cuts[k]:cuts[k + 1]]) for k in range(n)]
        self.gshot = None

    def reset_state(self):
        for g in self.groups:
            g.reset_state()

    def gshot_buffer(self):
        if self.gshot is None:
            self.gshot = torch.zeros((self.S, 2, self.NI, self.NJ), dtype=torch.float64,
                                     device=self.traces.device)
        else:
            self.gshot.zero_()
        for g in self.groups:
            g.gshot = self.gshot[g.lo:g.hi]
        return self.gshot


_SIDE_STREAMS: dict = {}


def _side_streams(device, n: int) -> list:
    """Per-device side streams (default priority) for shot groups."""
    key = torch.device(device).index or 0
    have = _SIDE_STREAMS.setdefault(key, [])
    while len(have) < n:
        have.append(torch.cuda.Stream(device))
    return have[:n]


def _fan_out(device, groups, run):
    """run(group, stream) for every group on its own side stream, ordered
    after the work queued so far on the current stream, which then waits for
    all of them."""
    main = torch.cuda.current_stream(device)
    start = torch.cuda.Event()
    start.record(main)
    streams = _side_streams(device, len(groups))
    for g, st in zip(groups, streams):
        st.wait_event(start)
        with torch.cuda.stream(st):
            run(g, ctypes.c_void_p(st.cuda_stream))
    for st in streams:
        done = torch.cuda.Event()
        done.record(st)
        main.wait_event(done)


def max_shot_groups() -> int:
    """Side streams that get a hardware queue of their own: HIP gives a
    process GPU_MAX_HW_QUEUES queues (4 by default) and the calling stream
    keeps one.  Streams beyond that share queues, and a shared queue runs
    its kernels in order: four groups on four side streams measured 93
    Gcell-updates/s on C2 against 150 for three."""
    try:
        q = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
    except ValueError:
        q = 4
    return max(1, q - 1)


_sweep_log = None


def release_graphs():
    """Free every captured graph the library holds.  Graphs are cached by
    their descriptor (device pointers and sizes), so a caller that keeps its
    buffers in place reuses them; one whose buffers move (a sweep over many
    shapes) accumulates them and can call this between calls.  Waits for all
    device work first, so no released graph is still running."""
    if torch.cuda.is_available():
        torch.cuda.synchronize()
    _native.check(_native.lib.fgpr_release_graphs(), "release graphs")


def graph_count() -> int:
    """Captured graphs the library's cache holds."""
    return int(_native.lib.fgpr_graph_count())


def set_graph_limit(n: int) -> int:
    """Bound the library's graph cache (least recently launched graphs are
    destroyed beyond it); returns the previous bound."""
    prev = int(_native.lib.fgpr_set_graph_limit(int(n)))
    if prev < 0:
        _native.check(prev, "set graph limit")
    return prev


def record_sweeps(on: bool):
    """Start recording a pair of HIP events on the calling stream around
    every full-history sweep (forward and reference adjoint: from before the
    fan-out to after the join), or stop and return the records
    [(kind, start, end), ...].  Measurement only."""
    global _sweep_log
    if on:
        _sweep_log = []
        return None
    out, _sweep_log = _sweep_log, None
    return out


def _sweep_begin(device):
    if _sweep_log is None:
        return None
    ev = torch.cuda.Event(enable_timing=True)
    ev.record(torch.cuda.current_stream(device))
    return ev


def _sweep_end(device, kind, start):
    if start is None or _sweep_log is None:
        return
    ev = torch.cuda.Event(enable_timing=True)
    ev.record(torch.cuda.current_stream(device))
    _sweep_log.append((kind, start, ev))


def shot_groups(G) -> int:
    """Shot groups of a full-history time-blocked sweep (option
    ``shot_streams``, default 3; 1 = one batch on the current stream),
    at most one per spare hardware queue and one per shot."""
    n = int(G.options.get("shot_streams", 3) or 1)
    if not G.time_block or G.options.get("gradient", "reference") == "exact":
        return 1
    return max(1, min(n, G.S, max_shot_groups()))


TB_GRAPH_LAUNCHES = 80


def sweep_chunk(G) -> int:
    """Steps per captured graph of a full-history sweep: ``graph_chunk`` when
    the caller set it or the per-step kernels run, else 80 time-blocked
    launches per graph (measured on C2: 16 launches 148, 32 150, 80 151-152,
    160-500 145-150 Gcell-updates/s)."""
    if G.options.get("graph_chunk_set") or not G.time_block or G.time_block == "auto":
        return G.graph_chunk
    return TB_GRAPH_LAUNCHES * int(G.time_block[2])


def source_scale(G) -> torch.Tensor:
    """w[n]*dl/(dx*dy*dz) per source row in float32, the expression of
    sourcereceiver.cu:73-75 with dl = dy = dz = dx (compute.py:571, 693)."""
    d = torch.tensor(G.dx32, dtype=torch.float32)
    w = G.srcwaves.to(torch.float32)
    return ((w * d) / ((d * d) * d)).contiguous()


def emitting_sources(G, b: Batch) -> torch.Tensor:
    """Device injection amounts of the emitting source rows; drops rows whose
    waveform is identically zero from ``b`` (they inject exactly nothing:
    Ez - uE4*0 == Ez)."""
    scale = source_scale(G)
    emit = (scale != 0).any(dim=1)
    if not bool(emit.all()):
        keep = to_device(emit.nonzero().flatten(), b.src.device)
        b.src = b.src.index_select(1, keep).contiguous()
        b.nsrc = int(b.src.shape[1])
        scale = scale[emit]
    scale = to_device(scale, G.device)
    b.srcscale = scale
    return scale


def forward(G, er, se, mr, store: bool = True):
    """Forward sweep over all nt steps.  Returns (traces (S, nt, nrx), batch)."""
    with _trace.span(f"fgpr forward {_trace.grid_tag(G)}"):
        return _forward(G, er, se, mr, store)


def _forward(G, er, se, mr, store):
    dev = G.device
    se_f = torch.clamp(se.detach(), min=0, max=1e6)          # compute.py:729
    coef = cell_coeffs(er, se_f, mr, G.dt32, G.dx32, False, dev)
    n = shot_groups(G)
    if n > 1:
        b = ShotGroups(G, store, G.time_block, n)
        probs = [g.problem(coef, srcscale=emitting_sources(G, g)) for g in b.groups]

        def run(g, st):
            check(lib.fgpr_forward(ctypes.byref(probs[b.groups.index(g)]), 0, g.nt,
                                   _MODES[G.exec_mode], sweep_chunk(G), st), "forward")
        t0 = _sweep_begin(dev)
        _fan_out(dev, b.groups, run)
        _sweep_end(dev, "forward", t0)
        return b.traces, b
    b = Batch(G, store=store, tb=G.time_block)
    scale = emitting_sources(G, b)
    p = b.problem(coef, srcscale=scale)
    t0 = _sweep_begin(dev)
    check(lib.fgpr_forward(ctypes.byref(p), 0, b.nt, _MODES[G.exec_mode], sweep_chunk(G),
                           _stream(dev)), "forward")
    _sweep_end(dev, "forward", t0)
    return b.traces, b


EPS0_KERNEL = 8.8541878128e-12   # the coefficient kernel's float constant (uc.cu:8)


def adjoint_exact(G, b: Batch, er, se, mr, grad_output, want_er=True, want_se=True):
    """Exact gradient dL/d er, dL/d se of the FP32 forward map (transpose of
    the forward stepping), unnormalised, float32 (NI, NJ, 1)."""
    with _trace.span(f"fgpr exact adjoint {_trace.grid_tag(G)}"):
        return _adjoint_exact(G, b, er, se, mr, grad_output, want_er, want_se)


def _adjoint_exact(G, b, er, se, mr, grad_output, want_er, want_se):
    dev = G.device
    se_f = torch.clamp(se.detach(), min=0, max=1e6)
    coef = cell_coeffs(er, se_f, mr, G.dt32, G.dx32, False, dev)
    b.reset_state()
    resid = grad_output.detach().to(device=dev, dtype=torch.float32).contiguous()
    if resid.shape != (b.S, b.nt, b.nrx):
        raise ValueError(f"grad_output shape {tuple(resid.shape)} != {(b.S, b.nt, b.nrx)}")
    NI, NJ = b.NI, b.NJ
    ge = torch.zeros((NI, NJ), dtype=torch.float64, device=dev)
    gs = torch.zeros((NI, NJ), dtype=torch.float64, device=dev)
    scratch = torch.empty((4, b.S, NI + 1, b.pitch), dtype=torch.float32, device=dev)
    p = b.problem(coef, resid=resid, grad_eps=ge, grad_sig=gs)
    p.tb_k = 0
    p.scratch = scratch.data_ptr()
    check(lib.fgpr_adjoint(ctypes.byref(p), b.nt - 1, 0, _native.GRAD_EXACT,
                           _MODES[G.exec_mode], G.graph_chunk, _stream(dev)), "exact adjoint")
    return finish_exact(G, ge, gs, want_er, want_se)


def finish_exact(G, ge, gs, want_er, want_se):
    """dL/d er = -(e0/dt) * sum uE4*lambda*dE, dL/d se = the sigma sum, as
    float32 (NI, NJ, 1); plus the raw FP64 sums."""
    dev = G.device
    NI, NJ = ge.shape
    scale_e = -float(np.float32(EPS0_KERNEL)) / G.dt32
    out = []
    for want, g, sc in ((want_er, ge, scale_e), (want_se, gs, 1.0)):
        if not want:
            out.append(None)
            continue
        o = torch.empty((NI, NJ, 1), dtype=torch.float32, device=dev)
        check(lib.fgpr_scale(_ptr(g), NI * NJ, sc, _ptr(o), _stream(dev)), "scale")
        out.append(o)
    return out[0], out[1], ge, gs


def adjoint(G, b: Batch, er, se, mr, grad_output, want_er=True, want_se=True):
    """Reference-mode adjoint sweep (compute.py:1020-1180).

    Returns (grad_er, grad_se) as float32 (NI, NJ, 1) normalised by their L2
    norms, None where not requested, plus the raw FP64 sums.
    """
    with _trace.span(f"fgpr adjoint {_trace.grid_tag(G)}"):
        return _adjoint(G, b, er, se, mr, grad_output, want_er, want_se)


def _adjoint(G, b, er, se, mr, grad_output, want_er, want_se):
    dev = G.device
    coef = cell_coeffs(er, se, mr, G.dt32, G.dx32, True, dev)      # Ucgeta
    b.reset_state()
    resid = grad_output.detach().to(device=dev, dtype=torch.float32).contiguous()
    if resid.shape != (b.S, b.nt, b.nrx):
        raise ValueError(f"grad_output shape {tuple(resid.shape)} != {(b.S, b.nt, b.nrx)}")
    NI, NJ = b.NI, b.NJ
    ge = torch.zeros((NI, NJ), dtype=torch.float64, device=dev)
    gs = torch.zeros((NI, NJ), dtype=torch.float64, device=dev)
    if isinstance(b, ShotGroups):
        b.gshot_buffer()

        # the shot reduction is deferred (null gradients) to one pass below
        probs = [g.problem(coef, srcscale=getattr(g, "srcscale", None), resid=resid[g.lo:g.hi])
                 for g in b.groups]

        def run(g, st):
            check(lib.fgpr_adjoint(ctypes.byref(probs[b.groups.index(g)]), g.nt - 1, 1,
                                   _native.GRAD_REFERENCE, _MODES[G.exec_mode], sweep_chunk(G), st),
                  "adjoint")
        if b.nt >= 2:
            t0 = _sweep_begin(dev)
            _fan_out(dev, b.groups, run)
            _sweep_end(dev, "adjoint", t0)
        p = b.groups[0].problem(coef, grad_eps=ge, grad_sig=gs)
        p.S = b.S
        p.gshot = b.gshot.data_ptr()
        check(lib.fgpr_reduce_shots(ctypes.byref(p), _stream(dev)), "reduce shots")
        return finish_reference(G, ge, gs, want_er, want_se)
    if b.tb:
        b.gshot_buffer()
    p = b.problem(coef, resid=resid, grad_eps=ge, grad_sig=gs)
    if b.nt >= 2:
        t0 = _sweep_begin(dev)
        check(lib.fgpr_adjoint(ctypes.byref(p), b.nt - 1, 1, _native.GRAD_REFERENCE,
                               _MODES[G.exec_mode], sweep_chunk(G), _stream(dev)), "adjoint")
        _sweep_end(dev, "adjoint", t0)
    return finish_reference(G, ge, gs, want_er, want_se)


def finish_reference(G, ge, gs, want_er, want_se):
    """g / ||g||_2 per requested parameter, float32 (NI, NJ, 1)
    (compute.py:1174-1180); plus the raw FP64 sums."""
    dev = G.device
    NI, NJ = ge.shape
    out = []
    for want, g in ((want_er, ge), (want_se, gs)):
        if not want:
            out.append(None)
            continue
        o = torch.empty((NI, NJ, 1), dtype=torch.float32, device=dev)
        check(lib.fgpr_normalise(_ptr(g), NI * NJ, 1.0, _ptr(o), _stream(dev)), "normalise")
        out.append(o)
    return out[0], out[1], ge, gs
