"""Drop-in replacement of the reference's ``src/compute.py`` call surface.

``compute`` keeps the reference signature (compute.py:33-37) and return value
``(traces (S, nt, nrx) float32, dt)``; ``GPRFWI`` is the autograd Function
(compute.py:877-942) whose backward returns the normalised reference
gradients (compute.py:1174-1180) plus the optional TV term
(compute.py:916-933).  The internals run on libfastgpr (HIP, gfx950).

Extra keyword arguments (all default to the reference behaviour):

* ``exec_mode``   -- "graph" (time loop replayed from captured HIP graphs) or
                     "eager" (one kernel sequence per step from the host loop).
* ``graph_chunk`` -- steps per captured graph (default: 32 for the per-step
                     kernels and checkpointed segments, 80 launches for a
                     full-history time-blocked sweep).
* ``time_block``  -- "auto" (the default: the autotuner picks a time-blocked
                     variant for this grid; where no variant holds the
                     geometry, a warning and the per-step kernels), None (one
                     H and one E kernel per step) or an explicit (region
                     rows, threads, steps per launch).  Traces are bitwise the
                     same every way; gradients agree to ~1e-12 (FP64 sums
                     blocked differently).
* ``gradient``    -- "reference" (the reference's normalised approximate
                     gradient) or "exact" (dL/d er, dL/d se of the discrete
                     forward map through its exact adjoint, unnormalised).
* ``checkpoint``  -- None (store every snapshot, as the reference), an
                     interval C (keep the last C + 1 snapshots, save the
                     solver state every C steps before them and recompute
                     those snapshots segment by segment in the adjoint), or
                     "auto" (the default: store every snapshot, as the
                     reference, unless the history would not fit in the free
                     device memory less a reserve for the rest of the
                     gradient; then keep as many of the last snapshots as
                     fit and recompute only the earlier ones).  Checkpointed
                     gradients are bitwise the stored ones.
* ``shot_streams`` -- shot groups of a time-blocked sweep, each on its own
                     HIP stream so that the launches of one fill the CUs the
                     other's launch tail leaves idle (default 3; 1 = one
                     batch on the current stream).  Traces are bitwise those
                     of one batch; the gradient is reduced over all shots in
                     shot order either way.
"""
from __future__ import annotations

import atexit
import time
import warnings

import torch
from torch import Tensor

from . import checkpoint as _ckpt
from . import grid as _grid
from . import solver as _solver
from .grid import FDTDGrid


# The reference prints a warning when traces, grad_output or a gradient hold a
# NaN (compute.py:898-899, 905-906, 924-925, 931-932).  Testing on the host
# at that point would stall the launch queue (the host waits for the GPU
# before it queues the adjoint); the tests here are queued on the device,
# their flags copied to pinned host memory behind an event, and the warnings
# printed once the flags have arrived: at the end of backward when they
# already have, else at the next call (or at exit).
_nan_pending: list = []


def _watch_nan(*pairs):
    """Queue NaN tests of (tensor, message) pairs without a host sync."""
    pairs = [(t, msg) for t, msg in pairs if t is not None]
    if not pairs:
        return
    dev = pairs[0][0].device
    if dev.type != "cuda":
        for t, msg in pairs:
            if torch.isnan(t).any():
                print(msg)
        return
    flags = torch.stack([torch.isnan(t).any().to(dev) for t, _ in pairs])
    host = torch.empty(flags.shape, dtype=torch.bool, pin_memory=True)
    host.copy_(flags, non_blocking=True)
    ev = torch.cuda.Event()
    ev.record(torch.cuda.current_stream(dev))
    _nan_pending.append((ev, host, [msg for _, msg in pairs]))


def report_nan_warnings(block: bool = False):
    """Print the warnings of every NaN test whose flags have arrived (all of
    them, waiting, when ``block``)."""
    keep = []
    for ev, host, msgs in _nan_pending:
        if block or ev.query():
            ev.synchronize()
            for flag, msg in zip(host.tolist(), msgs):
                if flag:
                    print(msg)
        else:
            keep.append((ev, host, msgs))
    _nan_pending[:] = keep


atexit.register(lambda: report_nan_warnings(block=True) if torch.cuda.is_initialized() else None)


class GPRFWIModel(torch.nn.Module):
    """nn.Module holding er / se as parameters (compute.py:16-29).  Its
    forward calls ``compute`` with keyword arguments, which fixes the
    reference's positional mix-up of ``time_windows`` into ``source``."""

    def __init__(self, se: Tensor, er: Tensor, serequires_grad: bool = False,
                 errequires_grad: bool = False):
        super().__init__()
        self.er = torch.nn.Parameter(er, requires_grad=errequires_grad)
        self.se = torch.nn.Parameter(se, requires_grad=serequires_grad)

    def forward(self, device, dx, ny, nx, nz, time_windows, source, source_location, source_step,
                receiver_location, receiver_step, pmlthick, source_apmlitudes, step, mr, nsrc, nrx,
                freq=None, dt=None, pml_formulation='HORIPML', total_step=None, tv=None):
        return compute(device, dx, ny, nx, nz, source, source_apmlitudes, step, self.er, self.se,
                       mr, nsrc, nrx, time_windows=time_windows, source_location=source_location,
                       source_step=source_step, receiver_location=receiver_location,
                       receiver_step=receiver_step, pmlthick=pmlthick, freq=freq, dt=dt,
                       pml_formulation=pml_formulation, total_step=total_step, tv=tv)


def compute(device, dx, ny, nx, nz, source, source_apmlitudes, step, er, se, mr, nsrc, nrx,
            time_windows=None, source_location=None, source_step=None,
            receiver_location=None, receiver_step=None, pmlthick=0, freq=None,
            dt=None, pml_formulation='HORIPML', total_step=None, tv=True,
            Fixedreceiver=False, nseg=None, customsource=None, customreceiver=None,
            exec_mode="graph", graph_chunk=None, time_block="auto", gradient="reference",
            checkpoint="auto", shot_streams=3):
    """Forward-model ``step`` shots; returns (traces (step, nt, nrx), dt).

    Same arguments and meaning as compute.py:33-50: ``nx`` is the depth
    (first axis) and ``ny`` the lateral size, ``step`` the number of shots,
    er / se / mr are (nx, ny, 1) tensors, locations and steps are metres.
    se is clamped at 0 and er at 1 (differentiably) before the solve.
    """
    report_nan_warnings()
    se = torch.clamp(se, min=0)
    er = torch.clamp(er, min=1)
    G = initialization(device, dx, ny, nx, nz, time_windows, source, source_location,
                       source_step, receiver_location, receiver_step, pmlthick,
                       source_apmlitudes, step, er, se, mr, nsrc, nrx, dt, pml_formulation,
                       freq, total_step, tv, Fixedreceiver, nseg, customsource, customreceiver,
                       exec_mode=exec_mode, graph_chunk=graph_chunk, time_block=time_block,
                       gradient=gradient, checkpoint=checkpoint, shot_streams=shot_streams)
    if G.time_block == "auto":
        from .autotune import NoVariant, choose
        try:
            G.time_block = choose(G, er, se, mr)
        except NoVariant as e:
            warnings.warn(f"{e}; using the per-step kernels", RuntimeWarning, stacklevel=2)
            G.time_block = None
    rxs = GPRFWI.apply(G, er, se, mr)
    return rxs, G.dt


def initialization(device, dx, ny, nx, nz, time_windows, source, source_location, source_step,
                   receiver_location, receiver_step, pmlthick, source_apmlitudes, step, er, se,
                   mr, nsrc, nrx, dt, pml_formulation, freq, total_step, tv, Fixedreceiver,
                   nseg, customsource=None, customreceiver=None, exec_mode="graph",
                   graph_chunk=None, time_block=None, gradient="reference",
                   checkpoint="auto", shot_streams=3) -> FDTDGrid:
    """Build the solve description (compute.py:726-872).

    ``nseg`` (the reference's unfinished segmenting, compute.py:543-553) and
    ``freq`` (whose filtered wavelet is discarded, compute.py:825-826) are
    accepted and have no effect, as in the reference.
    """
    device = _solver.require_hip(device)
    if pml_formulation != 'HORIPML':
        raise ValueError("only the HORIPML formulation is implemented (compute.py:193)")
    if source_apmlitudes[0] != 'ricker':
        raise ValueError("only the 'ricker' waveform is implemented (compute.py:393)")
    if source[1] != 'z':
        raise ValueError("2-D TMz supports z-polarised sources only")
    if exec_mode not in ("graph", "eager"):
        raise ValueError("exec_mode must be 'graph' or 'eager'")
    if gradient not in ("reference", "exact"):
        raise ValueError("gradient must be 'reference' or 'exact'")
    if int(shot_streams) < 1:
        raise ValueError("shot_streams must be >= 1")
    NI, NJ = int(nx), int(ny)
    if int(nz) != 1:
        raise ValueError("2-D TMz solver: nz must be 1")
    S = int(step)
    if tuple(er.shape[:2]) != (NI, NJ):
        raise ValueError(f"er shape {tuple(er.shape)} does not match (nx, ny) = {(NI, NJ)}")
    Gdt = _grid.resolve_dt(dx, dt)
    ts = S if total_step is None else int(total_step)
    src, rx = _grid.geometry(nsrc, nrx, ts, source_location, source_step, receiver_location,
                             receiver_step, dx, Fixedreceiver, customsource, customreceiver)
    if time_windows is None:
        raise ValueError("time_windows is required")
    tw = time_windows if isinstance(time_windows, torch.Tensor) else torch.tensor(time_windows)
    nt = _grid.n_iterations(tw, Gdt)
    amp, f0 = source_apmlitudes[1], source_apmlitudes[2]
    w = _grid.ricker(nt, Gdt, amp, f0, tw)
    srcwaves = torch.zeros((src.shape[1], nt), dtype=torch.float32)
    if src.shape[1] > 0:
        srcwaves[0] = w          # one dipole object: only row 0 emits (compute.py:563-564)
    P = int(pmlthick)
    pml = _grid.pml_profile(P, dx, Gdt, er.detach(), mr.detach()) if P > 0 else None
    return FDTDGrid(device=device, NI=NI, NJ=NJ, S=S, nsrc=int(src.shape[1]), nrx=int(rx.shape[1]),
                    dx=dx, dt=Gdt, iterations=nt, pmlthick=P, source=src, receiver=rx,
                    srcwaves=srcwaves, pml=pml, tv=tv, exec_mode=exec_mode,
                    graph_chunk=32 if graph_chunk is None else int(graph_chunk),
                    time_block=time_block if time_block in (None, "auto") else tuple(time_block),
                    options={"gradient": gradient, "checkpoint": checkpoint,
                             "shot_streams": int(shot_streams),
                             "graph_chunk_set": graph_chunk is not None})


class GPRFWI(torch.autograd.Function):
    """Forward returns the Ez traces (S, nt, nrx); backward returns the
    L2-normalised reference gradients for er / se (compute.py:877-942)."""

    @staticmethod
    def forward(ctx, G, er, se, mr):
        ctx.save_for_backward(er, se, mr)
        ctx.G = G
        need_grad = er.requires_grad or se.requires_grad
        hist = _ckpt.history(G, G.options.get("checkpoint")) if need_grad else None
        ctx.fw = None
        if hist is not None and not hist.full:
            traces, ctx.fw = _ckpt.forward(G, er, se, mr, hist)
            ctx.batch = None
        else:
            traces, batch = _solver.forward(G, er, se, mr, store=need_grad)
            ctx.batch = batch if need_grad else None
        _watch_nan((traces, "⚠️ rxs_gpu contains NaN!"))
        return traces

    @staticmethod
    def backward(ctx, grad_output):
        _watch_nan((grad_output, "⚠️ grad_output contains NaN!"))
        er, se, mr = ctx.saved_tensors
        G = ctx.G
        exact = G.options.get("gradient") == "exact"
        if ctx.fw is not None:
            g_er, g_se, _, _ = _ckpt.adjoint(G, ctx.fw, er, se, mr, grad_output,
                                             want_er=er.requires_grad, want_se=se.requires_grad,
                                             exact=exact)
        else:
            sweep = _solver.adjoint_exact if exact else _solver.adjoint
            g_er, g_se, _, _ = sweep(G, ctx.batch, er, se, mr, grad_output,
                                     want_er=er.requires_grad, want_se=se.requires_grad)
        ctx.batch = None
        ctx.fw = None
        if G.tv:
            from .tv import total_variation
            t0 = time.time()
            if er.requires_grad:
                uer = total_variation(er.clone())
                g_er += 0.01 * (2 * (er - uer))
                _watch_nan((g_er, "⚠️ epsilon.grad contains NaN!"))
            if se.requires_grad:
                use = total_variation(se.clone())
                g_se += 0.01 * (2 * (se - use))
                _watch_nan((g_se, "⚠️ sigma.grad contains NaN!"))
            print(f"Total variation regularization took {time.time() - t0:.4f} seconds.")
        report_nan_warnings()
        return None, g_er, g_se, None
