"""The inversion loop of the reference's driver (example.ipynb, cell 1):
two-phase Adam (ε alone, then ε and σ), L1 trace misfit x 100, gradients
zeroed on the first and last ``pmlthick`` columns, ε clamped at 1 and σ to
[0, 0.006] after every step, and the per-epoch frequency schedule with the
multiscale low-pass of the observed traces.

As in the notebook, the filtered observed traces are computed for the
scheduled frequency but the misfit is taken against the unfiltered ones
unless ``use_filtered=True``; ``freq`` is passed to ``compute``, which (like
the reference) does not use it.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import torch

from .compute import compute
from .multiscale import apply_filter


@dataclass
class InversionConfig:
    n_epochs: int = 202
    phase1_epochs: int = 50          # epochs updating ε only
    lr_er_phase1: float = 0.2
    lr_er_phase2: float = 0.1
    lr_se_phase2: float = 1e-4
    pmlthick: int = 10               # gradient columns zeroed at both ends of axis 1
    er_min: float = 1.0
    se_min: float = 0.0
    se_max: float = 0.006
    loss_scale: float = 100.0        # L1 misfit x 100
    target_freq: list = field(default_factory=list)   # empty: the source frequency throughout
    change_freq: int = 1000          # epochs per scheduled frequency
    use_filtered: bool = False       # misfit against the low-passed observed traces


def invert(compute_args: tuple, compute_kwargs: dict, er0: torch.Tensor, se0: torch.Tensor,
           mr: torch.Tensor, observed: torch.Tensor, cfg: InversionConfig | None = None,
           callback=None) -> dict:
    """Run the notebook's inversion loop.

    ``compute_args`` = (device, dx, ny, nx, nz, source, source_amplitudes,
    step) and ``compute_kwargs`` = the keyword arguments of ``compute`` after
    ``er, se, mr, nsrc, nrx`` (passed as ``compute_kwargs['nsrc']`` /
    ``['nrx']``).  ``callback(epoch, er, se, loss)`` runs after each step.
    Returns the final models, the per-epoch losses and timings."""
    cfg = cfg or InversionConfig()
    kw = dict(compute_kwargs)
    nsrc, nrx = kw.pop("nsrc"), kw.pop("nrx")
    source_amplitudes = compute_args[6]
    dt = kw.get("dt")
    er = er0.detach().clone()
    se = se0.detach().clone()
    opt1 = torch.optim.Adam([{"params": er, "lr": cfg.lr_er_phase1}])
    opt2 = torch.optim.Adam([{"params": er, "lr": cfg.lr_er_phase2},
                             {"params": se, "lr": cfg.lr_se_phase2}])
    l1 = torch.nn.L1Loss()
    losses, times = [], []
    P = cfg.pmlthick
    for epoch in range(cfg.n_epochs):
        t0 = time.time()
        if epoch < cfg.phase1_epochs:
            opt = opt1
            se.requires_grad = False
            er.requires_grad = True
        else:
            opt = opt2
            se.requires_grad = True
            er.requires_grad = True
        freq, obs = source_amplitudes[2], observed
        tf = cfg.target_freq
        if tf:
            n = epoch // cfg.change_freq
            if not (n >= len(tf) or len(tf) == 1):
                freq = tf[n]
                filtered = apply_filter(observed.detach(), 1 / dt, freq)
                if cfg.use_filtered:
                    obs = filtered
        opt.zero_grad()
        result, _ = compute(*compute_args, er, se, mr, nsrc, nrx, freq=freq, **kw)
        loss = l1(result, obs) * cfg.loss_scale
        loss.backward()
        for p in (er, se):
            if p.requires_grad and p.grad is not None and P > 0:
                p.grad[:, :P, :] = 0
                p.grad[:, -P:, :] = 0
        opt.step()
        losses.append(loss.item())
        with torch.no_grad():
            se.data.clamp_(min=cfg.se_min, max=cfg.se_max)
            er.data.clamp_(min=cfg.er_min)
        times.append(time.time() - t0)
        if callback is not None:
            callback(epoch, er, se, losses[-1])
    return {"er": er.detach(), "se": se.detach(), "losses": losses, "epoch_seconds": times}
