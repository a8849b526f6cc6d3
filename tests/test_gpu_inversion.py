"""The notebook's inversion loop (fastgpr.inversion, example.ipynb cell 1)
on a small crosswell problem: the misfit falls, σ stays fixed through phase
1, the zeroed gradient columns never move, clamps hold, and the scheduled
low-pass runs on the observed traces."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_inversion_loop_behaviour(hip_device):
    from fastgpr.compute import compute
    from fastgpr.inversion import InversionConfig, invert
    from fastgpr.workloads import small_case
    dev = hip_device
    wl = small_case(NI=64, NJ=48, P=8, S=2, nt=320, nrx=6, seed=11)
    t = lambda a: torch.tensor(np.asarray(a, np.float32))[..., None].to(dev)
    args = (dev, wl.dx, wl.NJ, wl.NI, 1, ['hertzian', 'z'], ['ricker', 1, wl.freq], wl.S)
    kw = dict(nsrc=wl.src.shape[1], nrx=wl.rx.shape[1], time_windows=wl.time_window, dt=wl.dt,
              pmlthick=wl.P, customsource=torch.tensor(wl.src), customreceiver=torch.tensor(wl.rx),
              tv=True, time_block="auto")
    with torch.no_grad():
        obs, _ = compute(*args, t(wl.er), t(wl.se), t(wl.mr), kw["nsrc"], kw["nrx"],
                         **{k: v for k, v in kw.items() if k not in ("nsrc", "nrx")})
    er0 = t(np.maximum(wl.er * 0.9, 1.0))
    se0 = t(np.clip(wl.se * 0.5, 0, 0.006))
    seen = []
    cfg = InversionConfig(n_epochs=8, phase1_epochs=3, pmlthick=4, target_freq=[wl.freq / 2, wl.freq],
                          change_freq=4)
    out = invert(args, kw, er0, se0, t(wl.mr), obs, cfg,
                 callback=lambda e, er, se, l: seen.append((er.detach().clone(), se.detach().clone())))
    L = out["losses"]
    assert len(L) == 8 and np.isfinite(L).all()
    assert min(L[1:]) < L[0]
    for e in range(3):                                   # phase 1: ε only
        assert torch.equal(seen[e][1], se0)
    assert not torch.equal(seen[-1][1], se0)             # phase 2 moves σ
    P = cfg.pmlthick
    for er, se in seen:                                  # zeroed gradient columns never move
        assert torch.equal(er[:, :P], er0[:, :P]) and torch.equal(er[:, -P:], er0[:, -P:])
        assert float(se.min()) >= 0.0 and float(se.max()) <= np.float32(0.006) and float(er.min()) >= 1.0
