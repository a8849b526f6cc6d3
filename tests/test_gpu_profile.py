"""fgpr_profile_kernel's graph replay (FGPR_PROFILE_GRAPH), which the
autotuner times: the captured launches must do exactly the work of the
same launches issued one by one, on the first call (capture + run) and on
the replays after it, so the tuner ranks the work the sweeps do."""
import ctypes
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grid(wl, dev, er, se, mr):
    import torch
    import fastgpr
    return fastgpr.initialization(dev, wl.dx, wl.NJ, wl.NI, 1, wl.time_window, ['hertzian', 'z'],
                                  None, None, None, None, wl.P, ['ricker', 1, wl.freq], wl.S,
                                  er, se, mr, 1, wl.rx.shape[1], wl.dt, 'HORIPML', None, None,
                                  False, False, None, torch.tensor(wl.src), torch.tensor(wl.rx))


@pytest.mark.parametrize("tb", [(64, 1024, 6), (32, 512, 4)])
def test_graph_replayed_blocks_equal_eager_blocks(hip_device, tb):
    import torch
    from fastgpr import _native, solver
    from fastgpr.workloads import small_case
    from helpers import product_tensors
    wl = small_case(nt=120)
    er, se, mr = product_tensors(wl, hip_device)
    G = dataclasses.replace(_grid(wl, hip_device, er, se, mr), time_block=tb)
    K = tb[2]
    print("case", wl.name, "S", wl.S, "variant", tb, flush=True)
    coef = solver.cell_coeffs(er, torch.clamp(se, min=0), mr, G.dt32, G.dx32, False, hip_device)
    rng = np.random.default_rng(3)
    resid = torch.tensor(rng.standard_normal((wl.S, G.iterations, wl.rx.shape[1])),
                         dtype=torch.float32, device=hip_device)
    st = ctypes.c_void_p(torch.cuda.current_stream(hip_device).cuda_stream)
    states = []
    for flag in (0, _native.PROFILE_GRAPH):
        b = solver.Batch(G, store=True, tb=tb)
        b.gshot_buffer()
        ge = torch.zeros((G.NI, G.NJ), dtype=torch.float64, device=hip_device)
        gs = torch.zeros_like(ge)
        p = b.problem(coef, srcscale=solver.emitting_sources(G, b), resid=resid,
                      grad_eps=ge, grad_sig=gs)
        b.snaps.zero_()
        for _ in range(2):            # graph mode: capture + run, then a replay
            for kid, n in ((_native.KERNEL_TB_FWD, K), (_native.KERNEL_TB_ADJ, 2 * K)):
                _native.check(_native.lib.fgpr_profile_kernel(ctypes.byref(p), kid | flag, n, 3, st),
                              "profile")
        torch.cuda.synchronize(hip_device)
        states.append([t.clone() for t in (b.fields, b.fields2, b.psi, b.psi2, b.gshot, b.snaps,
                                           b.traces)])
    eager, graph = states
    assert float(eager[0].abs().max()) > 0 and float(eager[4].abs().max()) > 0
    for a, g in zip(eager, graph):
        assert torch.equal(a, g)
    solver.release_graphs()
