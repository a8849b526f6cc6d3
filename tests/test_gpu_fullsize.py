"""Parity at full size on the cross-hole batch (400 x 200 cells at 0.025 m,
20-cell CPML, 3000 steps, 13 shots, 80 receivers), run through the product
configuration the benchmark times: time-blocked (64, 1024, 6) kernels,
three shot streams, captured graphs.  The checker is the C restatement of
the oracle on the host cores.  Traces must lie within 1e-5 relative L2 of
its FP32 mode, which is bitwise the numpy float32 oracle.  The gradient at the
smoothed initial model, for the residual (initial - true traces), must lie
within 1e-4 of its FP64 reference-mode gradient."""
import time

import numpy as np
import pytest

from helpers import oracle_problem, rel_l2, run_product
from oracle import c_oracle as CO
from oracle import fdtd_oracle as O

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (CO.available() and CO.available(np.float32)),
                                 reason="oracle/_build/liboracle*.so not built (make -C oracle)")]

BENCH = dict(time_block=(64, 1024, 6), shot_streams=3, graph_chunk=None)


def test_c2_crosshole_full_size_traces_and_gradient(hip_device):
    from fastgpr.workloads import c2_crosshole
    wl = c2_crosshole()
    print("case", wl.name, "threads", CO.threads(), flush=True)
    obs, _ = run_product(wl, hip_device, **BENCH)

    t0 = time.perf_counter()
    pb = oracle_problem(wl)
    f = np.float32
    c, pml = O.setup(pb, f)
    tr32, _ = CO.forward(c, pml, pb.P, pb.dx, pb.src, O.source_scale(pb.wavelet, pb.dx, f),
                         pb.rx, pb.nt, store=False, dtype=f)
    e32 = rel_l2(obs, tr32)
    print(f"traces rel-L2 vs FP32 oracle {e32:.3e} ({time.perf_counter() - t0:.1f} s)", flush=True)
    assert np.abs(tr32).max() > 0
    assert e32 <= 1e-5

    init = (wl.er0, wl.se0)
    pred, _ = run_product(wl, hip_device, model=init, **BENCH)
    resid = (pred - obs).astype(np.float32)
    _, g = run_product(wl, hip_device, resid=resid, model=init, **BENCH)

    t0 = time.perf_counter()
    pb0 = oracle_problem(wl, model=init)
    c, pml = O.setup(pb0, np.float64)
    _, snaps = CO.forward(c, pml, pb0.P, pb0.dx, pb0.src,
                          O.source_scale(pb0.wavelet, pb0.dx, np.float64), pb0.rx, pb0.nt, True)
    cadj, _ = O.setup(pb0, np.float64, adjoint=True)
    ge, gs = CO.adjoint_reference(cadj, pml, pb0.P, pb0.dx, pb0.rx,
                                  resid.astype(np.float64), snaps)
    del snaps
    print(f"FP64 oracle gradient {time.perf_counter() - t0:.1f} s", flush=True)
    for k, ref in (("er", O.normalise(ge)), ("se", O.normalise(gs))):
        e = rel_l2(g[k], ref)
        print(f"gradient {k} rel-L2 vs FP64 oracle {e:.3e}", flush=True)
        assert e <= 1e-4, k
