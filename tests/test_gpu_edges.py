"""Edge cases of the HIP path (GPU): no CPML, receivers and sources on the
domain's first and last rows and columns and in its corners, colliding
receivers, a dead (out-of-plane) source, odd grid sizes whose last tiles are
ragged.  The time-blocked variants must reproduce the per-step kernels
bitwise (traces) and the oracle within the usual tolerances."""
import numpy as np
import pytest

from helpers import rel_l2, run_oracle, run_product

pytestmark = pytest.mark.gpu

VARIANTS = [(64, 1024, 6), (32, 512, 4), (64, 512, 3)]


def _edge_workload(P):
    from fastgpr.workloads import Workload, small_case
    wl = small_case(NI=67, NJ=53, P=P, S=2, nt=131, nrx=6, seed=21)
    NI, NJ = wl.NI, wl.NJ
    rx = np.zeros((2, 6, 3), np.int32)
    rx[0, :, :2] = [(0, 0), (0, NJ - 1), (NI - 1, 0), (NI - 1, NJ - 1), (0, 17), (33, 0)]
    rx[1, :, :2] = [(NI - 1, 20), (10, NJ - 1), (1, 1), (NI - 2, NJ - 2), (1, 1), (1, 1)]
    src = np.zeros((2, 2, 3), np.int32)
    src[0, 0, :2] = (1, 1)                          # next to the corner
    src[0, 1] = (30, 30, 1)                         # k != 0: dead (compute.py:790-796)
    src[1, 0, :2] = (NI - 2, NJ // 2)
    src[1, 1, :2] = (5, NJ - 2)
    return Workload(wl.name + f"_edges_P{P}", wl.er, wl.se, wl.mr, wl.er0, wl.se0, wl.dx, wl.dt,
                    wl.nt, P, wl.freq, src, rx)


@pytest.mark.parametrize("P", [0, 5])
def test_edges_time_blocked_match_per_step_and_oracle(hip_device, P):
    wl = _edge_workload(P)
    print("case", wl.name, flush=True)
    resid = np.random.default_rng(P + 1).standard_normal((wl.S, wl.nt, 6)).astype(np.float32)
    a, ga = run_product(wl, hip_device, resid=resid)
    o32, _ = run_oracle(wl, np.float32)
    _, o64 = run_oracle(wl, np.float64, resid=resid)
    assert np.abs(o32).max() > 0
    assert rel_l2(a, o32) <= 1e-5
    for k in ("er", "se"):
        assert rel_l2(ga[k], o64[k]) <= 1e-4, k
    for tb in VARIANTS:
        print("variant", tb, flush=True)
        b, gb = run_product(wl, hip_device, resid=resid, time_block=tb, graph_chunk=12)
        assert np.array_equal(a, b), tb
        for k in ("er", "se"):
            assert rel_l2(gb[k], ga[k]) < 1e-12, (tb, k)


def test_colliding_receivers_give_reproducible_gradients(hip_device):
    """Three receivers on one cell (shot 1) and a colliding pair: the
    per-step adjoint applies their residuals by one thread in list order, so
    two runs give bitwise the same gradients (no dependence on the order in
    which threads reach the cell)."""
    wl = _edge_workload(5)
    resid = np.random.default_rng(9).standard_normal((wl.S, wl.nt, 6)).astype(np.float32)
    _, g1 = run_product(wl, hip_device, resid=resid)
    _, g2 = run_product(wl, hip_device, resid=resid)
    for k in ("er", "se"):
        assert np.array_equal(g1[k], g2[k]), k
