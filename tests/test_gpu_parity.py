"""Parity of the HIP path (through fastgpr.compute -> libfastgpr C ABI)
against the oracle on identical inputs.

Tolerances:
* traces: <= 1e-5 relative L2 against the reference FP32 result (the oracle's
  FP32 mode, which follows the reference's FP32 operation order);
* gradients: <= 1e-4 relative L2 against the FP64 oracle.
"""
import numpy as np
import pytest

from helpers import rel_l2, run_oracle, run_product

pytestmark = pytest.mark.gpu

TRACE_TOL_FP32 = 1e-5
TRACE_TOL_FP64 = 1e-4
GRAD_TOL_FP64 = 1e-4


def _resid(wl, seed=11):
    from helpers import host_grid
    nt, _ = host_grid(wl)
    rng = np.random.default_rng(seed)
    return rng.standard_normal((wl.S, nt, wl.rx.shape[1])).astype(np.float32)


def test_small_forward_matches_oracle(hip_device):
    from fastgpr.workloads import small_case
    wl = small_case()
    print("case", wl.name, "S", wl.S, "nt", wl.nt, flush=True)
    tr, _ = run_product(wl, hip_device)
    o32, _ = run_oracle(wl, np.float32)
    o64, _ = run_oracle(wl, np.float64)
    e32, e64 = rel_l2(tr, o32), rel_l2(tr, o64)
    print(f"trace rel-L2 vs fp32 oracle {e32:.3e}, vs fp64 {e64:.3e}")
    assert np.abs(o64).max() > 0
    assert e32 <= TRACE_TOL_FP32
    assert e64 <= TRACE_TOL_FP64


def test_graph_and_eager_are_bitwise_equal(hip_device):
    from fastgpr.workloads import small_case
    wl = small_case(nt=130)
    a, _ = run_product(wl, hip_device, exec_mode="graph", graph_chunk=16)
    b, _ = run_product(wl, hip_device, exec_mode="eager")
    assert np.array_equal(a, b)


def test_small_gradient_matches_oracle(hip_device):
    from fastgpr.workloads import small_case
    wl = small_case(nt=240)
    resid = _resid(wl)
    _, g = run_product(wl, hip_device, resid=resid)
    _, o = run_oracle(wl, np.float64, resid=resid)
    for k in ("er", "se"):
        e = rel_l2(g[k], o[k])
        print(f"grad {k} rel-L2 vs fp64 oracle {e:.3e}")
        assert np.isfinite(g[k]).all()
        assert abs(np.linalg.norm(g[k]) - 1.0) < 1e-5
        assert e <= GRAD_TOL_FP64


def test_single_parameter_gradient(hip_device):
    """Only er requires grad: se gets no gradient (compute.py:1177-1178)."""
    from fastgpr.workloads import small_case
    wl = small_case(nt=120)
    _, g = run_product(wl, hip_device, resid=_resid(wl), want=("er",))
    assert "se" not in g and "er" in g


def test_c1_traces_and_gradient(hip_device):
    """Full C1 configuration: 200x100, 20-cell CPML, 1500 steps; residual =
    traces of the smoothed initial model minus the true-model traces."""
    from fastgpr.workloads import c1_plain
    wl = c1_plain()
    obs, _ = run_product(wl, hip_device)
    o32, _ = run_oracle(wl, np.float32)
    e32 = rel_l2(obs, o32)
    print(f"C1 trace rel-L2 vs fp32 oracle {e32:.3e}")
    assert e32 <= TRACE_TOL_FP32
    init = (wl.er0, wl.se0)
    pred, _ = run_product(wl, hip_device, model=init)
    resid = (pred - obs).astype(np.float32)
    _, g = run_product(wl, hip_device, resid=resid, model=init)
    _, o = run_oracle(wl, np.float64, resid=resid, model=init)
    for k in ("er", "se"):
        e = rel_l2(g[k], o[k])
        print(f"C1 grad {k} rel-L2 vs fp64 oracle {e:.3e}")
        assert e <= GRAD_TOL_FP64


TB_VARIANTS = [(32, 256, 1), (32, 256, 4), (64, 512, 3), (64, 512, 8), (64, 256, 4), (32, 512, 4),
               (64, 1024, 6), (64, 1024, 8)]


@pytest.mark.parametrize("tb", TB_VARIANTS)
def test_time_blocked_matches_per_step_kernels(hip_device, tb):
    """The time-blocked kernels perform the same per-cell arithmetic in the
    same order as the per-step kernels: traces and gradients agree bitwise
    (graph replay with a ragged tail, nt not a multiple of the block)."""
    from fastgpr.workloads import small_case
    wl = small_case(NI=70, NJ=90, nt=157)
    resid = _resid(wl)
    print("variant", tb, flush=True)
    a, ga = run_product(wl, hip_device, resid=resid)
    b, gb = run_product(wl, hip_device, resid=resid, time_block=tb, graph_chunk=16)
    assert np.array_equal(a, b)
    for k in ("er", "se"):
        assert rel_l2(gb[k], ga[k]) < 1e-12


@pytest.mark.parametrize("tb", [(64, 1024, 6), (32, 512, 4), (64, 512, 3)])
def test_time_blocked_step_loop_copies_match_per_step_kernels(hip_device, tb):
    """200 x 160 cells with a 10-cell CPML: some regions have clean columns
    (no CPML column, no column 0) and some waves clean rows, so all four
    compiled copies of the step loop run (clean rows and columns, dirty
    rows only, dirty columns only, both dirty) in one launch; traces are
    bitwise those of the per-step kernels."""
    from fastgpr import _native
    from fastgpr.workloads import small_case
    wl = small_case(NI=200, NJ=160, P=10, S=2, nt=61, nrx=5, seed=8)
    RI, _, K = tb
    TJ = 64 - 2 * K - 1
    starts = [tj * TJ - K for tj in range(_native.tb_tiles(200, 160, *tb)[3])]
    assert any(g >= wl.P + 1 and g + 64 <= wl.NJ - wl.P for g in starts)     # clean columns
    assert any(g < wl.P + 1 or g + 64 > wl.NJ - wl.P for g in starts)        # CPML columns
    resid = _resid(wl)
    print("variant", tb, flush=True)
    a, ga = run_product(wl, hip_device, resid=resid)
    b, gb = run_product(wl, hip_device, resid=resid, time_block=tb, graph_chunk=8)
    assert np.array_equal(a, b)
    for k in ("er", "se"):
        assert rel_l2(gb[k], ga[k]) < 1e-12


@pytest.mark.parametrize("groups", [2, 3, 5])
def test_shot_groups_on_streams_are_bitwise_one_batch(hip_device, groups):
    """Shot groups swept on separate streams give the traces and the
    (shot-order reduced) gradients of one batch, bitwise."""
    from fastgpr.workloads import small_case
    wl = small_case(NI=70, NJ=90, nt=157, S=5)
    resid = _resid(wl)
    print("groups", groups, "shots", wl.S, flush=True)
    a, ga = run_product(wl, hip_device, resid=resid, time_block=(64, 1024, 6), shot_streams=1)
    b, gb = run_product(wl, hip_device, resid=resid, time_block=(64, 1024, 6), shot_streams=groups)
    assert np.array_equal(a, b)
    assert np.array_equal(ga["er"], gb["er"]) and np.array_equal(ga["se"], gb["se"])


def test_released_graphs_are_captured_again(hip_device):
    """fastgpr.release_graphs() frees the cached graphs; the next call
    captures new ones and gives the same traces and gradients bitwise."""
    import fastgpr
    from fastgpr.workloads import small_case
    wl = small_case(nt=90)
    resid = _resid(wl)
    a, ga = run_product(wl, hip_device, resid=resid, time_block=(64, 1024, 6))
    fastgpr.release_graphs()
    b, gb = run_product(wl, hip_device, resid=resid, time_block=(64, 1024, 6))
    assert np.array_equal(a, b)
    for k in ("er", "se"):
        assert np.array_equal(ga[k], gb[k])


def test_time_blocked_eager_equals_graph(hip_device):
    from fastgpr.workloads import small_case
    wl = small_case(NI=50, NJ=130, nt=90)
    resid = _resid(wl)
    a, ga = run_product(wl, hip_device, resid=resid, time_block=(64, 512, 4), exec_mode="eager")
    b, gb = run_product(wl, hip_device, resid=resid, time_block=(64, 512, 4), graph_chunk=24)
    assert np.array_equal(a, b)
    assert np.array_equal(ga["er"], gb["er"]) and np.array_equal(ga["se"], gb["se"])


def test_autotuned_c1_gradient_matches_oracle(hip_device):
    from fastgpr.workloads import c1_plain
    wl = c1_plain()
    init = (wl.er0, wl.se0)
    obs, _ = run_product(wl, hip_device, time_block="auto")
    pred, _ = run_product(wl, hip_device, model=init, time_block="auto")
    resid = (pred - obs).astype(np.float32)
    _, g = run_product(wl, hip_device, resid=resid, model=init, time_block="auto")
    _, o = run_oracle(wl, np.float64, resid=resid, model=init)
    for k in ("er", "se"):
        e = rel_l2(g[k], o[k])
        print(f"C1 auto-tuned grad {k} rel-L2 vs fp64 oracle {e:.3e}")
        assert e <= GRAD_TOL_FP64


def test_exact_gradient_matches_oracle_exact_adjoint(hip_device):
    """gradient='exact': the HIP transpose of the FP32 forward stepping
    against the oracle's FP64 exact adjoint (itself pinned by finite
    differences and the dot-product test)."""
    from fastgpr.workloads import small_case
    from helpers import oracle_problem
    from oracle import fdtd_oracle as O
    wl = small_case(nt=200)
    resid = _resid(wl)
    _, g = run_product(wl, hip_device, resid=resid, gradient="exact")
    pb = oracle_problem(wl)
    _, sn = O.run_forward(pb)
    ge, gs = O.run_gradient_exact(pb, resid.astype(np.float64), sn)
    for k, o in (("er", ge), ("se", gs)):
        e = rel_l2(g[k], o)
        print(f"exact grad {k} rel-L2 vs fp64 oracle {e:.3e}")
        assert e <= GRAD_TOL_FP64


def test_exact_gradient_central_differences_16_cells(hip_device):
    """Central differences of the GPU (FP32) forward map on 16 seeded cells
    for each of er and se, loss = <r, traces> in FP64.  er uses a relative
    step of 1e-3.  For se a 1e-3 step changes uE0 by ~1e-5 relative, which
    float32 resolves to only ~0.6%, so se uses 5e-2 here; the 1e-3 check of
    both parameters runs on the FP64 forward map in the oracle suite."""
    from fastgpr.workloads import small_case
    wl = small_case(nt=200)
    resid = _resid(wl).astype(np.float64)
    _, g = run_product(wl, hip_device, resid=resid.astype(np.float32), gradient="exact")
    rng = np.random.default_rng(3)
    NI, NJ = wl.NI, wl.NJ
    fd, ad = {"er": [], "se": []}, {"er": [], "se": []}
    cells = []
    while len(cells) < 16:
        c = (int(rng.integers(1, NI)), int(rng.integers(0, NJ)))
        if wl.se[c] < 100 and c not in cells:
            cells.append(c)
    for (i, j) in cells:
        for k in ("er", "se"):
            base = (wl.er if k == "er" else wl.se).astype(np.float64)
            h = (1e-3 if k == "er" else 5e-2) * base[i, j]
            vals = []
            for sgn in (1, -1):
                m = base.copy()
                m[i, j] += sgn * h
                model = (m, wl.se) if k == "er" else (wl.er, m)
                tr, _ = run_product(wl, hip_device, model=model)
                vals.append(np.sum(tr.astype(np.float64) * resid))
            fd[k].append((vals[0] - vals[1]) / (2 * h))
            ad[k].append(g[k][i, j])
    for k in ("er", "se"):
        a, f = np.array(ad[k]), np.array(fd[k])
        e = rel_l2(a, f)
        print(f"FD check {k}: rel-L2(adjoint, central differences) over 16 cells = {e:.3e}")
        assert e <= 2e-2


def test_nan_warnings_without_host_sync(hip_device, capsys):
    """The reference's NaN warnings (compute.py:898-906) are still printed:
    queued on the device, reported once their flags have arrived."""
    import torch
    from fastgpr.compute import report_nan_warnings
    from fastgpr.workloads import small_case
    wl = small_case(nt=60)
    er0 = wl.er0.copy()
    er0[wl.NI // 2, wl.NJ // 2] = np.nan
    resid = _resid(wl)
    tr, _ = run_product(wl, hip_device, resid=resid, model=(er0, wl.se0), time_block=(64, 1024, 6))
    report_nan_warnings(block=True)
    out = capsys.readouterr().out
    assert np.isnan(tr).any()
    assert "rxs_gpu contains NaN" in out
    # a clean run prints nothing
    run_product(wl, hip_device, resid=resid, time_block=(64, 1024, 6))
    report_nan_warnings(block=True)
    assert "NaN" not in capsys.readouterr().out


@pytest.mark.parametrize("P,NJ", [(10, 24), (20, 200), (3, 7)])
def test_pml_profile_on_device_matches_host(hip_device, P, NJ):
    """The CPML profile computed from device-resident er / mr (no copy to the
    host) equals the host computation (sequential sums, same expressions) to
    an ulp or two (the device's pow / division round differently)."""
    import torch
    from fastgpr import grid as g
    rng = np.random.default_rng(P + NJ)
    er = torch.tensor(1.0 + 9.0 * rng.random((2 * P + 5, NJ, 1)), dtype=torch.float32)
    mr = torch.tensor(1.0 + 0.5 * rng.random((2 * P + 5, NJ, 1)), dtype=torch.float32)
    host = g.pml_profile(P, 0.025, 4.1e-11, er, mr)
    dev = g.pml_profile(P, 0.025, 4.1e-11, er.to(hip_device), mr.to(hip_device))
    assert dev.device.type == "cuda" and dev.shape == host.shape
    np.testing.assert_allclose(dev.cpu().numpy(), host.numpy(), rtol=1.2e-6, atol=0)


def test_graph_cache_stays_bounded(hip_device):
    """With the cache bound at one graph every sweep evicts the other
    direction's graph (after its last launch has finished) and captures its
    own again; with buffers that move between calls the cache never exceeds
    its bound.  Every gradient is bitwise the first one."""
    import torch
    from fastgpr import solver
    from fastgpr.workloads import small_case
    wl = small_case(NI=40, NJ=36, P=6, S=2, nt=120, nrx=4)
    resid = np.random.default_rng(3).standard_normal((wl.S, wl.nt, wl.rx.shape[1])).astype(np.float32)
    kw = dict(resid=resid, time_block=(64, 1024, 6), shot_streams=1, graph_chunk=12)
    solver.release_graphs()
    _, first = run_product(wl, hip_device, **kw)
    prev = solver.set_graph_limit(1)
    try:
        assert solver.graph_count() <= 1
        for limit, calls in ((1, 4), (3, 10)):
            solver.set_graph_limit(limit)
            hold = []
            for k in range(calls):
                hold.append(torch.empty(4096 * (k + 1), dtype=torch.uint8, device=hip_device))
                _, g = run_product(wl, hip_device, **kw)
                assert solver.graph_count() <= limit, (limit, k)
                for n in ("er", "se"):
                    assert np.array_equal(g[n], first[n]), (limit, k, n)
    finally:
        solver.set_graph_limit(prev)
        solver.release_graphs()


class _PoisonedEmpty:
    """The torch module with every float32 ``empty`` filled with NaN."""

    def __init__(self, real):
        self._real = real

    def __getattr__(self, name):
        return getattr(self._real, name)

    def empty(self, *args, **kwargs):
        t = self._real.empty(*args, **kwargs)
        if t.dtype == self._real.float32:
            t.fill_(float("nan"))
        return t


@pytest.mark.parametrize("gradient,time_block,checkpoint", [
    ("exact", None, None), ("exact", None, 40), ("reference", None, None),
    ("reference", (64, 1024, 6), None), ("reference", (64, 1024, 6), 40)])
def test_gradient_reads_no_uninitialised_buffer(hip_device, monkeypatch, gradient, time_block,
                                                checkpoint):
    """Every float32 buffer the solver takes uninitialised (the exact
    adjoint's scratch, snapshot histories and rings, ...) is written before
    it is read: with each such buffer filled with NaN first, the traces and
    the gradient are bitwise the usual ones."""
    import torch
    from fastgpr import checkpoint as ck
    from fastgpr import solver
    from fastgpr.workloads import small_case
    wl = small_case(nt=160)
    resid = _resid(wl)
    kw = dict(resid=resid, gradient=gradient, time_block=time_block, checkpoint=checkpoint)
    print("case", wl.name, gradient, time_block, "checkpoint", checkpoint, flush=True)
    tr_ref, ref = run_product(wl, hip_device, **kw)
    poisoned = _PoisonedEmpty(torch)
    monkeypatch.setattr(solver, "torch", poisoned)
    monkeypatch.setattr(ck, "torch", poisoned)
    tr, got = run_product(wl, hip_device, **kw)
    assert np.array_equal(tr, tr_ref)
    for k in ("er", "se"):
        assert np.isfinite(got[k]).all()
        assert np.array_equal(got[k], ref[k]), k


def _compute_default(wl, dev, er, se, mr):
    """fastgpr.compute with the reference's arguments only (no build extras)."""
    import torch
    import fastgpr
    return fastgpr.compute(dev, wl.dx, wl.NJ, wl.NI, 1, ['hertzian', 'z'], ['ricker', 1, wl.freq],
                           wl.S, er, se, mr, wl.src.shape[1], wl.rx.shape[1],
                           time_windows=wl.time_window, dt=wl.dt, pmlthick=wl.P, tv=False,
                           customsource=torch.tensor(wl.src),
                           customreceiver=torch.tensor(wl.rx))[0]


def _default_vs_per_step(wl, dev):
    import torch
    from helpers import product_tensors
    resid = torch.tensor(_resid(wl)).to(dev)
    er, se, mr = product_tensors(wl, dev)
    er.requires_grad_(True)
    se.requires_grad_(True)
    tr = _compute_default(wl, dev, er, se, mr)
    tr.backward(resid)
    torch.cuda.synchronize()
    ref_tr, ref_g = run_product(wl, dev, resid=resid.cpu().numpy())      # time_block=None
    return (tr.detach().cpu().numpy(), er.grad.cpu().numpy()[..., 0],
            se.grad.cpu().numpy()[..., 0], ref_tr, ref_g)


def test_default_call_runs_autotuned_time_blocked_kernels(hip_device):
    """compute() called with the reference's arguments only takes the
    autotuned time-blocked path: one new tuner entry, traces bitwise the
    per-step kernels', gradients within 1e-12."""
    from fastgpr import autotune
    from fastgpr.workloads import small_case
    wl = small_case(nt=200)
    before = set(autotune._cache)
    tr, ge, gs, ref_tr, ref_g = _default_vs_per_step(wl, hip_device)
    new = set(autotune._cache) - before
    print("tuner entries added", new, flush=True)
    assert len(new) == 1
    assert np.array_equal(tr, ref_tr)
    assert rel_l2(ge, ref_g["er"]) < 1e-12 and rel_l2(gs, ref_g["se"]) < 1e-12


def test_default_call_falls_back_to_per_step_kernels(hip_device):
    """200 receivers on a 20 x 10 block: no time-blocked variant's region
    lists hold them, so the default call warns and runs the per-step
    kernels, bitwise."""
    from fastgpr.workloads import Workload, small_case
    base = small_case(NI=40, NJ=40, P=5, S=1, nt=120, nrx=6, seed=31)
    rx = np.zeros((1, 200, 3), np.int32)
    rx[0, :, 0] = 10 + np.arange(200) // 10
    rx[0, :, 1] = 10 + np.arange(200) % 10
    wl = Workload(base.name + "_dense_rx", base.er, base.se, base.mr, base.er0, base.se0, base.dx,
                  base.dt, base.nt, base.P, base.freq, base.src, rx)
    with pytest.warns(RuntimeWarning, match="per-step kernels"):
        tr, ge, gs, ref_tr, ref_g = _default_vs_per_step(wl, hip_device)
    assert np.array_equal(tr, ref_tr)
    assert np.array_equal(ge, ref_g["er"]) and np.array_equal(gs, ref_g["se"])
