"""Checkpoint / recompute gradients (GPU): bitwise equal to full snapshot
storage, and the memory-bound configuration (2000 x 1000, 8000 steps,
4 shots) with the exact adjoint checked by central differences and against
the FP64 oracle on a 250 x 125 crop."""
import numpy as np
import pytest

from helpers import rel_l2, run_product

pytestmark = pytest.mark.gpu


def _resid(wl, seed=11):
    from helpers import host_grid
    nt, _ = host_grid(wl)
    rng = np.random.default_rng(seed)
    return rng.standard_normal((wl.S, nt, wl.rx.shape[1])).astype(np.float32)


@pytest.mark.parametrize("gradient,tb", [("reference", None), ("reference", (32, 512, 4)),
                                         ("reference", (64, 1024, 6)), ("exact", None)])
def test_checkpointed_gradient_equals_full_storage(hip_device, gradient, tb):
    from fastgpr.workloads import small_case
    wl = small_case(NI=70, NJ=90, nt=157)
    resid = _resid(wl)
    print("case", gradient, tb, flush=True)
    _, ga = run_product(wl, hip_device, resid=resid, gradient=gradient, time_block=tb)
    _, gb = run_product(wl, hip_device, resid=resid, gradient=gradient, time_block=tb,
                        checkpoint=37, graph_chunk=12)
    for k in ("er", "se"):
        assert np.array_equal(ga[k], gb[k]), k


def test_c4_checkpointed_exact_gradient_finite_differences(hip_device):
    """C4 at full size (2000 x 1000, 8000 steps, 4 shots, 200 receivers): the
    checkpointed (interval 96) exact gradient against central differences of
    the GPU forward map, loss <r, traces> in FP64, for 16 seeded cells per
    parameter.  A single cell of a 2-million-cell model moves the traces by
    ~1e-6 relative at a 1e-3 step, which is the FP32 noise floor of an
    8000-step run, so each check perturbs a Gaussian bump (radius 2 cells)
    centred on the seeded cell: peak relative step 1e-3 for er, 0.2 for se
    (se barely changes uE0).  Single-cell central differences at relative
    step 1e-3 for both parameters run on the FP64 oracle (oracle suite)."""
    import time
    from fastgpr.workloads import c4_memory
    wl = c4_memory()
    rng = np.random.default_rng(3)
    resid = _resid(wl, seed=4)
    t0 = time.time()
    _, g = run_product(wl, hip_device, resid=resid, gradient="exact", checkpoint=96,
                       graph_chunk=48, time_block="auto")
    print(f"C4 checkpointed exact gradient in {time.time() - t0:.1f} s", flush=True)
    for k in ("er", "se"):
        assert np.isfinite(g[k]).all() and np.abs(g[k]).max() > 0
    cells = []
    while len(cells) < 16:
        c = (int(rng.integers(300, 1700)), int(rng.integers(60, 940)))
        if c not in cells:
            cells.append(c)
    r64 = resid.astype(np.float64)
    ii, jj = np.meshgrid(np.arange(wl.NI), np.arange(wl.NJ), indexing="ij")
    fd = {"er": [], "se": []}
    ad = {"er": [], "se": []}
    for (i, j) in cells:
        bump = np.exp(-((ii - i) ** 2 + (jj - j) ** 2) / (2 * 2.0 ** 2))
        bump[bump < 1e-6] = 0
        for k, rel in (("er", 1e-3), ("se", 0.2)):
            base = (wl.er if k == "er" else wl.se).astype(np.float64)
            h = rel * base[i, j]
            vals = []
            for sgn in (1, -1):
                m = base + sgn * h * bump
                model = (m, wl.se) if k == "er" else (wl.er, m)
                tr, _ = run_product(wl, hip_device, model=model, time_block="auto")
                vals.append(np.sum(tr.astype(np.float64) * r64))
            fd[k].append((vals[0] - vals[1]) / (2 * h))
            ad[k].append(float(np.sum(g[k].astype(np.float64) * bump)))
        print("cell", (i, j), "done", flush=True)
    for k in ("er", "se"):
        e = rel_l2(np.array(ad[k]), np.array(fd[k]))
        print(f"C4 FD check {k}: rel-L2 over 16 cells = {e:.3e}")
        assert e <= 3e-2


def test_c4_crop_exact_gradient_against_fp64_oracle(hip_device):
    """250 x 125 crop of the C4 model around one source: the GPU exact
    gradient (checkpointed) against the oracle's FP64 exact adjoint."""
    from fastgpr.workloads import c4_memory, crop
    from helpers import oracle_problem
    from oracle import fdtd_oracle as O
    big = c4_memory()
    wl = crop(big, 300, 40, 250, 125, nt=1200, shots=[0],
              receivers=[(30 + 10 * r, 100) for r in range(20)])
    assert wl.S == 1 and wl.rx.shape[1] >= 1
    resid = _resid(wl)
    _, g = run_product(wl, hip_device, resid=resid, gradient="exact", checkpoint=64,
                       graph_chunk=32)
    pb = oracle_problem(wl)
    _, sn = O.run_forward(pb)
    ge, gs = O.run_gradient_exact(pb, resid.astype(np.float64), sn)
    for k, o in (("er", ge), ("se", gs)):
        e = rel_l2(g[k], o)
        print(f"crop exact grad {k} rel-L2 vs fp64 oracle {e:.3e}")
        assert e <= 1e-4
