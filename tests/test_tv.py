"""TV denoiser (src/tv.py:135-217): the torch restatement and the HIP
kernels against input/output vectors written from the reference's own
``total_variation`` (tests/golden/make_tv_golden.py).  The reference sums the
4-neighbourhood with conv2d, whose summation order is its library's, so the
comparison is to float32 rounding (relative L2 <= 1e-6), not bitwise."""
from pathlib import Path

import numpy as np
import pytest
import torch

from helpers import rel_l2

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "tv_golden.npz")
CASES = range(len([k for k in GOLD.files if k.startswith("in")]))


def _params(k):
    lam, rho, nit, tol = GOLD[f"par{k}"]
    return float(lam), float(rho), int(nit), float(tol)


@pytest.mark.parametrize("k", CASES)
def test_torch_tv_matches_reference_vectors(k):
    from fastgpr.tv import total_variation_torch
    lam, rho, nit, tol = _params(k)
    out = total_variation_torch(torch.tensor(GOLD[f"in{k}"]), lam, rho, nit, tol).numpy()
    assert out.shape == GOLD[f"out{k}"].shape
    assert rel_l2(out, GOLD[f"out{k}"]) <= 1e-6


def test_tv_squeezes_its_argument_in_place_like_the_reference():
    from fastgpr.tv import total_variation
    x = torch.tensor(GOLD["in0"])
    out = total_variation(x)
    assert x.shape == GOLD["in0"].shape[:2] and out.shape == GOLD["in0"].shape


@pytest.mark.gpu
@pytest.mark.parametrize("k", CASES)
def test_hip_tv_matches_reference_vectors(k, hip_device):
    from fastgpr.tv import total_variation
    lam, rho, nit, tol = _params(k)
    out = total_variation(torch.tensor(GOLD[f"in{k}"]).to(hip_device), lam, rho, nit, tol)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), GOLD[f"out{k}"]) <= 1e-6


@pytest.mark.gpu
def test_hip_tv_large_grid_matches_torch(hip_device):
    from fastgpr.tv import total_variation, total_variation_torch
    from fastgpr.workloads import layered_blob
    er, _ = layered_blob(400, 200, seed=1)
    x = torch.tensor(er.astype(np.float32))[..., None].to(hip_device)
    a = total_variation(x.clone()).cpu().numpy()
    b = total_variation_torch(x.clone()).cpu().numpy()
    assert rel_l2(a, b) <= 1e-6


@pytest.mark.gpu
def test_gradient_with_tv_term(hip_device):
    """GPRFWI.backward with tv=True adds 0.01 * 2 * (m - TV(m)) to each
    normalised gradient (compute.py:916-933)."""
    from fastgpr.tv import total_variation_torch
    from fastgpr.workloads import small_case
    from helpers import run_product
    wl = small_case(NI=50, NJ=40, P=6, S=2, nt=150, nrx=4)
    resid = np.random.default_rng(1).standard_normal((wl.S, wl.nt, wl.rx.shape[1])).astype(np.float32)
    _, g0 = run_product(wl, hip_device, resid=resid, tv=False)
    _, g1 = run_product(wl, hip_device, resid=resid, tv=True)
    for k, m in (("er", np.maximum(wl.er, 1.0)), ("se", np.maximum(wl.se, 0.0))):
        mt = torch.tensor(m.astype(np.float32))[..., None]
        expect = g0[k] + (0.01 * (2 * (mt - total_variation_torch(mt.clone()))))[..., 0].numpy()
        assert rel_l2(g1[k], expect) <= 1e-6, k
