"""Multiscale FIR filtering (src/multiscale.py): the taps are pinned by the
filter's defining properties (unit DC gain, linear phase, pass band near 1,
Hamming stop band), the filter itself by its impulse response, DC
preservation through the reflect padding, and an FP64 numpy restatement
(np.pad 'reflect' + correlation)."""
import numpy as np
import pytest
import torch

from helpers import rel_l2


def _response(h, f, fs):
    n = np.arange(len(h))
    return abs(np.sum(np.asarray(h, np.float64) * np.exp(-2j * np.pi * f / fs * n)))


@pytest.mark.parametrize("cutoff,fs", [(1e8, 1e10), (2e8, 1e10), (1e8, 1.5e10)])
def test_fir_taps_properties(cutoff, fs):
    """The short filter apply_filter builds (fs / cutoff taps): unit DC gain,
    linear phase, a response falling through the cut-off to a Hamming stop
    band."""
    from fastgpr.multiscale import design_fir_filter
    L = int(fs / cutoff)
    h = design_fir_filter(cutoff, fs, L).numpy()
    assert h.dtype == np.float32 and h.shape == (L,)
    assert abs(h.astype(np.float64).sum() - 1.0) < 1e-6                # unit DC gain
    # linear phase; for an even count the reference overwrites tap (L-1)//2
    # with 2 fc / fs although the centre falls between two taps, so that pair
    # is the one asymmetry (multiscale.py:11-12)
    c = (L - 1) // 2
    keep = np.ones(L, bool)
    if L % 2 == 0:
        keep[[c, L - 1 - c]] = False
    np.testing.assert_allclose(h[keep], h[::-1][keep], rtol=0, atol=1e-6 * np.abs(h).max())
    r = [_response(h, f, fs) for f in (cutoff / 4, cutoff / 2, cutoff, cutoff + 4.0 * fs / L)]
    assert 1.0 > r[0] > r[1] > r[2] > r[3] and r[3] < 1e-2


def test_long_fir_approaches_the_ideal_low_pass():
    """With many taps the windowed sinc tends to the ideal low-pass: gain 1
    below the cut-off, 1/2 at it, a Hamming stop band (< -50 dB) above it --
    which fixes the 2 fc / fs scaling and the window."""
    from fastgpr.multiscale import design_fir_filter
    fs, fc, L = 1.0, 0.01, 1001
    h = design_fir_filter(fc, fs, L).numpy()
    assert abs(_response(h, 0.005, fs) - 1.0) < 5e-3
    assert abs(_response(h, fc, fs) - 0.5) < 1e-2
    assert _response(h, 0.016, fs) < 3e-3


def _numpy_filter(x, h):
    """FP64: reflect-pad L-1 samples on the left of the time axis, correlate."""
    L = len(h)
    x = np.asarray(x, np.float64)
    xp = np.pad(x, [(0, 0), (L - 1, 0), (0, 0)], mode="reflect")
    out = np.zeros_like(x)
    for k in range(L):
        out += h[k] * xp[:, k:k + x.shape[1], :]
    return out


def _traces(S=3, nt=400, nrx=7, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((S, nt, nrx)).astype(np.float32)


def test_apply_filter_cpu_matches_fp64():
    from fastgpr.multiscale import apply_filter, design_fir_filter
    x = _traces()
    out = apply_filter(torch.tensor(x), 1e10, 2e8).numpy()
    h = design_fir_filter(2e8, 1e10, 50).numpy().astype(np.float64)
    assert out.shape == x.shape
    assert rel_l2(out, _numpy_filter(x, h)) <= 1e-6


def _impulse_and_dc(filt, device):
    from fastgpr.multiscale import design_fir_filter
    L, t0 = 50, 120
    h = design_fir_filter(2e8, 1e10, L).numpy()
    x = np.zeros((2, 300, 3), np.float32)
    x[:, t0, :] = 1.0
    out = filt(torch.tensor(x).to(device), 1e10, 2e8).cpu().numpy()
    # out[t0 + m] = h[L - 1 - m], exactly (one non-zero product per sample)
    assert np.array_equal(out[:, t0:t0 + L, :], np.broadcast_to(h[::-1][None, :, None], (2, L, 3)))
    assert not out[:, :t0, :].any() and not out[:, t0 + L:, :].any()
    c = np.full((1, 200, 2), 3.25, np.float32)
    dc = filt(torch.tensor(c).to(device), 1e10, 2e8).cpu().numpy()
    np.testing.assert_allclose(dc, c, rtol=2e-6)
    one = filt(torch.tensor(x[0, :, 0]).to(device), 1e10, 2e8).cpu().numpy()
    assert one.shape == (300,) and np.array_equal(one, out[0, :, 0])


def test_apply_filter_cpu_impulse_and_dc():
    from fastgpr.multiscale import apply_filter
    _impulse_and_dc(apply_filter, torch.device("cpu"))


@pytest.mark.gpu
def test_apply_filter_hip_matches_fp64(hip_device):
    from fastgpr.multiscale import apply_filter, design_fir_filter
    x = _traces(S=4, nt=1500, nrx=40, seed=2)
    out = apply_filter(torch.tensor(x).to(hip_device), 1e10, 1e8).cpu().numpy()
    h = design_fir_filter(1e8, 1e10, 100).numpy().astype(np.float64)
    assert rel_l2(out, _numpy_filter(x, h)) <= 1e-6


@pytest.mark.gpu
def test_apply_filter_hip_impulse_and_dc(hip_device):
    from fastgpr.multiscale import apply_filter
    _impulse_and_dc(apply_filter, hip_device)


@pytest.mark.gpu
def test_apply_filter_hip_rejects_short_traces(hip_device):
    from fastgpr.multiscale import apply_filter
    with pytest.raises(ValueError):
        apply_filter(torch.zeros((1, 40, 2), device=hip_device), 1e10, 2e8)
