"""Write tests/golden/tv_golden.npz: input/output vectors of the reference's
TV denoiser ``total_variation`` (src/tv.py:135-217), which no closed form or
invariant pins.  Loads that one file by path (it imports only torch), on the
CPU, in float32; run from the repository root where /root/reference exists:

    python tests/golden/make_tv_golden.py
"""
import importlib.util
from pathlib import Path

import numpy as np
import torch

REF = Path("/root/reference/src/tv.py")
OUT = Path(__file__).resolve().parent / "tv_golden.npz"


def main():
    spec = importlib.util.spec_from_file_location("ref_tv", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    cases = [  # (M, N, lamda, rho, num_iter, tol, seed)
        (24, 16, 0.02, 1.0, 500, 1e-5, 0),
        (40, 30, 0.02, 1.0, 60, 1e-5, 1),
        (33, 47, 0.05, 2.0, 200, 1e-4, 2),
        (64, 48, 0.02, 1.0, 500, 1e-5, 3),
    ]
    arrays = {}
    for k, (M, N, lam, rho, nit, tol, seed) in enumerate(cases):
        rng = np.random.default_rng(seed)
        base = 4.0 + 6.0 * (np.arange(M)[:, None] >= M // 2) + 0.0 * np.arange(N)[None, :]
        data = (base + rng.normal(0.0, 0.6, (M, N))).astype(np.float32)[..., None]
        out = mod.total_variation(torch.tensor(data.copy()), lamda=lam, rho=rho,
                                  num_iter=nit, tol=tol)
        arrays[f"in{k}"] = data
        arrays[f"out{k}"] = out.numpy().astype(np.float32)
        arrays[f"par{k}"] = np.array([lam, rho, nit, tol], np.float64)
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, {k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main()
