"""Write tests/golden/notebook_models.npz: the models and recorded losses of
the reference's example notebook (example.ipynb cell 0 builds the models,
cell 1 runs the inversion and prints one loss per epoch).

Run in a container that holds the reference checkout:

    python tests/golden/make_notebook_models.py /root/reference

Reads only data files of the reference: ``newmodel.npy`` (numpy.load,
allow_pickle=False) and the text outputs of ``example.ipynb`` (JSON).  The
notebook's model construction is restated with numpy and scipy, step by step
(example.ipynb cell 0):

* ``resize(ep_true, (200, 100), order=1, mode='edge', anti_aliasing=True,
  preserve_range=True)`` on the (200, 100) model is the identity: the
  anti-aliasing sigma is (factor - 1) / 2 = 0 and an order-1 zoom by 1
  samples the input grid exactly; the same holds for the conductivity;
* ``filters.gaussian(x, sigma=20)`` is ``gaussian_filter(x, 20,
  mode='nearest', truncate=4.0)`` (float input, no rescaling);
* ``ep_truef = np.fliplr(ep_true)`` is a view, so ``ep_truef[:100, :] =
  ep_true[:100, :]`` writes into ``ep_true`` itself (its first 100 rows are
  mirrored left-right); the padded true permittivity was taken before that,
  the smoothed initial permittivity after it -- reproduced by running the
  same numpy statements.
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

import numpy as np
from scipy.ndimage import gaussian_filter

OUT = Path(__file__).resolve().parent / "notebook_models.npz"


def pad10(m: np.ndarray) -> np.ndarray:
    """(200, 100) -> (220, 120): the model in the middle, its first/last row
    repeated above/below, then the first/last column repeated (cell 0)."""
    p = np.zeros((220, 120))
    p[10:-10, 10:-10] = m
    p[:10, 10:-10] = m[0:1, :].repeat(10, axis=0)
    p[-10:, 10:-10] = m[-1:, :].repeat(10, axis=0)
    p[:, :10] = p[:, 10:11].repeat(10, axis=1)
    p[:, -10:] = p[:, -11:-10].repeat(10, axis=1)
    return p


def build(newmodel: np.ndarray) -> dict:
    ep_true = np.array(newmodel, dtype=np.float64)          # resize to its own shape: identity
    assert ep_true.shape == (200, 100)
    padded_ep = pad10(ep_true)
    ep_truef = np.fliplr(ep_true)                           # a view of ep_true
    ep_truef[:100, :] = ep_true[:100, :]                    # writes into ep_true
    se_true = ep_truef * 0.001                              # resize: identity
    se_true[(0.0008 <= se_true) & (se_true <= 0.0011)] = 0
    padded_s = pad10(se_true)
    epj = ep_true.copy()
    epj[epj <= 2.1] = 5
    sej = se_true.copy()
    sej[sej <= 0.0021] = 0.005
    epg = gaussian_filter(epj, 20, mode="nearest", truncate=4.0)
    sig = gaussian_filter(sej, 20, mode="nearest", truncate=4.0)
    ep_init = padded_ep.copy()
    ep_init[10:-10, 10:-10] = epg
    se_init = padded_s.copy()
    se_init[10:-10, 10:-10] = sig
    return {"ep_true": padded_ep, "se_true": padded_s, "ep_init": ep_init, "se_init": se_init}


def recorded_losses(nb_path: Path) -> np.ndarray:
    """'Epoch k, Loss: x' lines of the inversion cell's output, in order."""
    nb = json.loads(nb_path.read_text())
    out = []
    for cell in nb["cells"]:
        for o in cell.get("outputs", []):
            text = "".join(o.get("text", ""))
            for m in re.finditer(r"Epoch (\d+), Loss: ([0-9.]+)", text):
                assert int(m.group(1)) == len(out)
                out.append(float(m.group(2)))
    return np.array(out)


def main(ref: Path):
    models = build(np.load(ref / "newmodel.npy", allow_pickle=False))
    losses = recorded_losses(ref / "example.ipynb")
    np.savez_compressed(OUT, **models, losses=losses)
    print(f"wrote {OUT}: {len(losses)} recorded losses, epoch 0 {losses[0]}, epoch 1 {losses[1]}")


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
