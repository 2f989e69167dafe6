"""Write golden vectors for the fixed-Sobol grid learning from the reference's own code.

Runs ONLY in the build container, where the reference tree is mounted
read-only; the vectors it writes (``fixed_sobol_reference_vectors.npz``) are
data and travel with the tests.  It imports the reference's pure-torch
``learn_triple_importance_grid`` (S/dis/fixed_sobol.py:161-210), which no
closed form pins, and runs it on the CPU in float64 over scrambled Sobol
training sets with a synthetic integrand (``synthetic`` below, restated in
numpy by the test), for two grid sizes, with and without the angle fold and
with an I1 weight other than one.  The other fixed-Sobol pieces are pinned by
mathematics and by the reference's own known answers, not by this script.

Usage: ``python tests/golden/make_fixed_sobol_golden.py`` (rewrites the .npz).
"""

from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np
import torch

REFERENCE_SRC = Path("/root/reference/torch_nlo_clean/src")
OUT = Path(__file__).with_name("fixed_sobol_reference_vectors.npz")

# (seed, training sets, points per set, bins, fold_angle, i1_weight, maxr)
CASES = (
    (5, 3, 1024, 8, True, 1.0, 5.0),
    (17, 2, 2048, 32, True, 0.5, 5.0),
    (23, 2, 512, 16, False, 2.0, 30.0),
)


def synthetic(x: torch.Tensor):
    """(regular I1, sensitive I1, I2, I3) of a smooth test integrand of the
    five coordinates with sign changes and angular dependence."""

    phi = 2 * math.pi * x[:, 4]
    regular = x[:, 0] * (1 + x[:, 3]) * torch.cos(phi) + 0.1
    sensitive = 0.01 * x[:, 1] * x[:, 2]
    i2 = torch.exp(-3 * x[:, 1]) * x[:, 2] - 0.05
    i3 = (x[:, 3] - 0.4) * x[:, 0] ** 2 * (1 + torch.cos(phi) ** 2)
    return regular, sensitive, i2, i3


def main() -> None:
    sys.path.insert(0, str(REFERENCE_SRC))
    from nlo_torch.dis.fixed_sobol import learn_triple_importance_grid

    out: dict[str, np.ndarray] = {}
    for k, (seed, sets, count, bins, fold, i1_weight, maxr) in enumerate(CASES):
        training = tuple(
            torch.quasirandom.SobolEngine(5, scramble=True, seed=seed + j).draw_base2(
                count.bit_length() - 1, dtype=torch.float64)
            for j in range(sets))

        def function(x: torch.Tensor):
            regular, sensitive, i2, i3 = synthetic(x)
            return regular, sensitive.double(), i2, i3

        edges = learn_triple_importance_grid(function, training, maxr=maxr, bins=bins, fold_angle=fold,
                                             i1_weight=i1_weight)
        out[f"case{k}_params"] = np.array([seed, sets, count, bins, int(fold), i1_weight, maxr], dtype=np.float64)
        out[f"case{k}_edges"] = edges.numpy()
    np.savez(OUT, **out)
    print(f"wrote {OUT} ({len(CASES)} cases)")


if __name__ == "__main__":
    main()
