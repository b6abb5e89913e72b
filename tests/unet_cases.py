"""Synthetic training samples and parameters shared by the learned-preconditioner tests."""
from __future__ import annotations

import numpy as np

from oracle import unet_oracle as U


def random_system(shape, seed):
    """Unit-diagonal 7-point operator on a masked box: couplings in [0, 1/6], zero on
    faces that touch a masked-out cell or the box edge (diagonally dominant, SPD)."""
    rng = np.random.default_rng(seed)
    mask = (rng.random(shape) < 0.85).astype(np.float64)
    cs = []
    for ax in range(3):
        c = rng.random(shape) / 6.0
        lo = np.roll(mask, 1, axis=ax)
        sl = [slice(None)] * 3
        sl[ax] = 0
        lo[tuple(sl)] = 0.0
        cs.append(c * mask * lo)
    return mask, cs


def training_sample(shape, seed):
    """(r, mask, cx, cy, cz, xstar) with r = A xstar on the unknowns."""
    mask, (cx, cy, cz) = random_system(shape, seed)
    rng = np.random.default_rng(seed + 100)
    xstar = rng.standard_normal(shape) * mask
    r = U.stencil(cx, cy, cz, xstar) * mask
    return r, mask, cx, cy, cz, xstar


def init_params(seed, bias_sd=0.05):
    """He-normal weights and small random biases in the library's parameter order."""
    rng = np.random.default_rng(seed)
    p = np.zeros(U.num_params())
    for (cin, cout, _), (wo, bo) in zip(U.LAYERS, U.layer_offsets()[0]):
        p[wo:bo] = rng.standard_normal(bo - wo) * np.sqrt(2.0 / (27 * cin))
        p[bo:bo + cout] = rng.standard_normal(cout) * bias_sd
    return p
