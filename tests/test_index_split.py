"""The liquid kernels split a flat 32-bit face or cell index q into (i, j, k) with
floor(q / n) = floor(m q / 2^64), m = floor((2^64 - 1) / n) + 1 computed on the host
(shiokaze_amd/csrc/liquid.hip, Div32 / quot; Lemire, Kaser and Kurz 2019).  Checked here
against integer division for the extents a grid can have, at the indices where an
off-by-one would show (multiples of n and their neighbours, the 32-bit ends) and at
random ones, and the whole (i, j, k) split against numpy's unravel_index."""
from __future__ import annotations

import numpy as np


def magic(n: int) -> int:
    return (2**64 - 1) // n + 1


def quot(q: int, n: int) -> int:
    return (magic(n) * q) >> 64 if n > 1 else q


def split(q: int, nx: int, ny: int) -> tuple[int, int, int]:
    t = quot(q, nx)
    k = quot(t, ny)
    return q - t * nx, t - k * ny, k


def test_quotient_matches_integer_division():
    rng = np.random.default_rng(3)
    extents = list(range(1, 700)) + [1023, 1024, 1025, 2049, 4097, 65535, 65536, 65537, 2**31 - 1, 2**31, 2**32 - 1]
    for n in extents:
        qs = {0, 1, n - 1, n, n + 1, 2**32 - 1, 2**32 - 2}
        for k in (1, 3, 1000, (2**32 - 1) // n):
            qs.update(k * n + r for r in (-1, 0, 1))
        qs.update(int(x) for x in rng.integers(0, 2**32, 64, dtype=np.uint64))
        for q in qs:
            if 0 <= q < 2**32:
                assert quot(q, n) == q // n, (q, n)


def test_split_matches_unravel_index():
    rng = np.random.default_rng(4)
    for shape in [(257, 256, 256), (256, 257, 256), (256, 256, 257), (41, 28, 36), (2, 2, 2), (513, 512, 512)]:
        nx, ny, nz = shape
        total = nx * ny * nz
        qs = np.concatenate([np.arange(min(total, 2000)), rng.integers(0, total, 2000), [total - 1]])
        k_ref, j_ref, i_ref = np.unravel_index(qs, (nz, ny, nx))
        for q, i, j, k in zip(qs.tolist(), i_ref.tolist(), j_ref.tolist(), k_ref.tolist()):
            assert split(q, nx, ny) == (i, j, k), (shape, q)
