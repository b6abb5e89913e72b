"""Shared test setup: the `gpu` marker, repo imports and scene/oracle fixtures."""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests through the C-ABI")


def _ensure_library():
    lib = ROOT / "shiokaze_amd" / "libshk_projection.so"
    if not lib.exists():
        from shiokaze_amd import build as b
        b.build()
    return lib


@pytest.fixture(scope="session")
def shk_lib():
    _ensure_library()
    from shiokaze_amd import _lib
    return _lib.load()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.load()
    return O


@pytest.fixture(scope="session")
def scenes(shk_lib):
    """Cache of generated scenes keyed by (kind, n, seed)."""
    from shiokaze_amd import scene
    cache = {}

    def get(kind, n, seed=0):
        key = (kind, n if np.isscalar(n) else tuple(n), seed)
        if key not in cache:
            cache[key] = scene.generate(kind, n, seed)
        return cache[key]

    return get


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
