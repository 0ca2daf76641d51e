import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def ctx():
    import milton_amd as m
    c = m.Context(0)
    yield c
    c.close()


def rays64(rays):
    """Structured float32 rays -> float64 (n, 6) for the oracle."""
    return np.concatenate([rays["o"], rays["d"]], axis=1).astype(np.float64)


def classify_hits(oracle_scene, rays, gpu_t, gpu_id, orc_t, orc_id):
    """Compares closest hits; returns (match fraction, non-tie mismatch
    fraction, relative t errors of matching hits).

    A disagreement counts as an exact tie when the double-precision t of the
    primitive the device chose equals the oracle's closest t to 1e-6 relative
    (shared edges and coincident faces)."""
    r = rays64(rays)
    same = gpu_id == orc_id
    diff = np.nonzero(~same)[0]
    non_tie = 0
    if len(diff):
        both_hit = (gpu_id[diff] != 0xFFFFFFFF) & (orc_id[diff] != 0xFFFFFFFF)
        t_alt = np.full(len(diff), np.inf)
        idx = diff[both_hit]
        if len(idx):
            t_alt[both_hit] = oracle_scene.prim_t(r[idx], gpu_id[idx])
        with np.errstate(invalid="ignore"):
            tie = both_hit & (np.abs(t_alt - orc_t[diff]) <= 1e-6 * np.maximum(orc_t[diff], 1e-3))
        non_tie = int((~tie).sum())
    hit = same & (orc_id != 0xFFFFFFFF)
    rel = np.abs(gpu_t[hit].astype(np.float64) - orc_t[hit]) / orc_t[hit]
    n = len(gpu_id)
    return same.mean(), non_tie / n, rel
