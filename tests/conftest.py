"""Shared test set-up.

Markers: ``gpu`` tests need an MI355X (they run the HIP path through the C
ABI); everything else runs on the CPU.  The CPU checkers come from oracle/
(see oracle/oracle.py); golden fixtures live in tests/golden/.
"""
import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpu-gridding_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP path)")
    config.addinivalue_line("markers", "slow: full-size workload")


def golden_names(prefix=""):
    """Fixture names: the gridding fixtures by default, the weighting ones
    (weights_*.npz) with prefix="weights_"."""
    names = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))
    if prefix:
        return [n for n in names if n.startswith(prefix)]
    return [n for n in names if not n.startswith("weights_")]


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    for k in ("num_w_planes", "oversample", "conv_size_half", "grid_size", "num_points"):
        d[k] = int(d[k])
    d["cell_size_rad"] = np.float32(d["cell_size_rad"])
    d["w_scale"] = np.float32(d["w_scale"])
    return d


def errors(test, truth):
    """Norm-wise errors of a complex grid ([..., 2] real or complex arrays):
    max|a-b| / max|b| and ||a-b||_2 / ||b||_2 (per-cell relative error is
    ill-posed where the grid cancels)."""
    a = np.asarray(test, np.float64)
    b = np.asarray(truth, np.float64)
    d = a - b
    bmax = np.max(np.abs(b))
    return (float(np.max(np.abs(d)) / bmax) if bmax > 0 else float(np.max(np.abs(d))),
            float(np.linalg.norm(d) / np.linalg.norm(b)) if np.linalg.norm(b) > 0
            else float(np.linalg.norm(d)))


def same_bits(a, b):
    """Bitwise equality, NaN-aware."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


def same_bits_or_nan(a, b):
    """Bitwise equality of the non-NaN values and NaN in the same places (the
    sign and payload of a NaN that an invalid operation creates differ
    between x86 and the GPU; IEEE leaves them open)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and a[~na].tobytes() == b[~nb].tobytes()


def same_scalar(a, b):
    """Equal, or both NaN."""
    a, b = float(a), float(b)
    return a == b or (a != a and b != b)


def close_scalar(a, b, rel):
    a, b = float(a), float(b)
    if a != a or b != b:
        return a != a and b != b
    return abs(a - b) <= rel * abs(b)


def nonfinite_split(test, truth):
    """For grids with NaN / Inf cells: True when the NaN cells and the +-Inf
    cells coincide; returns (ok, finite test values, finite truth values)."""
    a = np.asarray(test, np.float64)
    b = np.asarray(truth, np.float64)
    ok = np.array_equal(np.isnan(a), np.isnan(b)) and \
        np.array_equal(np.isposinf(a), np.isposinf(b)) and \
        np.array_equal(np.isneginf(a), np.isneginf(b))
    fin = np.isfinite(b)
    return ok, a[fin], b[fin]


@pytest.fixture(scope="session")
def gpu():
    """torch + the gridder library on cuda:0; fails (does not skip) when the
    library is missing, skips only where no GPU is visible."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from mi355_gridding import _lib
    _lib.load()
    from mi355_gridding import torch_ops
    return torch, torch_ops
