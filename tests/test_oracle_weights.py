"""Pins the oracle's restatement of the uniform-weighting pre-pass
(oracle_grid_weights_write/read_f/_d in oracle/grid_oracle.c, following
oskar_grid_weights.c) before the GPU path is checked against it.

1. Golden fixtures written by the reference's compiled functions
   (tests/golden/make_golden.py, weights_*.npz): two accumulating writes and
   a read, bit for bit in FP32 and FP64, with cells on the rounding and grid
   edges, non-finite coordinates and weights, and cells that cancel to 0.
2. oracle/_ref/libref.so, where it was built: bit-identical on fresh inputs.
3. Independent numpy statements: the cell of a point (round-half-away
   written with floor and an exact fraction test), the per-cell sums as
   np.add.at (unbuffered, in index order, so the same float additions in the
   same order), unit weights giving per-cell counts, read after write giving
   1/count, and untouched outputs for points off the grid.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_names, same_bits
from oracle import oracle


def _golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _run(c, real, use_ref=False):
    G = int(c["grid_size"])
    cell = real(c["cell_size_rad"])
    grid = np.zeros((G, G), real)
    sk = [oracle.weights_write(c["uu_" + b].astype(real), c["vv_" + b].astype(real),
                               c["w_" + b].astype(real), cell, G, grid, use_ref=use_ref)
          for b in ("a", "b")]
    wo = np.full(c["uu_a"].size, -7.0, real)
    sk.append(oracle.weights_read(c["uu_a"].astype(real), c["vv_a"].astype(real),
                                  c["w_a"].astype(real), wo, cell, G, grid, use_ref=use_ref))
    return grid, wo, sk


def numpy_cells(uu, vv, cell, G):
    """Cell of each point, independently of the oracle: position in the input
    precision, round half away from zero as floor(|x|) + (frac >= 1/2) (the
    fraction |x| - floor(|x|) is exact), x86 conversion (INT_MIN outside
    int range) and a wrapping add of the centre."""
    real = uu.dtype.type
    gs = real(G) * real(cell)

    def rnd(x):
        with np.errstate(invalid="ignore"):
            a = np.abs(x)
            f = np.floor(a)
            r = np.copysign(f + ((a - f) >= real(0.5)), x)
            ok = np.isfinite(r) & (r >= -2.0 ** 31) & (r < 2.0 ** 31)
            out = np.full(x.shape, -2 ** 31, np.int64)
            out[ok] = r[ok].astype(np.int64)
        return out

    with np.errstate(invalid="ignore", over="ignore"):
        pu = -uu * gs
        pv = vv * gs
    gu = (rnd(pu) + G // 2 + 2 ** 31) % 2 ** 32 - 2 ** 31
    gv = (rnd(pv) + G // 2 + 2 ** 31) % 2 ** 32 - 2 ** 31
    inside = (gu >= 0) & (gu < G) & (gv >= 0) & (gv < G)
    return gu, gv, inside


# ---------------------------------------------------------------------------
# 1. golden fixtures
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", golden_names("weights_"))
@pytest.mark.parametrize("tag,real", [("f", np.float32), ("d", np.float64)])
def test_oracle_matches_reference_fixture(name, tag, real):
    c = _golden(name)
    grid, wo, sk = _run(c, real)
    assert same_bits(grid, c["grid_" + tag])
    assert same_bits(wo, c["weight_out_" + tag])
    assert sk == [int(x) for x in c["skipped_" + tag]]


def test_fixtures_cover_the_edges():
    c = _golden("weights_random_g96")
    assert np.isnan(c["grid_f"]).any() and np.isinf(c["grid_f"]).any()
    assert (c["weight_out_f"] == 0).any() and (c["weight_out_f"] == -7).any()
    sizes = set()
    for name in ("weights_special_g64", "weights_special_g63"):
        c = _golden(name)
        assert c["skipped_f"][0] > 0
        sizes.add(int(c["grid_size"]) % 2)
    assert sizes == {0, 1}            # even and odd grid sizes


# ---------------------------------------------------------------------------
# 2. compiled reference
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("real", [np.float32, np.float64])
def test_oracle_bit_identical_to_compiled_reference(seed, real):
    if oracle.ref() is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    rng = np.random.default_rng(seed)
    G, N = 200 + seed, 50_000
    c = dict(grid_size=G, cell_size_rad=1.0 / (G * 3.0))
    for b in ("a", "b"):
        c["uu_" + b] = rng.normal(0, G, N)
        c["vv_" + b] = rng.normal(0, G, N)
        c["w_" + b] = rng.lognormal(0, 1, N)
    g1, w1, s1 = _run(c, real)
    g2, w2, s2 = _run(c, real, use_ref=True)
    assert same_bits(g1, g2) and same_bits(w1, w2) and s1 == s2


# ---------------------------------------------------------------------------
# 3. independent statements
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("real", [np.float32, np.float64])
def test_cells_sums_and_counts_match_numpy(real):
    rng = np.random.default_rng(11)
    G, N = 77, 40_000
    cell = real(1.0 / (G * 2.0))
    uu = rng.uniform(-1.3, 1.3, N).astype(real) * real(G)
    vv = rng.uniform(-1.3, 1.3, N).astype(real) * real(G)
    w = rng.uniform(0.1, 5.0, N).astype(real)
    gu, gv, inside = numpy_cells(uu, vv, cell, G)
    t = (gv * G + gu)[inside]

    grid = np.zeros((G, G), real)
    sk = oracle.weights_write(uu, vv, w, cell, G, grid)
    assert sk == int((~inside).sum())
    expect = np.zeros(G * G, real)
    np.add.at(expect, t, w[inside])            # in index order, one float add each
    assert same_bits(grid.ravel(), expect)

    ones = np.zeros((G, G), real)
    oracle.weights_write(uu, vv, np.ones(N, real), cell, G, ones)
    counts = np.bincount(t, minlength=G * G).astype(real)
    assert same_bits(ones.ravel(), counts)

    wo = np.full(N, real(-3.0))
    sk = oracle.weights_read(uu, vv, np.ones(N, real), wo, cell, G, ones)
    assert sk == int((~inside).sum())
    assert same_bits(wo[inside], real(1.0) / counts[t])
    assert np.all(wo[~inside] == real(-3.0))


def test_zero_cell_total_reads_zero_and_order_matters():
    """A cell whose weights cancel reads 0; summing in another order than the
    input's changes the FP32 total (so the restatement's order is pinned)."""
    real = np.float32
    G = 8
    uu = np.zeros(4, real)
    vv = np.zeros(4, real)
    w = np.array([1e8, 1.0, -1e8, 1.0], real)
    grid = np.zeros((G, G), real)
    oracle.weights_write(uu, vv, w, real(0.1), G, grid)
    assert grid[G // 2, G // 2] == real(1.0)        # ((1e8 + 1) - 1e8) + 1
    assert (real(1e8) + real(-1e8) + real(1.0) + real(1.0)) == real(2.0)
    wz = np.array([2.5, -2.5], real)
    gz = np.zeros((G, G), real)
    oracle.weights_write(uu[:2], vv[:2], wz, real(0.1), G, gz)
    wo = np.full(2, real(9.0))
    oracle.weights_read(uu[:2], vv[:2], wz, wo, real(0.1), G, gz)
    assert np.all(wo == 0)
