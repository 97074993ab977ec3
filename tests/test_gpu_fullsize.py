"""HIP path at the full 31.4M-visibility / 601-plane / support-72 / 16384^2
shape, checked through properties that do not need the serial gridder to run
over all visibilities.

Bars (written in each test):
  - FP32 grid vs the FP64 truth with the same selection (computed on the GPU
    by mi355_grid_wproj_fsel_d_dev, itself checked against the oracle at
    small sizes): norm-wise max|G - G64| / max|G64| <= 1e-5, plain and
    compensated;
  - selection of every visibility and num_skipped: bit-exact against the
    oracle's FP32 selection (oskar_grid_wproj.c:158-187);
  - linearity: doubling every amplitude doubles the grid bit for bit (the
    scaling by 2 is exact in every product and sum); two runs are bitwise
    identical;
  - the densest grid tiles and three sampled ones against the oracle's FP64
    tile restatement: <= 1e-5 of the tile's largest value.
"""
import numpy as np
import pytest

from conftest import errors
from oracle import oracle
from mi355_gridding import synth, _lib

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


@pytest.fixture(scope="module")
def full(gpu):
    torch, ops = gpu
    c = synth.make(2)
    ks = ops.KernelStack(c["support"], c["oversample"], c["conv_size_half"], c["conv_func"])
    d = {k: torch.from_numpy(np.ascontiguousarray(c[k])).cuda()
         for k in ("uu", "vv", "ww", "vis", "weight")}
    return torch, ops, c, ks, d


def _grid(full, vis=None, mode="plain"):
    torch, ops, c, ks, d = full
    G = c["grid_size"]
    ctx = ops.context()
    dtype = torch.float64 if mode == "fsel_d" else torch.float32
    g = torch.zeros((G, G, 2), dtype=dtype, device="cuda")
    norm = torch.zeros(1, dtype=torch.float64, device="cuda")
    skipped = torch.zeros(1, dtype=torch.int64, device="cuda")
    args = (ks, d["uu"], d["vv"], d["ww"], d["vis"] if vis is None else vis, d["weight"],
            c["cell_size_rad"], c["w_scale"], G, g, norm, skipped)
    ctx.set_option(_lib.OPT_COMPENSATED, 1 if mode == "comp" else 0)
    try:
        if mode == "fsel_d":
            ops.grid_fsel_d(*args, ctx=ctx)
        else:
            ops.grid(*args, ctx=ctx)
    finally:
        ctx.set_option(_lib.OPT_COMPENSATED, 0)
    torch.cuda.synchronize()
    return g, float(norm.item()), int(skipped.item())


def _rel_err(torch, g, t):
    return float((g.double() - t).abs().max().item() / t.abs().max().item())


def test_full_grid_vs_fp64_truth(full):
    torch = full[0]
    t, tn, ts = _grid(full, mode="fsel_d")
    g, n, s = _grid(full)
    assert s == ts and abs(n - tn) <= 1e-12 * abs(tn)
    e_plain = _rel_err(torch, g, t)
    del g
    gc, _, _ = _grid(full, mode="comp")
    e_comp = _rel_err(torch, gc, t)
    assert e_plain <= 1e-5, e_plain
    assert e_comp <= 1e-5 and e_comp <= e_plain, (e_comp, e_plain)


def test_full_selection_bitexact(full):
    torch, ops, c, ks, d = full
    ref = oracle.select_f(c)
    got = ops.select(c["support"], c["oversample"], d["uu"], d["vv"], d["ww"],
                     c["cell_size_rad"], c["w_scale"], c["grid_size"])
    for k in ("grid_u", "grid_v", "off_u", "off_v", "plane", "support", "conj", "skip"):
        assert np.array_equal(got[k].cpu().numpy(), ref[k].astype(np.int32)), k
    _, _, s = _grid(full)
    assert s == int(np.count_nonzero(ref["skip"]))


def test_full_linearity_and_determinism(full):
    torch, ops, c, ks, d = full
    g1, n1, _ = _grid(full)
    g1b, n1b, _ = _grid(full)
    assert torch.equal(g1, g1b) and n1 == n1b
    g2, _, _ = _grid(full, vis=d["vis"] * 2)
    assert torch.equal(g2, g1 * 2)


def test_full_tiles_vs_oracle(full):
    torch, ops, c, ks, d = full
    g, _, _ = _grid(full)
    G, T = c["grid_size"], 64
    # tile occupancy from the oracle's selection: densest tile + 3 sampled
    sel = oracle.select_f(c)
    keep = sel["skip"] == 0
    tid = (sel["grid_v"][keep] // T) * (G // T) + sel["grid_u"][keep] // T
    counts = np.bincount(tid, minlength=(G // T) ** 2)
    rng = np.random.default_rng(5)
    occupied = np.flatnonzero(counts)
    picks = [int(np.argmax(counts))] + [int(x) for x in rng.choice(occupied, 3, replace=False)]
    tiles = np.array([[p % (G // T), p // (G // T)] for p in picks], np.int32)
    ref = oracle.tiles_fsel_d(c, tiles, tile=T, threads=8)
    for (tx, ty), r in zip(tiles, ref):
        got = g[ty * T:(ty + 1) * T, tx * T:(tx + 1) * T].cpu().numpy()
        assert errors(got, r)[0] <= 1e-5, (tx, ty, errors(got, r))


# ---------------------------------------------------------------------------
# Configs 0 and 1 at their full sizes
# ---------------------------------------------------------------------------
def _device_grid(torch, ops, c, mode="plain"):
    ks = ops.KernelStack(c["support"], c["oversample"], c["conv_size_half"], c["conv_func"])
    d = {k: torch.from_numpy(np.ascontiguousarray(c[k])).cuda()
         for k in ("uu", "vv", "ww", "vis", "weight")}
    G = c["grid_size"]
    g = torch.zeros((G, G, 2), dtype=torch.float32, device="cuda")
    norm = torch.zeros(1, dtype=torch.float64, device="cuda")
    skipped = torch.zeros(1, dtype=torch.int64, device="cuda")
    ctx = ops.context()
    ctx.set_option(_lib.OPT_COMPENSATED, 1 if mode == "comp" else 0)
    try:
        ops.grid(ks, d["uu"], d["vv"], d["ww"], d["vis"], d["weight"], c["cell_size_rad"],
                 c["w_scale"], G, g, norm, skipped, ctx=ctx)
    finally:
        ctx.set_option(_lib.OPT_COMPENSATED, 0)
    torch.cuda.synchronize()
    return g.cpu().numpy(), float(norm.item()), int(skipped.item())


def test_config0_full_vs_oracle(gpu):
    """All 200,000 visibilities of the 2048^2 / 64-plane workload against the
    serial FP64 truth with the same selection (whole grid, <= 1e-5 norm-wise,
    plain and compensated), the serial FP32 gridder's own grid (<= 1e-5 of
    the truth's largest value: both are FP32 sums of the same products in
    different orders), num_skipped exact and the norm to 1e-12 of the serial
    FP32 norm."""
    torch, ops = gpu
    c = synth.make(0)
    assert c["num_points"] == 200_000 and c["grid_size"] == 2048
    truth, tnorm, tskip = oracle.grid_fsel_d(c)
    ref_f, rnorm, rskip = oracle.grid_f(c, use_ref=oracle.ref() is not None)
    assert rskip == tskip and tskip > 0            # the exponential tail leaves the grid
    for mode in ("plain", "comp"):
        g, n, s = _device_grid(torch, ops, c, mode)
        assert s == tskip
        assert abs(n - rnorm) <= 1e-12 * abs(rnorm), (mode, n, rnorm)
        e = errors(g, truth)[0]
        assert e <= 1e-5, (mode, e)
        m = float(np.max(np.abs(truth)))
        assert float(np.max(np.abs(g.astype(np.float64) - ref_f))) <= 1e-5 * m, mode


def test_config1_full_tiles_vs_oracle(gpu):
    """All 2,152,800 visibilities of the 8192^2 / 714-plane / support-36
    workload: selection bit-exact against the oracle, num_skipped exact, the
    norm to 1e-12 of the serial FP32 norm, and the densest 64x64 tile plus
    three sampled occupied tiles within 1e-5 of the tile's largest value
    against the FP64 tile truth with the same selection."""
    torch, ops = gpu
    c = synth.make(1)
    assert c["num_points"] == 2_152_800 and c["num_w_planes"] == 714
    sel = oracle.select_f(c)
    rnorm, rskip = oracle.norm_f(c)
    g, n, s = _device_grid(torch, ops, c)
    assert s == rskip == int(np.count_nonzero(sel["skip"]))
    assert abs(n - rnorm) <= 1e-12 * abs(rnorm), (n, rnorm)
    G, T = c["grid_size"], 64
    keep = sel["skip"] == 0
    tid = (sel["grid_v"][keep] // T) * (G // T) + sel["grid_u"][keep] // T
    counts = np.bincount(tid, minlength=(G // T) ** 2)
    rng = np.random.default_rng(9)
    occ = np.flatnonzero(counts)
    picks = [int(np.argmax(counts))] + [int(x) for x in rng.choice(occ, 3, replace=False)]
    tiles = np.array([[p % (G // T), p // (G // T)] for p in picks], np.int32)
    ref = oracle.tiles_fsel_d(c, tiles, tile=T, threads=8)
    for (tx, ty), r in zip(tiles, ref):
        got = g[ty * T:(ty + 1) * T, tx * T:(tx + 1) * T]
        assert errors(got, r)[0] <= 1e-5, (tx, ty, errors(got, r))
