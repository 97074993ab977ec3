"""HIP path vs the oracle, through the C ABI.

Bars (written in each test):
  - selection (cell, offsets, plane, support, conj, skip) and num_skipped:
    bit-exact against the serial code's arithmetic;
  - FP32 grid: norm-wise max|G - G64| / max|G64| <= 1e-5 against the FP64
    truth with the same (FP32) selection; the same bar against the
    reference's own FP32 grid;
  - FP64 grid (oskar_grid_wproj_d drop-in) and the FP32-selection/FP64
    path: <= 1e-12;
  - norm: relative 1e-12 (FP64 sums in a different order).
"""
import numpy as np
import pytest

from conftest import (close_scalar, errors, golden_names, load_golden, nonfinite_split,
                      same_bits)
from oracle import oracle
from mi355_gridding import synth, wproj, _lib

pytestmark = pytest.mark.gpu

FIELDS = ("grid_u", "grid_v", "off_u", "off_v", "plane", "support", "conj", "skip")


def dev(torch, a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def device_case(torch, ops, c, dtype=None):
    dtype = dtype or torch.float32
    ks = ops.KernelStack(c["support"], c["oversample"], c["conv_size_half"], c["conv_func"],
                         dtype=dtype)
    d = {k: dev(torch, c[k], dtype) for k in ("uu", "vv", "ww", "vis", "weight")}
    return ks, d


def run_grid(torch, ops, c, dtype=None, mode="plain", ctx=None, visits_per_item=None):
    dtype = dtype or torch.float32
    ks, d = device_case(torch, ops, c, dtype)
    G = c["grid_size"]
    out_dtype = torch.float64 if mode == "fsel_d" else dtype
    grid = torch.zeros((G, G, 2), dtype=out_dtype, device="cuda")
    norm = torch.zeros(1, dtype=torch.float64, device="cuda")
    skipped = torch.zeros(1, dtype=torch.int64, device="cuda")
    ctx = ctx or ops.context()
    vpi0 = ctx.get_option(_lib.OPT_VISITS_PER_ITEM)
    ctx.set_option(_lib.OPT_COMPENSATED, 1 if mode == "comp" else 0)
    if visits_per_item:
        ctx.set_option(_lib.OPT_VISITS_PER_ITEM, visits_per_item)
    try:
        if mode == "fsel_d":
            ops.grid_fsel_d(ks, d["uu"], d["vv"], d["ww"], d["vis"], d["weight"],
                            c["cell_size_rad"], c["w_scale"], G, grid, norm, skipped, ctx=ctx)
        else:
            ops.grid(ks, d["uu"], d["vv"], d["ww"], d["vis"], d["weight"], c["cell_size_rad"],
                     c["w_scale"], G, grid, norm, skipped, ctx=ctx)
    finally:
        ctx.set_option(_lib.OPT_COMPENSATED, 0)
        ctx.set_option(_lib.OPT_VISITS_PER_ITEM, vpi0)
    torch.cuda.synchronize()
    return grid.cpu().numpy(), float(norm.item()), int(skipped.item())


def small_config(cfg, N, G, P=None, smax=None, seed=None):
    c = synth.make(cfg, num_points=N, grid_size=G, num_w_planes=P, max_support=smax, seed=seed)
    return c


# ---------------------------------------------------------------------------
# selection: bit-exact
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", golden_names())
def test_selection_bitexact_on_fixtures(gpu, name):
    torch, ops = gpu
    c = load_golden(name)
    for real, sel_fn in ((torch.float32, oracle.select_f), (torch.float64, oracle.select_d)):
        ref = sel_fn(c)
        got = ops.select(c["support"], c["oversample"], dev(torch, c["uu"], real),
                         dev(torch, c["vv"], real), dev(torch, c["ww"], real),
                         c["cell_size_rad"], c["w_scale"], c["grid_size"])
        for k in FIELDS:
            assert np.array_equal(got[k].cpu().numpy(), ref[k]), (name, str(real), k)


@pytest.mark.parametrize("cfg", [0, 1, 2])
def test_selection_bitexact_config_shapes(gpu, cfg):
    """Full-size coordinates of configs 0-2 plus adversarial positions placed
    on half-integers and offset-rounding boundaries."""
    torch, ops = gpu
    c = synth.make(cfg) if cfg != 2 else synth.make(cfg, num_points=4_000_000)
    gs = np.float32(c["grid_size"]) * c["cell_size_rad"]
    k = np.arange(0, c["num_points"], 7)
    c["uu"][k] = (-(np.round(c["uu"][k] * gs) + 0.5) / gs).astype(np.float32)
    k = np.arange(3, c["num_points"], 11)
    c["vv"][k] = ((np.round(c["vv"][k] * gs) + 0.375) / gs).astype(np.float32)
    ref = oracle.select_f(c)
    got = ops.select(c["support"], c["oversample"], dev(torch, c["uu"]), dev(torch, c["vv"]),
                     dev(torch, c["ww"]), c["cell_size_rad"], c["w_scale"], c["grid_size"])
    for k in FIELDS:
        assert np.array_equal(got[k].cpu().numpy(), ref[k]), k


# ---------------------------------------------------------------------------
# drop-in (host pointers) against the reference's own outputs
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", golden_names())
def test_dropin_f_against_reference_fixture(gpu, name):
    c = load_golden(name)
    G = c["grid_size"]
    args = (c["num_w_planes"], c["support"], c["oversample"], c["conv_size_half"],
            c["conv_func"], c["num_points"], c["uu"], c["vv"], c["ww"], c["vis"], c["weight"],
            c["cell_size_rad"], c["w_scale"], G)
    if "grid0" in c:
        g = c["grid0"].copy()
        s, n = wproj.grid_wproj_f(*args, float(c["norm0"]), g)
        s, n = wproj.grid_wproj_f(*args, n, g)
    else:
        g = np.zeros((G, G, 2), np.float32)
        s, n = wproj.grid_wproj_f(*args, 0.0, g)
    assert s == int(c["skipped_f"])
    assert close_scalar(n, c["norm_f"], 1e-12)
    # FP64 truth with the same selection (FP32 selection, FP64 sums).
    if "grid0" in c:
        t, _, _ = oracle.grid_fsel_d(c, grid=c["grid0"].astype(np.float64), norm=0.0)
        t, _, _ = oracle.grid_fsel_d(c, grid=t, norm=0.0)
    else:
        t, _, _ = oracle.grid_fsel_d(c)
    # Non-finite amplitudes: the same cells must be NaN / +-Inf as in the
    # reference's own grid; the finite cells are held to the usual bar.
    ok, gf, tf = nonfinite_split(g, c["grid_f"])
    assert ok
    assert errors(gf, nonfinite_split(t, c["grid_f"])[2])[0] <= 1e-5
    assert errors(gf, tf)[0] <= 1e-5


@pytest.mark.parametrize("name", golden_names())
def test_dropin_d_against_reference_fixture(gpu, name):
    c = load_golden(name)
    G = c["grid_size"]
    args = (c["num_w_planes"], c["support"], c["oversample"], c["conv_size_half"],
            c["conv_func"].astype(np.float64), c["num_points"], c["uu"].astype(np.float64),
            c["vv"].astype(np.float64), c["ww"].astype(np.float64), c["vis"].astype(np.float64),
            c["weight"].astype(np.float64), float(c["cell_size_rad"]), float(c["w_scale"]), G)
    if "grid0" in c:
        g = c["grid0"].astype(np.float64)
        s, n = wproj.grid_wproj_d(*args, float(c["norm0"]), g)
        s, n = wproj.grid_wproj_d(*args, n, g)
    else:
        g = np.zeros((G, G, 2), np.float64)
        s, n = wproj.grid_wproj_d(*args, 0.0, g)
    assert s == int(c["skipped_d"])
    assert close_scalar(n, c["norm_d"], 1e-12)
    ok, gf, tf = nonfinite_split(g, c["grid_d"])
    assert ok
    assert errors(gf, tf)[0] <= 1e-12


# ---------------------------------------------------------------------------
# device path against the FP64 oracle, config shapes scaled down
# ---------------------------------------------------------------------------
CASES = [(0, 30_000, 1024, None, None), (1, 20_000, 2048, 96, 36), (2, 20_000, 2048, 80, 72),
         (4, 30_000, 1024, 64, 32)]


@pytest.mark.parametrize("cfg,N,G,P,smax", CASES)
def test_device_grid_against_fp64_oracle(gpu, cfg, N, G, P, smax):
    torch, ops = gpu
    c = small_config(cfg, N, G, P, smax)
    t, tn, ts = oracle.grid_fsel_d(c)
    f, fn, fs = oracle.grid_f(c)
    g, n, s = run_grid(torch, ops, c)
    assert s == ts == fs
    assert abs(n - tn) <= 1e-12 * abs(tn)
    emax, el2 = errors(g, t)
    assert emax <= 1e-5 and el2 <= 1e-5, (emax, el2)
    assert errors(g, f)[0] <= 1e-5
    gd, nd, sd = run_grid(torch, ops, c, mode="fsel_d")
    assert errors(gd, t)[0] <= 1e-12 and sd == ts and abs(nd - tn) <= 1e-12 * abs(tn)
    gc, nc, sc = run_grid(torch, ops, c, mode="comp")
    ec = errors(gc, t)
    assert ec[0] <= 1e-5 and ec[1] <= el2 * 1.01 + 1e-9


# Every gridding kernel: 16-column strips with tables staged for every run
# that fits (min run 1), with no run staged (direct reads only), and the
# 64 x 64 staged / direct kernels.  Bar as above: <= 1e-5 vs the FP64 truth.
KERNELS = [(0, 1), (0, 1 << 30), (0, 4), (2, 2), (1, 2)]


@pytest.mark.parametrize("kernel,min_run", KERNELS)
def test_kernel_variants_against_fp64_oracle(gpu, kernel, min_run):
    torch, ops = gpu
    c = small_config(2, 30_000, 1024, 80, 72, seed=7)
    t, tn, ts = oracle.grid_fsel_d(c)
    ctx = ops.context()
    msr0 = ctx.get_option(_lib.OPT_MIN_STAGE_RUN)
    ctx.set_option(_lib.OPT_GRID_KERNEL, kernel)
    ctx.set_option(_lib.OPT_MIN_STAGE_RUN, min_run)
    try:
        g, n, s = run_grid(torch, ops, c, ctx=ctx)
        g2, _, _ = run_grid(torch, ops, c, ctx=ctx, visits_per_item=53)
    finally:
        ctx.set_option(_lib.OPT_GRID_KERNEL, 0)
        ctx.set_option(_lib.OPT_MIN_STAGE_RUN, msr0)
    assert s == ts and abs(n - tn) <= 1e-12 * abs(tn)
    emax, el2 = errors(g, t)
    assert emax <= 1e-5 and el2 <= 1e-5, (kernel, min_run, emax, el2)
    assert errors(g2, t)[0] <= 1e-5


CLASS_OPTS = (_lib.OPT_MIN_STAGE_RUN_W33, _lib.OPT_MIN_STAGE_RUN_W65, _lib.OPT_MIN_STAGE_RUN_W97,
              _lib.OPT_MIN_STAGE_RUN_W111)


@pytest.mark.parametrize("classes", [(1, 2, 8, 12), (12, 1, 1, 2)])
def test_per_class_stage_thresholds_against_fp64_oracle(gpu, classes):
    """Staged-run thresholds set per support class (supports up to 72: all
    four classes occur, plus tables too wide to stage): same bar."""
    torch, ops = gpu
    c = small_config(2, 30_000, 1024, 80, 72, seed=7)
    t, tn, ts = oracle.grid_fsel_d(c)
    ctx = ops.context()
    try:
        for k, v in zip(CLASS_OPTS, classes):
            ctx.set_option(k, v)
        assert [ctx.get_option(k) for k in CLASS_OPTS] == list(classes)
        g, n, s = run_grid(torch, ops, c, ctx=ctx)
    finally:
        for k in CLASS_OPTS:
            ctx.set_option(k, 0)
    assert s == ts and abs(n - tn) <= 1e-12 * abs(tn)
    emax, el2 = errors(g, t)
    assert emax <= 1e-5 and el2 <= 1e-5, (classes, emax, el2)


def test_device_fp64_path_against_oracle_d(gpu):
    torch, ops = gpu
    c = small_config(0, 20_000, 1024)
    t, tn, ts = oracle.grid_d(c)
    g, n, s = run_grid(torch, ops, c, dtype=torch.float64)
    assert s == ts and abs(n - tn) <= 1e-12 * abs(tn)
    assert errors(g, t)[0] <= 1e-12


@pytest.mark.parametrize("min_run", [1, 0])
def test_device_fp64_staging_widths_against_oracle_d(gpu, min_run):
    """FP64 through the 64x64 kernel with supports up to 40 (tables up to 81
    wide: those up to 79 fit its staging stride, 81 is read directly), every
    run staged (min run 1) or the automatic threshold: <= 1e-12 against the
    serial FP64 oracle."""
    torch, ops = gpu
    c = small_config(2, 30_000, 1024, 80, 40, seed=11)
    t, tn, ts = oracle.grid_d(c)
    ctx = ops.context()
    msr0 = ctx.get_option(_lib.OPT_MIN_STAGE_RUN)
    ctx.set_option(_lib.OPT_MIN_STAGE_RUN, min_run)
    try:
        g, n, s = run_grid(torch, ops, c, dtype=torch.float64, ctx=ctx)
    finally:
        ctx.set_option(_lib.OPT_MIN_STAGE_RUN, msr0)
    assert s == ts and abs(n - tn) <= 1e-12 * abs(tn)
    assert errors(g, t)[0] <= 1e-12


def test_deterministic_and_item_split_invariant(gpu):
    torch, ops = gpu
    c = small_config(2, 40_000, 2048, 60, 40)
    g1, n1, _ = run_grid(torch, ops, c)
    g2, n2, _ = run_grid(torch, ops, c)
    assert same_bits(g1, g2) and n1 == n2
    g3, n3, _ = run_grid(torch, ops, c, visits_per_item=37)
    assert errors(g3, g1)[0] <= 1e-5 and n3 == n1
    c1, _, _ = run_grid(torch, ops, c, mode="comp")
    c2, _, _ = run_grid(torch, ops, c, mode="comp", visits_per_item=37)
    assert same_bits(c1, run_grid(torch, ops, c, mode="comp")[0])
    assert errors(c2, c1)[0] <= 1e-6
    # which workgroup takes which item (XCD groups, counters) does not touch
    # the result: each item sums on its own, partials merge in item order
    ctx = ops.context()
    g0 = ctx.get_option(_lib.OPT_ITEM_GROUP)
    try:
        for grp in (8, 3):
            ctx.set_option(_lib.OPT_ITEM_GROUP, grp)
            gg, ng, _ = run_grid(torch, ops, c, ctx=ctx)
            assert same_bits(gg, g1) and ng == n1, grp
            gg, _, _ = run_grid(torch, ops, c, ctx=ctx, visits_per_item=37)
            assert same_bits(gg, g3), grp
    finally:
        ctx.set_option(_lib.OPT_ITEM_GROUP, g0)


def test_automatic_visits_per_item(gpu):
    """Option 0 (the default) picks the work-item size from the visibility
    count: 1024 for a call of fewer than 6 x 2048 visibilities per CU; an
    explicit size is used as given; the grid does not depend on it beyond
    FP32 rounding (items merge in order)."""
    torch, ops = gpu
    c = small_config(2, 40_000, 2048, 60, 40)
    ctx = _lib.Context(torch.cuda.current_device())
    try:
        assert ctx.get_option(_lib.OPT_VISITS_PER_ITEM) == 0
        ga, na, _ = run_grid(torch, ops, c, ctx=ctx)
        st = ctx.stats()
        assert st["visits_per_item"] == 1024, st
        assert st["num_items"] >= -(-st["num_visits"] // 1024)
        gb, nb, _ = run_grid(torch, ops, c, ctx=ctx, visits_per_item=4096)
        assert ctx.stats()["visits_per_item"] == 4096
        assert ctx.get_option(_lib.OPT_VISITS_PER_ITEM) == 0      # (restored)
        assert errors(ga, gb)[0] <= 1e-5 and abs(na - nb) <= 1e-12 * abs(nb)
        with pytest.raises(_lib.GridderError):
            ctx.set_option(_lib.OPT_VISITS_PER_ITEM, -1)
    finally:
        ctx.close()


def test_device_accumulates_across_calls(gpu):
    """grid and norm accumulate, num_skipped is set per call (oskar_grid_wproj.c:151-210)."""
    torch, ops = gpu
    c = small_config(0, 10_000, 512)
    ks, d = device_case(torch, ops, c)
    G = c["grid_size"]
    grid = torch.zeros((G, G, 2), dtype=torch.float32, device="cuda")
    norm = torch.full((1,), 5.0, dtype=torch.float64, device="cuda")
    skipped = torch.full((1,), 99, dtype=torch.int64, device="cuda")
    for _ in range(2):
        ops.grid(ks, d["uu"], d["vv"], d["ww"], d["vis"], d["weight"], c["cell_size_rad"],
                 c["w_scale"], G, grid, norm, skipped)
    t, tn, ts = oracle.grid_fsel_d(c)
    assert int(skipped.item()) == ts
    assert abs(norm.item() - (5.0 + 2 * tn)) <= 1e-12 * abs(2 * tn)
    assert errors(grid.cpu().numpy(), 2 * t)[0] <= 1e-5


@pytest.mark.parametrize("G,smax", [(1001, 24), (130, 12), (300, 36)])
def test_grid_sizes_off_the_region_grid(gpu, G, smax):
    """Grid sizes that are no multiple of the strip regions (256 x 128, and
    192 x 64 compensated) or smaller than one region: edge regions are cut
    by the grid, and footprints that leave it are skipped (counted
    bit-exactly).  Plain and compensated against the FP64 truth."""
    torch, ops = gpu
    c = small_config(0, 12_000, G, 32, smax, seed=5)
    t, tn, ts = oracle.grid_fsel_d(c)
    for mode in ("plain", "comp"):
        g, n, s = run_grid(torch, ops, c, mode=mode)
        assert s == ts and abs(n - tn) <= 1e-12 * abs(tn), mode
        emax, el2 = errors(g, t)
        assert emax <= 1e-5 and el2 <= 1e-5, (mode, emax, el2)


def test_many_planes_shortened_table_ids_against_fp64_oracle(gpu):
    """4000 w-planes at oversample 4: 36000 kernel tables, more than the 15
    bits of table id a visit record holds, so the ids are shortened for the
    sort (runs then compare table bases) and the direct reads do not fold
    symmetric tables.  Strip kernel and 64 x 64 staged kernel against the
    FP64 truth; num_skipped bit-exact, norm to 1e-12."""
    torch, ops = gpu
    c = small_config(2, 20_000, 512, 4000, 8, seed=5)
    assert len(c["support"]) == 4000
    t, tn, ts = oracle.grid_fsel_d(c)
    ctx = ops.context()
    try:
        for kernel in (0, 2):
            ctx.set_option(_lib.OPT_GRID_KERNEL, kernel)
            g, n, s = run_grid(torch, ops, c, ctx=ctx)
            assert s == ts and abs(n - tn) <= 1e-12 * abs(tn), (kernel, s, ts, n, tn)
            emax, el2 = errors(g, t)
            assert emax <= 1e-5 and el2 <= 1e-5, (kernel, emax, el2)
    finally:
        ctx.set_option(_lib.OPT_GRID_KERNEL, 0)


@pytest.mark.parametrize("smax,N", [(100, 4000), (160, 3000)])
def test_large_supports_against_fp64_oracle(gpu, smax, N):
    """Supports beyond the shapes above (16 planes, 4 .. smax): footprints up
    to 201 and 321 wide -- wider than any staged table (only direct reads in
    the strip kernel), taller than a 128-row region and, at 321, wider than
    a 256-column one, so a visit is bucketed into several regions in both
    directions.  Every kernel (strips plain and compensated, 64 x 64 direct
    and staged) and the FP32-selection / FP64 path against the FP64 truth;
    num_skipped bit-exact, norm to 1e-12."""
    torch, ops = gpu
    c = small_config(0, N, 1024, 16, smax, seed=11)
    assert int(np.max(c["support"])) == smax
    t, tn, ts = oracle.grid_fsel_d(c)
    ctx = ops.context()
    try:
        for kernel, mode in ((0, "plain"), (0, "comp"), (1, "plain"), (2, "plain")):
            ctx.set_option(_lib.OPT_GRID_KERNEL, kernel)
            g, n, s = run_grid(torch, ops, c, mode=mode, ctx=ctx)
            assert s == ts and abs(n - tn) <= 1e-12 * abs(tn), (kernel, mode, s, ts, n, tn)
            emax, el2 = errors(g, t)
            assert emax <= 1e-5 and el2 <= 1e-5, (kernel, mode, emax, el2)
    finally:
        ctx.set_option(_lib.OPT_GRID_KERNEL, 0)
    gd, nd, sd = run_grid(torch, ops, c, mode="fsel_d")
    assert errors(gd, t)[0] <= 1e-12 and sd == ts and abs(nd - tn) <= 1e-12 * abs(tn)
