"""Bucketing by grid region, pinned directly: the sorted (region, kernel
table) keys, the visibility of every sorted visit, the per-region visit
counts and the staged / run-start classification that the strip kernel
reads, all bit-exact against a restatement built from the oracle's FP32
selection (oskar_grid_wproj.c:158-187).

The restatement: every gridded visibility i (in order) visits the 256 x 128
regions its footprint [u - s, u + s] x [v - s, v + s] touches, rows of
regions outer and columns inner; its key is (region << bits) | table, with
table = plane * 9 + |off_u| * 3 + |off_v| (oversample 4: |off| <= 2); the
visits are sorted stably by key.  A run is a maximal stretch of equal keys;
its visits are staged when it has at least min_stage_run of them and the
table is at most 111 wide; the threshold may differ per support class
(tables at most 33, 65, 97, 111 wide).  This replaces the per-tile count and scatter of
oskar_grid_wproj_gpu.cu:359-649, so grid parity alone would not catch, for
example, a visit dropped from one region and duplicated in another.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle
from mi355_gridding import _lib, synth

pytestmark = pytest.mark.gpu

RW, RH, MAXW = 256, 128, 111


def restate(c, min_runs):
    sel = oracle.select_f(c)
    G = c["grid_size"]
    keep = np.flatnonzero(sel["skip"] == 0)
    s = sel["support"][keep].astype(np.int64)
    gu = sel["grid_u"][keep].astype(np.int64)
    gv = sel["grid_v"][keep].astype(np.int64)
    tx0, tx1 = (gu - s) // RW, (gu + s) // RW
    ty0, ty1 = (gv - s) // RH, (gv + s) // RH
    nx, ny = tx1 - tx0 + 1, ty1 - ty0 + 1
    cnt = nx * ny
    vis = np.repeat(keep, cnt)
    k = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)   # index within visit
    rnx = np.repeat(nx, cnt)
    ty = np.repeat(ty0, cnt) + k // rnx
    tx = np.repeat(tx0, cnt) + k % rnx
    ntx = (G + RW - 1) // RW
    region = ty * ntx + tx
    table = (sel["plane"][vis].astype(np.int64) * 9 + np.abs(sel["off_u"][vis]) * 3 +
             np.abs(sel["off_v"][vis]))
    ntables = c["num_w_planes"] * 9
    bits = int(np.ceil(np.log2(ntables))) if ntables > 1 else 0
    key = (region << bits) | table
    order = np.argsort(key, kind="stable")
    key, vis = key[order], vis[order]
    start = np.r_[True, key[1:] != key[:-1]]
    run_id = np.cumsum(start) - 1
    run_len = np.bincount(run_id)[run_id]
    W = 2 * sel["support"][vis] + 1
    thr = np.select([W <= 33, W <= 65, W <= 97], list(min_runs[:3]), min_runs[3])
    staged = (run_len >= thr) & (W <= MAXW)
    flags = staged.astype(np.int32) | (start.astype(np.int32) << 1)
    return key.astype(np.uint32), vis.astype(np.uint32), flags, bits


CLASS_OPTS = (_lib.OPT_MIN_STAGE_RUN_W33, _lib.OPT_MIN_STAGE_RUN_W65, _lib.OPT_MIN_STAGE_RUN_W97,
              _lib.OPT_MIN_STAGE_RUN_W111)


@pytest.mark.parametrize("cfg,N,G,P,smax,seed,classes", [(2, 30_000, 1024, 80, 72, 7, None),
                                                        (1, 60_000, 2048, 714, 36, 8, None),
                                                        (2, 30_000, 1024, 80, 72, 9, (1, 2, 6, 12))])
def test_buckets_bitexact_against_selection(gpu, cfg, N, G, P, smax, seed, classes):
    torch, ops = gpu
    c = synth.make(cfg, num_points=N, grid_size=G, num_w_planes=P, max_support=smax, seed=seed)
    ks = ops.KernelStack(c["support"], c["oversample"], c["conv_size_half"], c["conv_func"])
    d = {k: torch.from_numpy(np.ascontiguousarray(c[k])).cuda()
         for k in ("uu", "vv", "ww", "vis", "weight")}
    grid = torch.zeros((G, G, 2), dtype=torch.float32, device="cuda")
    norm = torch.zeros(1, dtype=torch.float64, device="cuda")
    sk = torch.zeros(1, dtype=torch.int64, device="cuda")
    ctx = ops.context()
    base = ctx.get_option(_lib.OPT_MIN_STAGE_RUN)
    try:
        for k, v in zip(CLASS_OPTS, classes or (0, 0, 0, 0)):
            ctx.set_option(k, v)
        ops.grid(ks, d["uu"], d["vv"], d["ww"], d["vis"], d["weight"], c["cell_size_rad"],
                 c["w_scale"], G, grid, norm, sk, ctx=ctx)
        torch.cuda.synchronize()
    finally:
        for k in CLASS_OPTS:
            ctx.set_option(k, 0)
    # (option 0: the automatic thresholds, 6 for tables up to 65 wide, 3 above)
    min_runs = classes or ((base,) * 4 if base else (6, 6, 3, 3))
    lib = _lib.load()
    n = ctypes.c_ulonglong(0)
    bits = ctypes.c_int(0)
    _lib.check(lib.mi355_ctx_get_buckets(ctx.handle, ctypes.byref(n), ctypes.byref(bits), None,
                                         None, None, 0))
    keys = np.zeros(n.value, np.uint32)
    vis = np.zeros(n.value, np.uint32)
    flags = np.zeros(n.value, np.int32)
    _lib.check(lib.mi355_ctx_get_buckets(ctx.handle, ctypes.byref(n), ctypes.byref(bits),
                                         ctypes.c_void_p(keys.ctypes.data),
                                         ctypes.c_void_p(vis.ctypes.data),
                                         ctypes.c_void_p(flags.ctypes.data), n.value))
    rkey, rvis, rflags, rbits = restate(c, min_runs)
    assert bits.value == rbits
    assert n.value == len(rkey)
    assert np.array_equal(keys, rkey)
    assert np.array_equal(vis, rvis)
    assert np.array_equal(flags, rflags)
    # per-region visit counts
    nreg = ((G + RW - 1) // RW) * ((G + RH - 1) // RH)
    assert np.array_equal(np.bincount(keys >> bits, minlength=nreg),
                          np.bincount(rkey >> rbits, minlength=nreg))
    # both staged and direct visits occur, so the classification is exercised
    assert 0 < np.count_nonzero(flags & 1) < len(flags)
