"""Batched gridding (many grids, one pipeline pass) against single calls.

Each job's sorted visit lists, work-item splits and per-chunk norm partials
are the same as in a single call, so the bar is bit-identity.
"""
import numpy as np
import pytest

from conftest import same_bits
from mi355_gridding import _lib, synth

pytestmark = pytest.mark.gpu


def test_batched_equals_sequential(gpu):
    torch, ops = gpu
    b = synth.make_batched(num_jobs=6, grid_size=1024, num_w_planes=64, max_support=20,
                           n_min=2_000, n_max=40_000)
    b["jobs"][2]["grid_size"] = 2048          # grids of different sizes in one pass
    ks = ops.KernelStack(b["support"], b["oversample"], b["conv_size_half"], b["conv_func"])
    cat = {k: torch.from_numpy(np.concatenate([j[k] for j in b["jobs"]])).cuda()
           for k in ("uu", "vv", "ww", "vis", "weight")}
    jobs, off = [], 0
    for j in b["jobs"]:
        G = j["grid_size"]
        n = len(j["uu"])
        jobs.append(dict(vis_offset=off, num_points=n, cell_size_rad=j["cell_size_rad"],
                         w_scale=j["w_scale"], grid_size=G,
                         grid=torch.zeros((G, G, 2), dtype=torch.float32, device="cuda"),
                         norm=torch.zeros(1, dtype=torch.float64, device="cuda"),
                         num_skipped=torch.zeros(1, dtype=torch.int64, device="cuda")))
        off += n
    ops.grid_batched(ks, jobs, cat["uu"], cat["vv"], cat["ww"], cat["vis"], cat["weight"])
    for j, job in zip(b["jobs"], jobs):
        G = j["grid_size"]
        d = {k: torch.from_numpy(np.ascontiguousarray(j[k])).cuda()
             for k in ("uu", "vv", "ww", "vis", "weight")}
        g = torch.zeros((G, G, 2), dtype=torch.float32, device="cuda")
        n = torch.zeros(1, dtype=torch.float64, device="cuda")
        s = torch.zeros(1, dtype=torch.int64, device="cuda")
        ops.grid(ks, d["uu"], d["vv"], d["ww"], d["vis"], d["weight"], j["cell_size_rad"],
                 j["w_scale"], G, g, n, s)
        assert same_bits(g.cpu().numpy(), job["grid"].cpu().numpy())
        assert n.item() == job["norm"].item()
        assert s.item() == job["num_skipped"].item()


def _batched_case(torch, ops, num_jobs=5, seed=3):
    b = synth.make_batched(num_jobs=num_jobs, grid_size=1024, num_w_planes=64, max_support=20,
                           n_min=2_000, n_max=30_000, seed=seed)
    b["jobs"][1]["grid_size"] = 2048
    ks = ops.KernelStack(b["support"], b["oversample"], b["conv_size_half"], b["conv_func"])
    cat = {k: torch.from_numpy(np.concatenate([j[k] for j in b["jobs"]])).cuda()
           for k in ("uu", "vv", "ww", "vis", "weight")}
    jobs, off = [], 0
    for j in b["jobs"]:
        G, n = j["grid_size"], len(j["uu"])
        jobs.append(dict(vis_offset=off, num_points=n, cell_size_rad=j["cell_size_rad"],
                         w_scale=j["w_scale"], grid_size=G,
                         grid=torch.zeros((G, G, 2), dtype=torch.float32, device="cuda"),
                         norm=torch.zeros(1, dtype=torch.float64, device="cuda"),
                         num_skipped=torch.zeros(1, dtype=torch.int64, device="cuda")))
        off += n
    return ks, cat, jobs


def _zero(jobs):
    for job in jobs:
        job["grid"].zero_()
        job["norm"].zero_()
        job["num_skipped"].fill_(-1)


def _snap(jobs):
    return [(job["grid"].cpu().numpy().copy(), job["norm"].item(), job["num_skipped"].item())
            for job in jobs]


@pytest.mark.parametrize("compensated", [0, 1])
def test_graph_replays_equal_eager(gpu, compensated):
    """Graph bar: every replay, with eager work and host synchronisations in
    between, equals the eager (non-graph) batched pass bit for bit; so does
    a replay after the inputs' values change in place.  Plain and
    compensated accumulation (each its own strip kernel geometry)."""
    torch, ops = gpu
    ks, cat, jobs = _batched_case(torch, ops)
    args = (ks, jobs, cat["uu"], cat["vv"], cat["ww"], cat["vis"], cat["weight"])
    ctx = ops.context()
    ctx.set_option(_lib.OPT_COMPENSATED, compensated)
    try:
        _graph_replays(torch, ops, ctx, args, cat, jobs)
    finally:
        ctx.set_option(_lib.OPT_COMPENSATED, 0)


def _graph_replays(torch, ops, ctx, args, cat, jobs):
    _zero(jobs)
    ops.grid_batched(*args, ctx=ctx)
    torch.cuda.synchronize()
    ref = _snap(jobs)
    for rep in range(4):              # first call builds, the others replay
        _zero(jobs)
        ops.grid_batched_graph(*args, ctx=ctx)
        torch.cuda.synchronize()
        got = _snap(jobs)
        for (g0, n0, s0), (g1, n1, s1) in zip(ref, got):
            assert same_bits(g0, g1) and n0 == n1 and s0 == s1, rep
        # unrelated eager work between replays
        x = torch.randn(1 << 20, device="cuda")
        _ = (x * 2).sum().item()
    st = ctx.stats()
    assert st["num_points"] == cat["uu"].numel() and st["num_visits"] > 0
    # new values in the same arrays: the replay grids them
    cat["vis"].mul_(2)
    _zero(jobs)
    ops.grid_batched_graph(*args, ctx=ctx)
    torch.cuda.synchronize()
    for (g0, n0, s0), job in zip(ref, jobs):
        assert same_bits(g0 * 2, job["grid"].cpu().numpy()) and s0 == job["num_skipped"].item()


TUNED = (_lib.OPT_GRID_KERNEL, _lib.OPT_VISITS_PER_ITEM, _lib.OPT_MIN_STAGE_RUN,
         _lib.OPT_ITEM_GROUP, _lib.OPT_MIN_STAGE_RUN_W33, _lib.OPT_MIN_STAGE_RUN_W65,
         _lib.OPT_MIN_STAGE_RUN_W97, _lib.OPT_MIN_STAGE_RUN_W111)


def test_autotune_keeps_results_and_sets_options(gpu):
    """The trials write private grids (the caller's outputs stay untouched);
    the choice is cached per workload shape: a second call with the same
    shape (other values) returns it without trials, a call with the cache
    switched off runs them again; a different shape is tuned afresh."""
    torch, ops = gpu
    ks, cat, jobs = _batched_case(torch, ops, num_jobs=3, seed=4)
    args = (ks, jobs, cat["uu"], cat["vv"], cat["ww"], cat["vis"], cat["weight"])
    ctx = _lib.Context(torch.cuda.current_device())
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    _zero(jobs)
    import time
    t0 = time.perf_counter()
    best = ops.autotune_batched(*args, ctx=ctx)
    t_first = time.perf_counter() - t0
    assert best > 0
    chosen = [ctx.get_option(k) for k in TUNED]
    # scramble the options: the cached answer puts the choice back
    ctx.set_option(_lib.OPT_VISITS_PER_ITEM, 333)
    ctx.set_option(_lib.OPT_MIN_STAGE_RUN_W65, 7)
    cat["vis"].mul_(3)                        # same shape, other values
    t0 = time.perf_counter()
    best2 = ops.autotune_batched(*args, ctx=ctx)
    t_cached = time.perf_counter() - t0
    assert best2 == best and [ctx.get_option(k) for k in TUNED] == chosen
    assert t_cached < 0.25 * t_first, (t_cached, t_first)
    ctx.set_option(_lib.OPT_AUTOTUNE_CACHE, 0)
    best3 = ops.autotune_batched(*args, ctx=ctx)
    assert best3 > 0 and best3 != best
    ctx.set_option(_lib.OPT_AUTOTUNE_CACHE, 1)
    # the caller's outputs were not touched by the trials
    for job in jobs:
        assert job["grid"].abs().max().item() == 0 and job["num_skipped"].item() == -1
    ops.grid_batched(*args, ctx=ctx)
    torch.cuda.synchronize()
    for job in jobs:
        assert job["num_skipped"].item() >= 0
    ctx.close()


def test_graph_survives_other_calls_on_the_context(gpu):
    """A captured batched graph is private to its context's graph state: single
    grid / degrid / selection calls and the autotuner on the same context in
    between (which rewrite the context's own job table, support and plane
    offsets, point jobs at other grids and grow or free buffers) leave later
    replays equal to the eager pass bit for bit."""
    torch, ops = gpu
    ks, cat, jobs = _batched_case(torch, ops, num_jobs=4, seed=6)
    args = (ks, jobs, cat["uu"], cat["vv"], cat["ww"], cat["vis"], cat["weight"])
    ctx = ops.context()
    _zero(jobs)
    ops.grid_batched(*args, ctx=ctx)
    torch.cuda.synchronize()
    ref = _snap(jobs)

    def check():
        _zero(jobs)
        ops.grid_batched_graph(*args, ctx=ctx)
        torch.cuda.synchronize()
        for (g0, n0, s0), (g1, n1, s1) in zip(ref, _snap(jobs)):
            assert same_bits(g0, g1) and n0 == n1 and s0 == s1

    check()                            # build
    check()                            # replay
    # other entry points on the same context, with a different (larger) stack
    b = synth.make_batched(num_jobs=1, grid_size=2048, num_w_planes=96, max_support=28,
                           n_min=60_000, n_max=60_001, seed=11)
    j = b["jobs"][0]
    ks2 = ops.KernelStack(b["support"], b["oversample"], b["conv_size_half"], b["conv_func"])
    d = {k: torch.from_numpy(np.ascontiguousarray(j[k])).cuda()
         for k in ("uu", "vv", "ww", "vis", "weight")}
    G = j["grid_size"]
    g = torch.zeros((G, G, 2), dtype=torch.float32, device="cuda")
    n = torch.zeros(1, dtype=torch.float64, device="cuda")
    s = torch.zeros(1, dtype=torch.int64, device="cuda")
    ops.grid(ks2, d["uu"], d["vv"], d["ww"], d["vis"], d["weight"], j["cell_size_rad"],
             j["w_scale"], G, g, n, s, ctx=ctx)
    check()
    ops.degrid(ks2, d["uu"], d["vv"], d["ww"], d["weight"], j["cell_size_rad"], j["w_scale"],
               G, g, ctx=ctx)
    ops.select(b["support"], b["oversample"], d["uu"], d["vv"], d["ww"], j["cell_size_rad"],
               j["w_scale"], G, ctx=ctx)
    check()
    # the autotuner runs trials on the context and leaves its choice in the
    # options; with the options put back, the key matches the graph built
    # above, which must replay unchanged
    keys = (_lib.OPT_GRID_KERNEL, _lib.OPT_VISITS_PER_ITEM, _lib.OPT_MIN_STAGE_RUN,
            _lib.OPT_ITEM_GROUP, _lib.OPT_MIN_STAGE_RUN_W33, _lib.OPT_MIN_STAGE_RUN_W65,
            _lib.OPT_MIN_STAGE_RUN_W97, _lib.OPT_MIN_STAGE_RUN_W111)
    saved = [ctx.get_option(k) for k in keys]
    ops.autotune_batched(*args, ctx=ctx)
    for k, v in zip(keys, saved):
        ctx.set_option(k, v)
    check()
    assert ctx.stats()["num_points"] == cat["uu"].numel()


def test_workspace_cap_refuses_before_allocating(gpu):
    """MI355_OPT_MAX_WORKSPACE: a call whose buffers would exceed the cap
    fails with MI355_ERR_OUT_OF_MEMORY (forced from the host, no device
    allocation is attempted) and leaves the outputs alone; lifting the cap
    lets the same call through."""
    torch, ops = gpu
    ks, cat, jobs = _batched_case(torch, ops, num_jobs=2, seed=7)
    ctx = _lib.Context(torch.cuda.current_device())
    try:
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        ctx.set_option(_lib.OPT_MAX_WORKSPACE, 1 << 20)
        assert ctx.get_option(_lib.OPT_MAX_WORKSPACE) == 1 << 20
        _zero(jobs)
        with pytest.raises(_lib.GridderError) as e:
            ops.grid_batched(ks, jobs, cat["uu"], cat["vv"], cat["ww"], cat["vis"],
                             cat["weight"], ctx=ctx)
        assert e.value.status == _lib.ERR_OUT_OF_MEMORY and "cap" in str(e.value)
        torch.cuda.synchronize()
        assert all(job["grid"].abs().max().item() == 0 for job in jobs)
        ctx.set_option(_lib.OPT_MAX_WORKSPACE, 0)
        ops.grid_batched(ks, jobs, cat["uu"], cat["vv"], cat["ww"], cat["vis"], cat["weight"],
                         ctx=ctx)
        torch.cuda.synchronize()
        assert all(job["num_skipped"].item() >= 0 for job in jobs)
    finally:
        ctx.close()
