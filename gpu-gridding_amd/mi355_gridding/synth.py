"""Seeded synthetic workloads with the shapes of the reference's data sets.

The generator itself is C++ (csrc/synth.cpp); this module fixes the five
workload shapes and returns numpy arrays laid out as the reference's
gridding call takes them (main.cpp:340-380).
"""
import ctypes

import numpy as np

from . import _lib

SEED_BASE = 0x4D49333535

# name, N, baselines, times, G, P, smax, oversample, radius law (0 exp, 1 log-normal)
CONFIGS = {
    0: dict(name="serial-cpu-200k", N=200_000, baselines=0, times=1, G=2048, P=64,
            smax=16, ov=4, dist=0),
    1: dict(name="synthetic-challenge-2.15M", N=2_152_800, baselines=21_528, times=100,
            G=8192, P=714, smax=36, ov=4, dist=0),
    2: dict(name="ska-low-el56-82-31.4M", N=31_395_840, baselines=130_816, times=240,
            G=16384, P=601, smax=72, ov=4, dist=1),
    3: dict(name="batched-64x4096", N=None, baselines=0, times=1, G=4096, P=339,
            smax=44, ov=4, dist=0, jobs=64),
    4: dict(name="differentiable-1M", N=1_000_000, baselines=0, times=1, G=4096, P=128,
            smax=32, ov=4, dist=0),
}


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def kernels(num_w_planes, max_support, oversample=4, conv_size_half=None):
    """(support int32 [P], conv float32 [P, csh, csh, 2], csh)."""
    lib = _lib.load_synth()
    csh = conv_size_half or (max_support + 1) * oversample
    support = np.zeros(num_w_planes, np.int32)
    conv = np.zeros((num_w_planes, csh, csh, 2), np.float32)
    if lib.mi355_synth_kernels(num_w_planes, max_support, oversample, csh, _p(support), _p(conv)):
        raise ValueError("bad kernel parameters")
    return support, conv, csh


def visibilities(seed, num_points, num_w_planes, dist=0, baselines=0, times=1):
    """dict(uu, vv, ww, vis[N,2], weight, w_scale) for a seed."""
    lib = _lib.load_synth()
    n = int(num_points)
    uu = np.empty(n, np.float32)
    vv = np.empty(n, np.float32)
    ww = np.empty(n, np.float32)
    vis = np.empty((n, 2), np.float32)
    weight = np.empty(n, np.float32)
    ws = ctypes.c_double(0.0)
    if lib.mi355_synth_vis(int(seed), int(dist), n, int(baselines), int(times),
                           int(num_w_planes), _p(uu), _p(vv), _p(ww), _p(vis), _p(weight),
                           ctypes.byref(ws)):
        raise ValueError("bad visibility parameters")
    return dict(uu=uu, vv=vv, ww=ww, vis=vis, weight=weight, w_scale=ws.value)


def cell_size(grid_size, max_support):
    return _lib.load_synth().mi355_synth_cell_size(int(grid_size), int(max_support))


def make(config, num_points=None, grid_size=None, num_w_planes=None, max_support=None,
         seed=None):
    """A complete single-grid workload for ``config`` (0, 1, 2 or 4).

    Any size may be overridden to get a smaller case of the same shape; with
    num_points overridden, track layouts fall back to independent draws.
    """
    c = dict(CONFIGS[config])
    N = int(num_points or c["N"])
    G = int(grid_size or c["G"])
    P = int(num_w_planes or c["P"])
    smax = int(max_support or c["smax"])
    ov = c["ov"]
    tracks = num_points is None and c["times"] > 1
    support, conv, csh = kernels(P, smax, ov)
    v = visibilities(SEED_BASE + config if seed is None else seed, N, P, c["dist"],
                     c["baselines"] if tracks else 0, c["times"] if tracks else 1)
    cell = cell_size(G, smax)
    return dict(name=c["name"], num_w_planes=P, support=support, oversample=ov,
                conv_size_half=csh, conv_func=conv, num_points=N, uu=v["uu"], vv=v["vv"],
                ww=v["ww"], vis=v["vis"], weight=v["weight"],
                cell_size_rad=np.float32(cell), w_scale=np.float32(v["w_scale"]),
                grid_size=G, max_support=smax)


def make_batched(num_jobs=64, grid_size=4096, num_w_planes=339, max_support=44,
                 n_min=50_000, n_max=2_000_000, seed=SEED_BASE + 3, counts=None):
    """The batched workload: num_jobs independent grids sharing one kernel stack, with
    visibility counts log-uniform in [n_min, n_max] (or the given counts)."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(n_min), np.log(n_max), num_jobs)).astype(np.int64) \
        if counts is None else np.asarray(counts, np.int64)
    num_jobs = len(counts)
    support, conv, csh = kernels(num_w_planes, max_support, 4)
    cell = cell_size(grid_size, max_support)
    jobs = []
    for j in range(num_jobs):
        v = visibilities(seed * 1000 + j, int(counts[j]), num_w_planes, 0)
        v["cell_size_rad"] = np.float32(cell)
        v["w_scale"] = np.float32(v["w_scale"])
        v["grid_size"] = grid_size
        jobs.append(v)
    return dict(num_w_planes=num_w_planes, support=support, oversample=4,
                conv_size_half=csh, conv_func=conv, jobs=jobs, max_support=max_support)
