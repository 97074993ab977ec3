"""Batched gridding (many grids, one pipeline pass) against single calls.

Each job's sorted visit lists, work-item splits and per-chunk norm partials
are the same as in a single call, so the bar is bit-identity.
"""
import numpy as np
import pytest

from conftest import same_bits
from mi355_gridding import synth

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
