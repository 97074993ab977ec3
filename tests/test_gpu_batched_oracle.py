"""The batched workload at its real shape (4096^2 grids, 339 w-planes,
support 44, oversample 4) against the oracle.

Each job stands for one block call of oskar_grid_wproj_f (main.cpp:243-381,
oskar_grid_wproj.c:126-212) on its own grid.  Bars, per job, for the eager
batched pass and for a replay of the captured graph:
  - num_skipped bit-exact against the oracle's FP32 selection count;
  - norm within 1e-12 (relative) of the oracle's serial FP32 norm
    (oracle_norm_f, bit-identical to oskar_grid_wproj_f's);
  - the densest 64x64 tile and three sampled occupied tiles within 1e-5 of
    the tile's largest value against the oracle's FP64 tile truth with the
    same selection (oracle_grid_wproj_fsel_d_tiles);
  - the graph replay bit-identical to the eager pass.
The full 64-job batch (log-uniform counts in [50k, 2M]) is checked through
properties: two replays are bitwise identical, and a replay with doubled
amplitudes doubles every grid bit for bit (scaling by 2 is exact in every
product and sum).
"""
import numpy as np
import pytest

from conftest import errors
from oracle import oracle
from mi355_gridding import synth

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

# at least one job of >= 1M visibilities and one of ~50k
COUNTS = [1_200_000, 50_000, 180_000, 420_000, 75_000, 900_000, 130_000, 60_000]
T = 64


def _setup(torch, ops, b):
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


def _case(b, j):
    """Job j of batch b as a single-call case for the oracle."""
    job = b["jobs"][j]
    return dict(num_w_planes=b["num_w_planes"], support=b["support"],
                oversample=b["oversample"], conv_size_half=b["conv_size_half"],
                conv_func=b["conv_func"], num_points=len(job["uu"]), uu=job["uu"],
                vv=job["vv"], ww=job["ww"], vis=job["vis"], weight=job["weight"],
                cell_size_rad=job["cell_size_rad"], w_scale=job["w_scale"],
                grid_size=job["grid_size"])


@pytest.fixture(scope="module")
def batch8(gpu):
    torch, ops = gpu
    b = synth.make_batched(counts=COUNTS)
    assert b["num_w_planes"] == 339 and b["max_support"] == 44 and b["oversample"] == 4
    ks, cat, jobs = _setup(torch, ops, b)
    # oracle side, per job: skipped count, serial norm, tiles to check
    truth = []
    rng = np.random.default_rng(33)
    for j in range(len(COUNTS)):
        c = _case(b, j)
        G = c["grid_size"]
        sel = oracle.select_f(c)
        keep = sel["skip"] == 0
        tid = (sel["grid_v"][keep] // T) * (G // T) + sel["grid_u"][keep] // T
        counts = np.bincount(tid, minlength=(G // T) ** 2)
        occ = np.flatnonzero(counts)
        picks = [int(np.argmax(counts))] + [int(x) for x in rng.choice(occ, 3, replace=False)]
        tiles = np.array([[p % (G // T), p // (G // T)] for p in picks], np.int32)
        norm, skipped = oracle.norm_f(c)
        assert skipped == int(np.count_nonzero(sel["skip"]))
        truth.append(dict(skipped=skipped, norm=norm, tiles=tiles,
                          ref=oracle.tiles_fsel_d(c, tiles, tile=T, threads=8)))
    return torch, ops, b, ks, cat, jobs, truth


def _check_jobs(torch, jobs, truth):
    for j, (job, t) in enumerate(zip(jobs, truth)):
        assert job["num_skipped"].item() == t["skipped"], j
        n = job["norm"].item()
        assert abs(n - t["norm"]) <= 1e-12 * abs(t["norm"]), (j, n, t["norm"])
        g = job["grid"]
        for (tx, ty), r in zip(t["tiles"], t["ref"]):
            got = g[ty * T:(ty + 1) * T, tx * T:(tx + 1) * T].cpu().numpy()
            e = errors(got, r)[0]
            assert e <= 1e-5, (j, tx, ty, e)


def test_batched_config_shape_vs_oracle(batch8):
    torch, ops, b, ks, cat, jobs, truth = batch8
    args = (ks, jobs, cat["uu"], cat["vv"], cat["ww"], cat["vis"], cat["weight"])
    _zero(jobs)
    ops.grid_batched(*args)
    torch.cuda.synchronize()
    _check_jobs(torch, jobs, truth)
    eager = [job["grid"].clone() for job in jobs]
    for rep in range(2):              # build (eager sync-free run), then a replay
        _zero(jobs)
        ops.grid_batched_graph(*args)
        torch.cuda.synchronize()
        _check_jobs(torch, jobs, truth)
        for j, job in enumerate(jobs):
            assert torch.equal(job["grid"], eager[j]), (rep, j)


def test_batched_64_jobs_replay_properties(gpu):
    torch, ops = gpu
    b = synth.make_batched()          # 64 jobs, counts log-uniform in [50k, 2M]
    assert len(b["jobs"]) == 64
    ks, cat, jobs = _setup(torch, ops, b)
    args = (ks, jobs, cat["uu"], cat["vv"], cat["ww"], cat["vis"], cat["weight"])
    runs = []
    for rep in range(3):              # build, replay, replay
        _zero(jobs)
        ops.grid_batched_graph(*args)
        torch.cuda.synchronize()
        if rep:
            runs.append([job["grid"].clone() for job in jobs])
        skipped = [job["num_skipped"].item() for job in jobs]
        assert min(skipped) >= 0
    for a, c in zip(*runs):
        assert torch.equal(a, c)
    cat["vis"].mul_(2)
    _zero(jobs)
    ops.grid_batched_graph(*args)
    torch.cuda.synchronize()
    for j, (a, job) in enumerate(zip(runs[0], jobs)):
        assert torch.equal(job["grid"], a * 2), j
    del runs
