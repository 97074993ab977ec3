"""Uniform-weighting pre-pass (oskar_grid_weights.c) on the GPU against the
oracle and the reference's fixtures, through the C ABI.

Bar: bit-exact grid, weight_out and num_skipped, in FP32 and FP64 (NaN
compared by position only).  The GPU adds each cell's weights in point
order -- the float additions of the serial loop, in its order
(oskar_grid_weights.c:131 / :62) -- and divides with IEEE division
(:163 / :94), so nothing is left to a tolerance.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_names, same_bits_or_nan
from oracle import oracle
from mi355_gridding import synth, wproj, _lib

pytestmark = pytest.mark.gpu


def _golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _oracle_run(blocks, read_block, real, cell, G, grid0=None):
    grid = np.zeros((G, G), real) if grid0 is None else grid0.copy()
    sk = [oracle.weights_write(u, v, w, cell, G, grid) for u, v, w in blocks]
    u, v, w = read_block
    wo = np.full(u.size, real(-7.0))
    sk.append(oracle.weights_read(u, v, w, wo, cell, G, grid))
    return grid, wo, sk


def _dev_run(torch, ops, blocks, read_block, real, cell, G, grid0=None):
    tdt = torch.float32 if real == np.float32 else torch.float64
    grid = torch.zeros((G, G), dtype=tdt, device="cuda") if grid0 is None else \
        torch.from_numpy(grid0).cuda()
    sk = torch.full((1,), 99, dtype=torch.int64, device="cuda")
    out = []
    for u, v, w in blocks:
        d = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (u, v, w)]
        ops.grid_weights_write(d[0], d[1], d[2], cell, G, grid, sk)
        out.append(int(sk.item()))
    u, v, w = (torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in read_block)
    wo = torch.full((u.numel(),), -7.0, dtype=tdt, device="cuda")
    ops.grid_weights_read(u, v, w, wo, cell, G, grid, sk)
    out.append(int(sk.item()))
    torch.cuda.synchronize()
    return grid.cpu().numpy(), wo.cpu().numpy(), out


@pytest.mark.parametrize("name", golden_names("weights_"))
@pytest.mark.parametrize("tag,real", [("f", np.float32), ("d", np.float64)])
def test_drop_ins_match_reference_fixture(gpu, name, tag, real):
    c = _golden(name)
    G = int(c["grid_size"])
    cell = real(c["cell_size_rad"])
    write = wproj.grid_weights_write_f if tag == "f" else wproj.grid_weights_write_d
    read = wproj.grid_weights_read_f if tag == "f" else wproj.grid_weights_read_d
    grid = np.zeros((G, G), real)
    sk = [write(c["uu_" + b].size, c["uu_" + b].astype(real), c["vv_" + b].astype(real),
                c["w_" + b].astype(real), cell, G, grid) for b in ("a", "b")]
    n = c["uu_a"].size
    wo = np.full(n, real(-7.0))
    sk.append(read(n, c["uu_a"].astype(real), c["vv_a"].astype(real), c["w_a"].astype(real),
                   wo, cell, G, grid))
    assert same_bits_or_nan(grid, c["grid_" + tag])
    assert same_bits_or_nan(wo, c["weight_out_" + tag])
    assert sk == [int(x) for x in c["skipped_" + tag]]


@pytest.mark.parametrize("name", golden_names("weights_"))
@pytest.mark.parametrize("real", [np.float32, np.float64])
def test_device_entry_points_match_reference_fixture(gpu, name, real):
    torch, ops = gpu
    c = _golden(name)
    tag = "f" if real == np.float32 else "d"
    G = int(c["grid_size"])
    blocks = [tuple(c[k + b].astype(real) for k in ("uu_", "vv_", "w_")) for b in ("a", "b")]
    grid, wo, sk = _dev_run(torch, ops, blocks, blocks[0], real, real(c["cell_size_rad"]), G)
    assert same_bits_or_nan(grid, c["grid_" + tag])
    assert same_bits_or_nan(wo, c["weight_out_" + tag])
    assert sk == [int(x) for x in c["skipped_" + tag]]


def _random_blocks(rng, real, N, G, spread=1.15, nblocks=2):
    out = []
    for _ in range(nblocks):
        u = (rng.normal(0, 0.35, N) * spread * G).astype(real)
        v = (rng.normal(0, 0.35, N) * spread * G).astype(real)
        w = rng.lognormal(0, 1.0, N).astype(real)
        out.append((u, v, w))
    return out


@pytest.mark.parametrize("real", [np.float32, np.float64])
@pytest.mark.parametrize("N,G", [(1, 1), (5000, 63), (300_000, 1024), (1_000_000, 4096)])
def test_random_blocks_bitexact_vs_oracle(gpu, real, N, G):
    torch, ops = gpu
    rng = np.random.default_rng(N + G)
    cell = real(1.0 / G)
    blocks = _random_blocks(rng, real, N, G)
    grid0 = rng.uniform(0, 2, (G, G)).astype(real)          # accumulate onto existing values
    want = _oracle_run(blocks, blocks[1], real, cell, G, grid0)
    got = _dev_run(torch, ops, blocks, blocks[1], real, cell, G, grid0)
    assert got[2] == want[2]
    assert same_bits_or_nan(got[0], want[0])
    assert same_bits_or_nan(got[1], want[1])


def test_long_cell_runs_and_host_window(gpu):
    """200k points in a handful of cells (long ordered sums), through the host
    drop-in whose window covers only those cells."""
    rng = np.random.default_rng(3)
    real, G, N = np.float32, 2048, 200_000
    cell = real(1.0 / G)
    u = rng.choice(np.array([3.2, -7.9, 100.4], real), N)
    v = rng.choice(np.array([-0.3, 55.1], real), N)
    w = rng.uniform(1e-3, 1e3, N).astype(real)
    g_ref = rng.uniform(0, 1, (G, G)).astype(real)
    g = g_ref.copy()
    assert oracle.weights_write(u, v, w, cell, G, g_ref) == \
        wproj.grid_weights_write_f(N, u, v, w, cell, G, g) == 0
    assert same_bits_or_nan(g, g_ref)
    wo_ref = np.full(N, real(5.0))
    wo = wo_ref.copy()
    oracle.weights_read(u, v, w, wo_ref, cell, G, g_ref)
    wproj.grid_weights_read_f(N, u, v, w, wo, cell, G, g)
    assert same_bits_or_nan(wo, wo_ref)


def test_zero_points_and_errors_on_device(gpu):
    torch, ops = gpu
    G = 16
    grid = torch.full((G, G), 3.0, dtype=torch.float32, device="cuda")
    sk = torch.full((1,), 7, dtype=torch.int64, device="cuda")
    e = torch.zeros(0, dtype=torch.float32, device="cuda")
    ops.grid_weights_write(e, e, e, 0.1, G, grid, sk)
    assert int(sk.item()) == 0 and bool((grid == 3.0).all())
    sk.fill_(7)
    ops.grid_weights_read(e, e, e, e, 0.1, G, grid, sk)
    assert int(sk.item()) == 0
    with pytest.raises(_lib.GridderError):
        big = torch.zeros(1, dtype=torch.float32, device="cuda")
        lib = _lib.load()
        import ctypes
        _lib.check(lib.mi355_grid_weights_write_f_dev(
            ops.context().handle, 1, ctypes.c_void_p(big.data_ptr()),
            ctypes.c_void_p(big.data_ptr()), ctypes.c_void_p(big.data_ptr()),
            ctypes.c_float(0.1), 70000, ctypes.c_void_p(sk.data_ptr()),
            ctypes.c_void_p(big.data_ptr())))


@pytest.mark.slow
def test_full_size_write_read_bitexact(gpu):
    """The 31.4M-point SKA-Low workload on its 16384^2 grid, FP32."""
    torch, ops = gpu
    c = synth.make(2)
    G = c["grid_size"]
    cell = np.float32(c["cell_size_rad"])
    blk = (c["uu"], c["vv"], c["weight"])
    want = _oracle_run([blk], blk, np.float32, cell, G)
    got = _dev_run(torch, ops, [blk], blk, np.float32, cell, G)
    assert got[2] == want[2]
    assert same_bits_or_nan(got[0], want[0])
    assert same_bits_or_nan(got[1], want[1])
