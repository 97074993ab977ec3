"""The allocation-free gridding entry point (mi355_grid_wproj_f_dev_noalloc)
through the C ABI against mi355_grid_wproj_f_dev on the same inputs: two
calls in a row, the second with fewer visibilities on a smaller grid (the
static scratch keeps the first call's contents), grid, norm and num_skipped
bit for bit; and a call too large for the static scratch refused with
MI355_ERR_TOO_LARGE, its outputs untouched."""
import ctypes

import numpy as np
import pytest

from mi355_gridding import _lib, synth

pytestmark = pytest.mark.gpu


def _call(torch, fn, ctx, c, ks_support, conv):
    G = c["grid_size"]
    d = {k: torch.from_numpy(np.ascontiguousarray(c[k])).cuda()
         for k in ("uu", "vv", "ww", "vis", "weight")}
    grid = torch.zeros((G, G, 2), dtype=torch.float32, device="cuda")
    norm = torch.full((1,), 0.25, dtype=torch.float64, device="cuda")
    sk = torch.full((1,), 7, dtype=torch.int64, device="cuda")
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    st = fn(ctx.handle, c["num_w_planes"], ctypes.c_void_p(ks_support.ctypes.data),
            c["oversample"], c["conv_size_half"], p(conv), c["num_points"], p(d["uu"]),
            p(d["vv"]), p(d["ww"]), p(d["vis"]), p(d["weight"]),
            ctypes.c_float(float(c["cell_size_rad"])), ctypes.c_float(float(c["w_scale"])), G,
            p(sk), p(norm), p(grid))
    torch.cuda.synchronize()
    return st, grid.cpu().numpy(), norm.item(), sk.item()


def test_noalloc_matches_device_entry_point(gpu):
    torch, ops = gpu
    lib = _lib.load()
    ctx = ops.context()
    cases = [synth.make(0, num_points=60_000, grid_size=1024, num_w_planes=48, max_support=14,
                        seed=21),
             synth.make(0, num_points=9_000, grid_size=512, num_w_planes=48, max_support=14,
                        seed=22)]
    for c in cases:                        # second call smaller than the first
        sup = np.ascontiguousarray(c["support"], np.int32)
        conv = torch.from_numpy(np.ascontiguousarray(c["conv_func"])).cuda()
        st0, g0, n0, s0 = _call(torch, lib.mi355_grid_wproj_f_dev, ctx, c, sup, conv)
        st1, g1, n1, s1 = _call(torch, lib.mi355_grid_wproj_f_dev_noalloc, ctx, c, sup, conv)
        assert st0 == 0 and st1 == 0, (st0, st1, lib.mi355_grid_last_error_message())
        assert g0.tobytes() == g1.tobytes() and n0 == n1 and s0 == s1
        assert np.abs(g1).max() > 0 and n1 != 0.25
        st = ctx.stats()                  # stats of the allocation-free call
        assert st["num_points"] == c["num_points"]


def test_noalloc_refuses_work_beyond_its_scratch(gpu):
    torch, ops = gpu
    lib = _lib.load()
    ctx = ops.context()
    c = synth.make(0, num_points=3_000_000, grid_size=4096, num_w_planes=64, max_support=16,
                   seed=23)
    sup = np.ascontiguousarray(c["support"], np.int32)
    conv = torch.from_numpy(np.ascontiguousarray(c["conv_func"])).cuda()
    st, g, n, s = _call(torch, lib.mi355_grid_wproj_f_dev_noalloc, ctx, c, sup, conv)
    assert st == _lib.ERR_TOO_LARGE
    assert b"static scratch" in lib.mi355_grid_last_error_message()
    assert np.all(g == 0) and n == 0.25 and s == 7       # outputs untouched
