"""C ABI checks that need no GPU: the library loads, exports every entry
point include/mi355_gridder.h declares, and rejects invalid arguments
before touching the device (the errors the reference's GPU path would abort
on, oskar_grid_wproj_gpu.cu:1833-1843, gpu_support.hpp:143-151).
Also the synthetic-workload generator's determinism and support schedule."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from mi355_gridding import _lib, synth, wproj

HEADER = os.path.join(ROOT, "include", "mi355_gridder.h")
HEADERS = [HEADER, os.path.join(ROOT, "include", "mi355_io.h")]


def declared_functions():
    text = "".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"MI355_EXPORT\s+[\w\s\*]+?\b(mi355_\w+)\s*\(", text)))


def test_header_declares_the_drop_ins():
    names = declared_functions()
    assert "mi355_grid_wproj_f" in names and "mi355_grid_wproj_d" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, "no ctypes prototype for " + name


def test_drop_in_signature_matches_reference_header():
    """Same parameter list as oskar_grid_wproj_f / _d (oskar_grid_wproj.h:68-131)."""
    text = open(HEADER).read()
    for prec, real in (("f", "float"), ("d", "double")):
        m = re.search(r"void mi355_grid_wproj_%s\((.*?)\);" % prec, text, re.S)
        params = [p.strip() for p in m.group(1).replace("\n", " ").split(",")]
        types = [re.sub(r"\s+", " ", p.rsplit(" ", 1)[0].replace("const ", "")).strip()
                 for p in params]
        assert types == ["size_t", "int*", "int", "int", real + "*", "size_t", real + "*",
                         real + "*", real + "*", real + "*", real + "*", real, real, "int",
                         "size_t*", "double*", real + "*"], types


def _tiny(P=2, smax=2, ov=4, N=4, G=32):
    support, conv, csh = synth.kernels(P, smax, ov)
    return dict(num_w_planes=P, support=support, oversample=ov, conv_size_half=csh,
                conv_func=conv, num_points=N, uu=np.zeros(N, np.float32),
                vv=np.zeros(N, np.float32), ww=np.zeros(N, np.float32),
                vis=np.ones((N, 2), np.float32), weight=np.ones(N, np.float32),
                cell_size_rad=np.float32(0.01), w_scale=np.float32(1.0), grid_size=G)


def _call(c, grid=None):
    G = c["grid_size"]
    grid = np.zeros((G, G, 2), np.float32) if grid is None else grid
    return wproj.grid_wproj_f(c["num_w_planes"], c["support"], c["oversample"],
                              c["conv_size_half"], c["conv_func"], c["num_points"], c["uu"],
                              c["vv"], c["ww"], c["vis"], c["weight"], c["cell_size_rad"],
                              c["w_scale"], G, 0.0, grid)


@pytest.mark.parametrize("field,value,status", [
    ("oversample", 0, _lib.ERR_INVALID_ARG),
    ("grid_size", 0, _lib.ERR_INVALID_ARG),
])
def test_invalid_scalars_rejected_before_device(field, value, status):
    c = _tiny()
    c[field] = value
    with pytest.raises(_lib.GridderError) as e:
        _call(c, grid=np.zeros((32, 32, 2), np.float32) if field == "grid_size" else None) \
            if field != "grid_size" else wproj.grid_wproj_f(
                c["num_w_planes"], c["support"], c["oversample"], c["conv_size_half"],
                c["conv_func"], c["num_points"], c["uu"], c["vv"], c["ww"], c["vis"],
                c["weight"], c["cell_size_rad"], c["w_scale"], 0, 0.0,
                np.zeros(0, np.float32))
    assert e.value.status == status


def test_kernel_too_small_rejected():
    c = _tiny()
    c["support"] = np.array([2, 4], np.int32)     # needs conv_size_half > 4*4+2
    with pytest.raises(_lib.GridderError) as e:
        _call(c)
    assert e.value.status == _lib.ERR_KERNEL_TOO_SMALL


def test_negative_support_rejected():
    c = _tiny()
    c["support"] = np.array([2, -1], np.int32)
    with pytest.raises(_lib.GridderError) as e:
        _call(c)
    assert e.value.status == _lib.ERR_INVALID_ARG


def test_zero_points_sets_num_skipped_without_device():
    """oskar_grid_wproj.c:151 zeroes num_skipped even for an empty block."""
    c = _tiny(N=0)
    g = np.full((32, 32, 2), 3.0, np.float32)
    skipped, norm = _call(c, g)
    assert skipped == 0 and norm == 0.0 and np.all(g == 3.0)


def test_null_pointers_rejected():
    lib = _lib.load()
    c = _tiny()
    sk = ctypes.c_size_t(7)
    nrm = ctypes.c_double(0)
    lib.mi355_grid_wproj_f(2, None, 4, c["conv_size_half"], None, 4, None, None, None, None,
                           None, ctypes.c_float(0.01), ctypes.c_float(1.0), 32,
                           ctypes.byref(sk), ctypes.byref(nrm), None)
    assert lib.mi355_grid_last_error() == _lib.ERR_INVALID_ARG
    assert b"NULL" in lib.mi355_grid_last_error_message()


def test_error_strings():
    lib = _lib.load()
    for code in range(7):
        assert lib.mi355_grid_error_string(code)


def test_synth_deterministic_and_schedule():
    a = synth.visibilities(123, 5000, 64, 0)
    b = synth.visibilities(123, 5000, 64, 0)
    for k in ("uu", "vv", "ww", "vis", "weight"):
        assert np.array_equal(a[k], b[k])
    c = synth.visibilities(124, 5000, 64, 0)
    assert not np.array_equal(a["uu"], c["uu"])
    support, conv, csh = synth.kernels(64, 16, 4)
    k = np.arange(64)
    assert np.array_equal(support, np.ceil(4 + 12 * np.sqrt(k / 63)).astype(np.int32))
    assert csh == 68 and conv.shape == (64, 68, 68, 2)
    assert abs(conv[0, 0, 0, 0] - 1.0) < 1e-6 and conv[0, 0, 0, 1] == 0.0


def test_synth_tracks_are_time_major():
    v = synth.visibilities(5, 30 * 7, 16, 1, baselines=30, times=7)
    r = np.hypot(v["uu"], v["vv"]).reshape(7, 30)
    assert np.allclose(r, r[0], rtol=1e-5)          # each baseline keeps its radius
    assert not np.allclose(v["uu"].reshape(7, 30), v["uu"][:30])   # and rotates


def test_weights_drop_in_signatures_match_reference_header():
    """Same parameter lists as oskar_grid_weights_write/read_f/_d
    (oskar_grid_weights.h:60-64, :84-88, :107-111, :131-135)."""
    text = open(HEADER).read()
    for prec, real in (("f", "float"), ("d", "double")):
        for op, want in (("write", ["size_t", real + "*", real + "*", real + "*", real, "int",
                                    "size_t*", real + "*"]),
                         ("read", ["size_t", real + "*", real + "*", real + "*", real + "*",
                                   real, "int", "size_t*", real + "*"])):
            m = re.search(r"void mi355_grid_weights_%s_%s\((.*?)\);" % (op, prec), text, re.S)
            params = [p.strip() for p in m.group(1).replace("\n", " ").split(",")]
            types = [re.sub(r"\s+", " ", p.rsplit(" ", 1)[0].replace("const ", "")).strip()
                     for p in params]
            assert types == want, (op, prec, types)


def test_weights_invalid_sizes_rejected_before_device():
    """grid_size outside [1, 65535] is refused before any array is read."""
    lib = _lib.load()
    u = np.zeros(3, np.float32)
    g = np.zeros(1, np.float32)
    p = ctypes.c_void_p(u.ctypes.data)
    sk = ctypes.c_size_t(9)
    for G, status in ((0, _lib.ERR_INVALID_ARG), (-5, _lib.ERR_INVALID_ARG),
                      (65536, _lib.ERR_TOO_LARGE)):
        lib.mi355_grid_weights_write_f(3, p, p, p, ctypes.c_float(0.01), G, ctypes.byref(sk),
                                       ctypes.c_void_p(g.ctypes.data))
        assert lib.mi355_grid_last_error() == status
        lib.mi355_grid_weights_read_f(3, p, p, p, p, ctypes.c_float(0.01), G, ctypes.byref(sk),
                                      ctypes.c_void_p(g.ctypes.data))
        assert lib.mi355_grid_last_error() == status


def test_weights_zero_points_and_nulls_without_device():
    """oskar_grid_weights.c:116 / :149 zero num_skipped for an empty block."""
    lib = _lib.load()
    g = np.full((8, 8), 3.0, np.float32)
    u = np.zeros(0, np.float32)
    assert wproj.grid_weights_write_f(0, u, u, u, 0.01, 8, g) == 0
    wo = np.zeros(0, np.float32)
    assert wproj.grid_weights_read_f(0, u, u, u, wo, 0.01, 8, g) == 0
    assert np.all(g == 3.0)
    sk = ctypes.c_size_t(5)
    lib.mi355_grid_weights_write_d(4, None, None, None, ctypes.c_double(0.1), 8,
                                   ctypes.byref(sk), None)
    assert lib.mi355_grid_last_error() == _lib.ERR_INVALID_ARG
    lib.mi355_grid_weights_read_f(4, None, None, None, None, ctypes.c_float(0.1), 8,
                                  ctypes.byref(sk), ctypes.c_void_p(g.ctypes.data))
    assert lib.mi355_grid_last_error() == _lib.ERR_INVALID_ARG


def _jobs(grids, norms, skips, G=64, n=10):
    jobs = (_lib.JobF * len(grids))()
    for j, (g, nr, sk) in enumerate(zip(grids, norms, skips)):
        jobs[j].vis_offset = j * n
        jobs[j].num_points = n
        jobs[j].cell_size_rad = 0.01
        jobs[j].w_scale = 1.0
        jobs[j].grid_size = G
        jobs[j].grid = g
        jobs[j].norm = nr
        jobs[j].num_skipped = sk
    return jobs


@pytest.mark.parametrize("case", ["same_grid", "overlapping_grid", "same_norm", "same_skipped"])
def test_batched_jobs_with_shared_outputs_rejected(case):
    """Jobs' outputs are written concurrently, so a job table whose jobs share
    a grid cell, a norm or a count is refused before the device is touched
    (the addresses below are never dereferenced)."""
    lib = _lib.load()
    G = 64
    gbytes = G * G * 8
    grids = [0x10000000, 0x20000000, 0x30000000]
    norms = [0x40000000, 0x40000008, 0x40000010]
    skips = [0x50000000, 0x50000008, 0x50000010]
    if case == "same_grid":
        grids[2] = grids[0]
    elif case == "overlapping_grid":
        grids[1] = grids[0] + gbytes - 8          # last cell of job 0 = first of job 1
    elif case == "same_norm":
        norms[1] = norms[0]
    else:
        skips[2] = skips[1]
    support, conv, csh = synth.kernels(2, 2, 4)
    arr = np.zeros(64, np.float32)
    p = ctypes.c_void_p(arr.ctypes.data)
    st = lib.mi355_grid_wproj_f_batched_dev(ctypes.c_void_p(1), 2,
                                            ctypes.c_void_p(support.ctypes.data), 4, csh,
                                            ctypes.c_void_p(conv.ctypes.data), 3,
                                            _jobs(grids, norms, skips, G), p, p, p, p, p)
    assert st == _lib.ERR_INVALID_ARG
    assert b"share" in lib.mi355_grid_last_error_message()
