"""The end-to-end driver (mi355_grid_vis_file and the mi355_grid_main
command) over a data set in the reference's file formats, against the
oracle's restatement of main.cpp's flow (main.cpp:202-381) on the same
files.

The files are written by the library's own writers, which tests/test_io.py
pins against the reference's OSKAR binary library and cfitsio.  The
restatement: uniform-weighting grid over every block
(oskar_grid_weights_write_f), then per block the weights read
(oskar_grid_weights_read_f) and the gridder, accumulated over the blocks.

Bars: num_skipped (last block, as main.cpp reports it) exact; FP32 grid
norm-wise max|G - G64| / max|G64| <= 1e-5 against the FP64 truth with the
same selection and weights; norm relative 1e-12; FP64 data: grid and norm
within 1e-12 of oskar_grid_wproj_d's semantics.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import errors, ROOT
from oracle import oracle
from mi355_gridding import _lib, synth, vis_io

pytestmark = pytest.mark.gpu

C = 299792458.0


def make_dataset(root, prec, S=24, T=40, tpb=10, G=512, P=32, smax=12, ov=4, seed=11):
    """Kernel cubes from the synthetic generator and a visibility file whose
    coordinates are in metres at freq = c (so wavelengths equal metres)."""
    support, conv, csh = synth.kernels(P, smax, ov)
    real = np.float64 if prec == _lib.BIN_DOUBLE else np.float32
    cell = 1.0 / (G * 1.5)
    rng = np.random.default_rng(seed)
    B = S * (S - 1) // 2
    umax = (G // 2 - smax - 2) * 1.5
    r = np.abs(rng.normal(0, 0.45 * umax, B))                # a few footprints leave the grid
    th0 = rng.uniform(0, 2 * np.pi, B)
    th = th0[None, :] + np.linspace(0, 0.5, T)[:, None]      # tracks rotate with time
    uu, vv = r * np.cos(th), r * np.sin(th)
    ww = rng.normal(0, 1, (T, B)) * 0.3 * r
    w_scale = (P - 1) ** 2 / np.percentile(np.abs(ww), 99.9)     # planes span 0 .. P-1
    vis = (rng.normal(size=(T, B)) + 1j * rng.normal(size=(T, B))) / np.sqrt(2)
    kre = np.ascontiguousarray(conv[..., 0])
    kim = np.ascontiguousarray(conv[..., 1])
    vis_io.write_kernel(root + "_KERNELS_REAL.fits", kre, support, ov, G, cell, w_scale)
    vis_io.write_kernel(root + "_KERNELS_IMAG.fits", kim, support, ov, G, cell, w_scale)
    vis_io.write_vis(root + ".vis", uu, vv, ww, vis, S, tpb, C, precision=prec)
    return dict(support=support, conv=conv, csh=csh, P=P, ov=ov, G=G, cell=cell,
                w_scale=w_scale, real=real)


def emulate_main(root, d):
    """main.cpp's two passes with the oracle, on what the files hold."""
    real = d["real"]
    G = d["G"]
    blocks = []
    with vis_io.BinaryFile(root + ".vis") as f:
        h = f.vis_header()
        for b in range(h.num_blocks):
            blocks.append(f.vis_block(h, b))
    cell = real(d["cell"])
    wgrid = np.zeros((G, G), real)
    for uu, vv, ww, vis in blocks:
        oracle.weights_write(uu, vv, np.ones(uu.size, real), cell, G, wgrid)
    if real == np.float32:
        grid, norm = np.zeros((G, G, 2), np.float64), 0.0
    else:
        grid, norm = np.zeros((G, G, 2), np.float64), 0.0
    weight = np.zeros(max(b[0].size for b in blocks), real)
    for uu, vv, ww, vis in blocks:
        n = uu.size
        w = weight[:n]
        oracle.weights_read(uu, vv, np.ones(n, real), w, cell, G, wgrid)
        case = dict(num_w_planes=d["P"], support=d["support"], oversample=d["ov"],
                    conv_size_half=d["csh"], conv_func=d["conv"].astype(real), num_points=n,
                    uu=uu, vv=vv, ww=ww, vis=vis.view(real), weight=w.copy(),
                    cell_size_rad=d["cell"], w_scale=d["w_scale"], grid_size=G)
        if real == np.float32:
            grid, norm, skipped = oracle.grid_fsel_d(case, grid=grid, norm=norm)
        else:
            grid, norm, skipped = oracle.grid_d(case, grid=grid, norm=norm)
    return grid, norm, skipped


@pytest.mark.parametrize("prec", [_lib.BIN_SINGLE, _lib.BIN_DOUBLE])
def test_driver_matches_main_flow(gpu, tmp_path, prec):
    root = str(tmp_path / "synthetic")
    d = make_dataset(root, prec)
    want, wnorm, wskip = emulate_main(root, d)
    grid, res = vis_io.grid_vis_file(root, check=True)
    assert res["num_blocks"] == 4 and res["grid_size"] == d["G"]
    assert res["num_skipped_last"] == wskip
    assert abs(res["norm"] - wnorm) <= 1e-12 * abs(wnorm)
    err = errors(grid.view(d["real"]).reshape(d["G"], d["G"], 2), want)[0]
    assert err <= (1e-5 if prec == _lib.BIN_SINGLE else 1e-12), err
    assert 0 <= res["check_max_rel"] <= (1e-5 if prec == _lib.BIN_SINGLE else 1e-12)


def test_driver_image_and_errors(gpu, tmp_path):
    root = str(tmp_path / "s")
    d = make_dataset(root, _lib.BIN_SINGLE, S=10, T=6, tpb=3, G=256, P=8, smax=6)
    img_path = str(tmp_path / "img.fits")
    grid, res = vis_io.grid_vis_file(root, image_path=img_path)
    assert os.path.getsize(img_path) % 2880 == 0
    # the FITS image is the image transform of the returned grid
    raw = np.frombuffer(open(img_path, "rb").read()[2880:2880 + 4 * 256 * 256], ">f4")
    want = oracle.grid_to_image(grid, norm=res["norm"])
    assert np.abs(raw.reshape(256, 256) - want).max() <= 1e-5 * np.abs(want).max()
    with pytest.raises(_lib.GridderError) as e:
        vis_io.grid_vis_file(str(tmp_path / "absent"))
    assert e.value.status == _lib.ERR_IO


def test_command_line_driver(gpu, tmp_path):
    root = str(tmp_path / "cli")
    make_dataset(root, _lib.BIN_SINGLE, S=12, T=8, tpb=4, G=256, P=8, smax=6)
    exe = os.path.join(ROOT, "gpu-gridding_amd", "lib", "mi355_grid_main")
    out = subprocess.run([exe, root, "--check"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "Checking grids" in out.stdout and "(ok)" in out.stdout
    bad = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert bad.returncode != 0 and "Usage" in bad.stderr
