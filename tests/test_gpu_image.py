"""Grid -> image on the GPU (hipFFT plus the HIP centring/scaling passes)
against the oracle's FP64 direct summation, through the C ABI.

Bars: FP32 norm-wise max|I - I64| / max|I64| <= 1e-5; FP64 <= 1e-12.  At
the 16384^2 size: point sources at known cells against their closed-form
cosine images at sampled pixels, same FP32 bar.
"""
import numpy as np
import pytest

from oracle import oracle
from mi355_gridding import wproj, _lib

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / np.abs(b).max())


@pytest.mark.parametrize("G", [2, 64, 96, 256, 500])
@pytest.mark.parametrize("prec", ["f", "d"])
def test_device_matches_oracle(gpu, G, prec):
    torch, ops = gpu
    rng = np.random.default_rng(G)
    g = rng.normal(size=(G, G)) + 1j * rng.normal(size=(G, G))
    if prec == "f":
        g = g.astype(np.complex64)      # the oracle then sees the grid the GPU sees
    want = oracle.grid_to_image(g, norm=3.25)
    rdt = torch.float32 if prec == "f" else torch.float64
    img = torch.empty((G, G), dtype=rdt, device="cuda")
    ops.grid_to_image(torch.from_numpy(g).cuda(), G, img, norm=3.25)
    torch.cuda.synchronize()
    assert _rel(img.cpu().numpy(), want) <= (1e-5 if prec == "f" else 1e-12)


@pytest.mark.parametrize("prec", ["f", "d"])
def test_host_drop_in_matches_device(gpu, prec):
    torch, ops = gpu
    G = 128
    rng = np.random.default_rng(7)
    g = (rng.normal(size=(G, G)) + 1j * rng.normal(size=(G, G))).astype(
        np.complex64 if prec == "f" else np.complex128)
    host = wproj.grid_to_image(g, norm=2.0)
    img = torch.empty((G, G), dtype=torch.float32 if prec == "f" else torch.float64,
                      device="cuda")
    ops.grid_to_image(torch.from_numpy(g).cuda(), G, img, norm=2.0)
    torch.cuda.synchronize()
    assert np.array_equal(host, img.cpu().numpy())


def test_odd_size_refused(gpu):
    torch, ops = gpu
    g = torch.zeros((5, 5), dtype=torch.complex64, device="cuda")
    img = torch.empty((5, 5), dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.GridderError) as e:
        ops.grid_to_image(g, 5, img)
    assert e.value.status == _lib.ERR_INVALID_ARG


@pytest.mark.slow
def test_full_size_point_sources(gpu):
    """16384^2, FP32: four point sources; 4096 sampled pixels against the
    closed form sum of a_k cos(2 pi (du_k (x-c) + dv_k (y-c)) / G) (real
    amplitudes), divided by the norm."""
    torch, ops = gpu
    G = 16384
    c = G // 2
    src = [(0, 0, 1.0), (3, -2, 0.5), (-1000, 17, -0.25), (4095, 8000, 2.0)]
    g = torch.zeros((G, G), dtype=torch.complex64, device="cuda")
    for du, dv, a in src:
        g[c + dv, c + du] = a
    img = torch.empty((G, G), dtype=torch.float32, device="cuda")
    ops.grid_to_image(g, G, img, norm=4.0)
    torch.cuda.synchronize()
    rng = np.random.default_rng(3)
    ys = rng.integers(0, G, 4096)
    xs = rng.integers(0, G, 4096)
    got = img[torch.from_numpy(ys).cuda(), torch.from_numpy(xs).cuda()].cpu().numpy()
    want = np.zeros(4096)
    for du, dv, a in src:
        k = (du * (xs - c) + dv * (ys - c)) % G        # exact phase index
        want += a * np.cos(2 * np.pi * k / G)
    want /= 4.0
    assert np.abs(got - want).max() <= 1e-5 * 3.75 / 4.0
