"""Pins the oracle's grid -> image restatement (oracle_grid_to_image_d: the
centred inverse 2-D DFT, real part, scaled by 1/norm; README:24-26 names the
step, the reference has no code for it) before the GPU transform is checked
against it:

1. closed forms: one value at the centre cell gives a constant image; a unit
   value at cell (c + du, c + dv) gives cos(2 pi (du (x-c) + dv (y-c)) / G);
2. an independent statement through numpy's FFT:
   image = (-1)^(x+y) * G^2 * ifft2((-1)^(u+v) grid), real part / norm,
   on random grids of power-of-two and other even sizes;
3. odd sizes are refused.
"""
import numpy as np
import pytest

from oracle import oracle


def numpy_image(g, norm):
    G = g.shape[0]
    s = (-1.0) ** np.add.outer(np.arange(G), np.arange(G))
    scale = 1.0 / norm if norm else 1.0
    return (s * np.fft.ifft2(s * g) * G * G).real * scale


def test_centre_value_gives_constant_image():
    G = 32
    g = np.zeros((G, G), np.complex128)
    g[G // 2, G // 2] = 3.0 - 2.0j
    im = oracle.grid_to_image(g, norm=1.5)
    assert np.allclose(im, 2.0, rtol=0, atol=1e-14)


@pytest.mark.parametrize("du,dv", [(1, 0), (0, 3), (-5, 2), (7, -7)])
def test_offset_value_gives_cosine(du, dv):
    G = 48
    c = G // 2
    g = np.zeros((G, G), np.complex128)
    g[c + dv, c + du] = 1.0
    im = oracle.grid_to_image(g)
    x = np.arange(G) - c
    want = np.cos(2 * np.pi * (du * x[None, :] + dv * x[:, None]) / G)
    assert np.abs(im - want).max() < 1e-13


@pytest.mark.parametrize("G", [2, 16, 50, 64, 96])
def test_matches_numpy_fft(G):
    rng = np.random.default_rng(G)
    g = rng.normal(size=(G, G)) + 1j * rng.normal(size=(G, G))
    im = oracle.grid_to_image(g, norm=0.7)
    ref = numpy_image(g, 0.7)
    assert np.abs(im - ref).max() <= 1e-12 * np.abs(ref).max()


def test_real_layout_and_zero_norm():
    G = 16
    rng = np.random.default_rng(1)
    g = rng.normal(size=(G, G, 2))
    a = oracle.grid_to_image(g)                      # [G, G, 2] real layout, norm 0 -> 1
    b = numpy_image(g[..., 0] + 1j * g[..., 1], 0.0)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_odd_size_refused():
    with pytest.raises(ValueError):
        oracle.grid_to_image(np.zeros((5, 5), np.complex128))
