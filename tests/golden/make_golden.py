"""Writes the golden fixtures in this directory from the reference's own
serial gridder.

The reference's gridding path has no tests and no vectors of its own, so the
vectors are made here: small seeded inputs (plus hand-built special cases of
the selection rules) are run through oskar_grid_wproj_f and
oskar_grid_wproj_d compiled unmodified from the reference sources
(oracle/_ref/libref.so, built by `make -C oracle ref` from
/root/reference/oskar_grid_wproj.c), and inputs plus outputs are saved as
.npz.  The weights_*.npz fixtures do the same for the uniform-weighting
pre-pass (oskar_grid_weights_write/read_f/_d, from
/root/reference/utils/oskar_grid_weights.c).  Run from the repository root:

    make -C oracle && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "gpu-gridding_amd"))

from oracle import oracle  # noqa: E402
from mi355_gridding import synth  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def kernel_case(P, smax, ov, G, cell, w_scale):
    support, conv, csh = synth.kernels(P, smax, ov)
    return dict(num_w_planes=P, support=support, oversample=ov, conv_size_half=csh,
                conv_func=conv, grid_size=G, cell_size_rad=np.float32(cell),
                w_scale=np.float32(w_scale))


def random_case(seed, N, P, smax, ov, G, spread=1.1):
    rng = np.random.default_rng(seed)
    cell = 1.0 / (G * 8.0)                 # grid_scale = G*cell = 1/8 exactly
    reach = (G // 2 - smax - 2) * 8.0      # |uu| that just fits
    c = kernel_case(P, smax, ov, G, cell, 0.0)
    uu = rng.uniform(-spread * reach, spread * reach, N).astype(np.float32)
    vv = rng.uniform(-spread * reach, spread * reach, N).astype(np.float32)
    ww = rng.normal(0.0, 40.0, N).astype(np.float32)
    c["w_scale"] = np.float32((P - 1) ** 2 / (2.5 * 40.0) ** 2)   # some planes clamp
    vis = rng.normal(0.0, np.sqrt(0.5), (N, 2)).astype(np.float32)
    weight = rng.uniform(0.5, 2.0, N).astype(np.float32)
    c.update(num_points=N, uu=uu, vv=vv, ww=ww, vis=vis, weight=weight)
    return c


def special_case():
    """Hand-built inputs on the selection rules' edges (grid_scale = 1, so
    pos_u = -uu and pos_v = vv exactly; support 3 on every plane)."""
    G, P, ov = 64, 4, 4
    c = kernel_case(P, 3, ov, G, 1.0 / G, 1.0)
    c["support"] = np.full(P, 3, np.int32)
    pts = []   # (uu, vv, ww)
    # exact half-integer positions: roundf goes away from zero, off = +-ov/2
    for n in (0, 1, 5, -1, -6):
        pts += [(-(n + 0.5), 0.25, 0.0), (n + 0.5, -(n + 0.5), 0.0)]
    # offsets on their own rounding boundaries: (round(p)-p)*4 = +-1.5, +-0.5
    for frac in (0.375, 0.625, 0.125, 0.875):
        pts += [(-(3 + frac), 2 + frac, 1.0), (3 + frac, -(2 + frac), -1.0)]
    # footprints touching the first and last cells (kept) and one beyond (skipped)
    pts += [(29.0, 0.0, 0.0), (30.0, 0.0, 0.0), (-28.0, 0.0, 0.0), (-29.0, 0.0, 0.0),
            (0.0, -29.0, 0.0), (0.0, -30.0, 0.0), (0.0, 28.0, 0.0), (0.0, 29.0, 0.0)]
    # w: zero, negative zero, signs, plane rounding, clamp, non-finite
    pts += [(2.0, 3.0, 0.0), (2.0, 3.0, -0.0), (2.0, 3.0, 1e-30), (2.0, 3.0, -1e-30),
            (1.0, -1.0, 2.25), (1.0, -1.0, 6.25), (1.0, -1.0, -6.25), (1.0, -1.0, 1e4),
            (1.0, -1.0, np.inf), (1.0, -1.0, -np.inf), (1.0, -1.0, np.nan),
            (1.0, -1.0, 3.4e38)]
    # coordinates outside any int: NaN, Inf, huge
    pts += [(np.nan, 0.0, 0.0), (0.0, np.nan, 0.0), (np.inf, 0.0, 0.0), (0.0, -np.inf, 0.0),
            (1e10, 0.0, 0.0), (-3e9, 0.0, 0.0)]
    # duplicates (same cell, same kernel)
    pts += [(4.2, -7.7, 3.0)] * 3
    a = np.array(pts, np.float64)
    N = a.shape[0]
    rng = np.random.default_rng(7)
    c.update(num_points=N, uu=a[:, 0].astype(np.float32), vv=a[:, 1].astype(np.float32),
             ww=a[:, 2].astype(np.float32),
             vis=rng.normal(0, 1, (N, 2)).astype(np.float32),
             weight=rng.uniform(0.5, 2.0, N).astype(np.float32))
    return c


def nonfinite_case():
    """Random case with a few non-finite amplitudes and weights: the cells
    they touch become NaN / +-Inf in the serial code."""
    c = random_case(4, 300, 8, 6, 4, 64)
    c["vis"][5, 0] = np.nan
    c["vis"][17, 1] = np.inf
    c["vis"][33, 0] = -np.inf
    c["weight"][41] = np.inf
    c["weight"][57] = np.nan
    return c


def synth_case(N=3000, G=256, P=32, smax=12):
    c = synth.make(0, num_points=N, grid_size=G, num_w_planes=P, max_support=smax,
                   seed=12345)
    c.pop("name")
    c.pop("max_support")
    return c


def with_outputs(c, accumulate=False):
    ref = oracle.ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libref.so not built (make -C oracle ref)")
    G = int(c["grid_size"])
    out = dict(c)
    if accumulate:
        rng = np.random.default_rng(99)
        g0 = rng.normal(0, 1, (G, G, 2)).astype(np.float32)
        n0 = 12.5
        out["grid0"] = g0
        out["norm0"] = np.float64(n0)
        gf, nf, sf = oracle.grid_f(c, use_ref=True, grid=g0.copy(), norm=n0)
        gf, nf, sf = oracle.grid_f(c, use_ref=True, grid=gf, norm=nf)   # second call
        gd, nd, sd = oracle.grid_d(c, use_ref=True, grid=g0.astype(np.float64), norm=n0)
        gd, nd, sd = oracle.grid_d(c, use_ref=True, grid=gd, norm=nd)
    else:
        gf, nf, sf = oracle.grid_f(c, use_ref=True)
        gd, nd, sd = oracle.grid_d(c, use_ref=True)
    out.update(grid_f=gf, norm_f=np.float64(nf), skipped_f=np.int64(sf),
               grid_d=gd, norm_d=np.float64(nd), skipped_d=np.int64(sd))
    return out


def save(name, d):
    path = os.path.join(OUT, name + ".npz")
    np.savez_compressed(path, **{k: np.asarray(v) for k, v in d.items()})
    print("wrote", path, os.path.getsize(path), "bytes")


def weights_random_case(seed, N, G):
    """Two blocks of points, a sixth of them off the grid, with cells whose
    weights cancel to exactly zero and a few non-finite weights."""
    rng = np.random.default_rng(seed)
    cell = 1.0 / (G * 4.0)                 # grid_scale = 1/4 exactly
    reach = (G / 2) * 4.0 * 1.2
    c = dict(grid_size=np.int64(G), cell_size_rad=np.float64(cell))
    for b in ("a", "b"):
        uu = rng.uniform(-reach, reach, N)
        vv = rng.uniform(-reach, reach, N)
        w = rng.uniform(0.1, 3.0, N)
        # exact cancellation: pairs of points in one cell with weights +x, -x
        uu[:8] = uu[8:16]
        vv[:8] = vv[8:16]
        w[:8] = -w[8:16]
        c["uu_" + b], c["vv_" + b], c["w_" + b] = uu, vv, w
    # non-finite weights on the grid, in cells that block a also uses
    for i, j, u, v, x in ((100, 300, 21.0, -9.0, np.inf), (200, 400, -33.0, 41.0, np.nan)):
        c["uu_b"][i], c["vv_b"][i], c["w_b"][i] = u, v, x
        c["uu_a"][j], c["vv_a"][j] = u + 0.5, v + 0.5
    return c


def weights_special_case(G):
    """Cells on the rounding and grid edges (grid_scale = 1: position = -uu,
    vv), non-finite and out-of-int coordinates, and signed zeros."""
    h = G // 2
    pts = []
    for n in (0, 1, 2, -1, -2, h - 1, -h):
        pts += [(-(n + 0.5), 0.0), (0.0, n + 0.5), (n - 0.5, -(n - 0.5))]
    pts += [(-(G - 1 - h), 0.0), (-(G - h), 0.0), (h, 0.0), (h + 1, 0.0),
            (0.0, G - 1 - h), (0.0, G - h), (0.0, -h), (0.0, -h - 1),
            (-(G - 1 - h) - 0.49, 0.2), (-(G - 1 - h) - 0.5, 0.2), (h + 0.5, 0.1), (h + 0.49, 0.1)]
    pts += [(np.nan, 0.0), (0.0, np.nan), (np.inf, 0.0), (0.0, -np.inf), (1e10, 0.0),
            (-3e9, 0.0), (0.0, 2.2e9), (-0.0, -0.0), (0.0, -0.0), (-0.4, 0.4)]
    a = np.array(pts, np.float64)
    rng = np.random.default_rng(G)
    w = rng.uniform(0.5, 2.0, a.shape[0])
    return dict(grid_size=np.int64(G), cell_size_rad=np.float64(1.0 / G),
                uu_a=a[:, 0], vv_a=a[:, 1], w_a=w,
                uu_b=a[::-1, 0].copy(), vv_b=a[::-1, 1].copy(), w_b=w[::-1] * 0.5)


def weights_with_outputs(c):
    """write(a), write(b) into one zeroed grid, then read(a) into a weight_out
    prefilled with -7 (entries of points off the grid stay -7)."""
    if oracle.ref() is None:
        raise SystemExit("oracle/_ref/libref.so not built (make -C oracle ref)")
    G = int(c["grid_size"])
    out = dict(c)
    for tag, real in (("f", np.float32), ("d", np.float64)):
        cell = real(c["cell_size_rad"])
        grid = np.zeros((G, G), real)
        sk_a = oracle.weights_write(c["uu_a"].astype(real), c["vv_a"].astype(real),
                                    c["w_a"].astype(real), cell, G, grid, use_ref=True)
        sk_b = oracle.weights_write(c["uu_b"].astype(real), c["vv_b"].astype(real),
                                    c["w_b"].astype(real), cell, G, grid, use_ref=True)
        wo = np.full(c["uu_a"].size, -7.0, real)
        sk_r = oracle.weights_read(c["uu_a"].astype(real), c["vv_a"].astype(real),
                                   c["w_a"].astype(real), wo, cell, G, grid, use_ref=True)
        out.update({"grid_" + tag: grid, "weight_out_" + tag: wo,
                    "skipped_" + tag: np.array([sk_a, sk_b, sk_r], np.int64)})
    return out


FIXTURES = {
    "random_g64": lambda: with_outputs(random_case(1, 400, 8, 6, 4, 64)),
    "random_g128_ov3": lambda: with_outputs(random_case(2, 600, 12, 5, 3, 128)),
    "special_g64": lambda: with_outputs(special_case()),
    "synth_g256": lambda: with_outputs(synth_case()),
    "accumulate_g64": lambda: with_outputs(random_case(3, 300, 8, 6, 4, 64), accumulate=True),
    "nonfinite_g64": lambda: with_outputs(nonfinite_case()),
    "weights_random_g96": lambda: weights_with_outputs(weights_random_case(21, 3000, 96)),
    "weights_special_g64": lambda: weights_with_outputs(weights_special_case(64)),
    "weights_special_g63": lambda: weights_with_outputs(weights_special_case(63)),
}


def main(names):
    """Writes the named fixtures (all of them when none are named)."""
    for name in names or FIXTURES:
        save(name, FIXTURES[name]())


if __name__ == "__main__":
    main(sys.argv[1:])
