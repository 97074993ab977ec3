"""Pins the CPU oracle (oracle/grid_oracle.c) before anything is checked
against it.

1. Golden fixtures written by the reference's own compiled serial gridder
   (tests/golden/make_golden.py): the restatement must reproduce grid, norm
   and num_skipped bit for bit, in FP32 and FP64, including accumulation
   across calls (oskar_grid_wproj.c:151, :205-210).
2. oracle/_ref/libref.so (the reference compiled from its sources), where it
   was built: bit-identical on fresh random inputs.
3. Closed forms that a plausible mistake breaks: an independent numpy
   selection (roundf written as sign*floor(|x|+0.5)), a delta kernel, a
   constant kernel, conjugation symmetry, linearity, the Hermitian mirror,
   and the norm as a sum of per-footprint kernel sums.
"""
import numpy as np
import pytest

from conftest import errors, golden_names, load_golden, same_bits, same_scalar
from oracle import oracle


# ---------------------------------------------------------------------------
# 1. golden fixtures
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_reference_fixture(name):
    c = load_golden(name)
    if "grid0" in c:   # two accumulating calls from a non-zero grid and norm
        g, n, s = oracle.grid_f(c, grid=c["grid0"].copy(), norm=float(c["norm0"]))
        g, n, s = oracle.grid_f(c, grid=g, norm=n)
        gd, nd, sd = oracle.grid_d(c, grid=c["grid0"].astype(np.float64), norm=float(c["norm0"]))
        gd, nd, sd = oracle.grid_d(c, grid=gd, norm=nd)
    else:
        g, n, s = oracle.grid_f(c)
        gd, nd, sd = oracle.grid_d(c)
    assert same_bits(g, c["grid_f"])
    assert same_scalar(n, c["norm_f"]) and s == int(c["skipped_f"])
    assert same_bits(gd, c["grid_d"])
    assert same_scalar(nd, c["norm_d"]) and sd == int(c["skipped_d"])


def test_fixtures_cover_selection_edges():
    c = load_golden("special_g64")
    sel = oracle.select_f(c)
    offs = set(sel["off_u"][sel["skip"] == 0]) | set(sel["off_v"][sel["skip"] == 0])
    assert {-2, 2} <= offs                       # half-integer positions
    assert sel["skip"].sum() >= 6 and (sel["skip"] == 0).sum() >= 20
    kept_u = sel["grid_u"][sel["skip"] == 0]
    assert kept_u.min() - 3 == 0 and kept_u.max() + 3 == 63   # footprints touch both edges
    assert int(c["skipped_f"]) == int(sel["skip"].sum())


# ---------------------------------------------------------------------------
# 2. compiled reference, fresh inputs
# ---------------------------------------------------------------------------
def _random_case(seed, N=500, G=128, P=16, smax=8, ov=4):
    from mi355_gridding import synth
    rng = np.random.default_rng(seed)
    support, conv, csh = synth.kernels(P, smax, ov)
    reach = (G // 2 - smax - 2) / 0.37
    return dict(num_w_planes=P, support=support, oversample=ov, conv_size_half=csh,
                conv_func=conv, grid_size=G, cell_size_rad=np.float32(0.37 / G),
                w_scale=np.float32((P - 1) ** 2 / 900.0), num_points=N,
                uu=rng.uniform(-1.05 * reach, 1.05 * reach, N).astype(np.float32),
                vv=rng.uniform(-1.05 * reach, 1.05 * reach, N).astype(np.float32),
                ww=rng.normal(0, 15, N).astype(np.float32),
                vis=rng.normal(0, 1, (N, 2)).astype(np.float32),
                weight=rng.uniform(0.1, 3, N).astype(np.float32))


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_oracle_bit_identical_to_compiled_reference(seed):
    if oracle.ref() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    c = _random_case(seed)
    for fn in (oracle.grid_f, oracle.grid_d):
        g1, n1, s1 = fn(c)
        g2, n2, s2 = fn(c, use_ref=True)
        assert same_bits(g1, g2) and n1 == n2 and s1 == s2


# ---------------------------------------------------------------------------
# 3. closed forms
# ---------------------------------------------------------------------------
def _roundf(x):
    x = np.asarray(x, np.float64)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def numpy_select(c):
    """Selection of oskar_grid_wproj.c:158-187 written independently."""
    f = np.float32
    G = c["grid_size"]
    gs = f(G) * f(c["cell_size_rad"])
    pos_u = (-c["uu"].astype(f)) * gs
    pos_v = c["vv"].astype(f) * gs
    ru, rv = _roundf(pos_u), _roundf(pos_v)
    off_u = _roundf(((ru.astype(f) - pos_u) * f(c["oversample"])).astype(np.float64)).astype(int)
    off_v = _roundf(((rv.astype(f) - pos_v) * f(c["oversample"])).astype(np.float64)).astype(int)
    wi = _roundf(np.sqrt(np.abs(c["ww"].astype(f) * f(c["w_scale"]))).astype(f)).astype(np.int64)
    plane = np.minimum(wi, c["num_w_planes"] - 1)
    s = c["support"][plane]
    gu = ru.astype(int) + G // 2
    gv = rv.astype(int) + G // 2
    skip = (gu + s >= G) | (gu - s < 0) | (gv + s >= G) | (gv - s < 0)
    return dict(grid_u=gu, grid_v=gv, off_u=off_u, off_v=off_v, plane=plane, support=s,
                skip=skip.astype(int), conj=(c["ww"] > 0).astype(int))


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_selection_matches_independent_numpy(seed):
    c = _random_case(seed, N=20000)
    # put a share of the positions exactly on half-integers and on the
    # +-1.5 offset rounding boundary
    gs = np.float32(c["grid_size"]) * c["cell_size_rad"]
    k = np.arange(0, 20000, 3)
    c["uu"][k] = (-(np.round(c["uu"][k] * gs) + 0.5) / gs).astype(np.float32)
    k = np.arange(1, 20000, 5)
    c["vv"][k] = ((np.round(c["vv"][k] * gs) + 0.375) / gs).astype(np.float32)
    ref = numpy_select(c)
    got = oracle.select_f(c)
    for key in ("grid_u", "grid_v", "plane", "support", "skip", "conj"):
        assert np.array_equal(got[key], ref[key]), key
    keep = ref["skip"] == 0
    for key in ("off_u", "off_v"):
        assert np.array_equal(got[key][keep], ref[key][keep]), key


def _constant_kernel_case(value, seed=5):
    c = _random_case(seed, N=300, G=96, P=6, smax=5)
    c["conv_func"] = np.zeros_like(c["conv_func"])
    c["conv_func"][..., 0] = value.real
    c["conv_func"][..., 1] = value.imag
    return c


def test_constant_kernel_closed_form():
    val = np.complex64(0.75 - 0.5j)
    c = _constant_kernel_case(val)
    sel = numpy_select(c)
    G = c["grid_size"]
    exp = np.zeros((G, G), np.complex128)
    norm = 0.0
    for i in np.nonzero(sel["skip"] == 0)[0]:
        s = sel["support"][i]
        w = float(c["weight"][i])
        v = complex(c["vis"][i, 0], c["vis"][i, 1])
        k = np.conj(val) if c["ww"][i] > 0 else val
        exp[sel["grid_v"][i] - s:sel["grid_v"][i] + s + 1,
            sel["grid_u"][i] - s:sel["grid_u"][i] + s + 1] += w * v * complex(k)
        norm += w * float(val.real) * (2 * int(s) + 1) ** 2
    g, n, s = oracle.grid_d(c)
    gc = g[..., 0] + 1j * g[..., 1]
    assert np.max(np.abs(gc - exp)) <= 1e-12 * np.max(np.abs(exp))
    assert abs(n - norm) <= 1e-12 * abs(norm)
    assert s == int(sel["skip"].sum())
    gf, nf, _ = oracle.grid_f(c)
    assert errors(gf, np.stack([exp.real, exp.imag], -1))[0] < 1e-5


def test_delta_kernel_closed_form():
    """Kernel 1 at tap (0, 0) of every plane, 0 elsewhere: only visibilities
    with zero offsets contribute, each exactly w*v at its own cell."""
    c = _random_case(6, N=4000, G=96, P=6, smax=5)
    c["conv_func"] = np.zeros_like(c["conv_func"])
    c["conv_func"][:, 0, 0, 0] = 1.0
    sel = numpy_select(c)
    G = c["grid_size"]
    exp = np.zeros((G, G), np.complex128)
    norm = 0.0
    for i in np.nonzero((sel["skip"] == 0) & (sel["off_u"] == 0) & (sel["off_v"] == 0))[0]:
        w = float(c["weight"][i])
        exp[sel["grid_v"][i], sel["grid_u"][i]] += w * complex(c["vis"][i, 0], c["vis"][i, 1])
        norm += w
    g, n, _ = oracle.grid_d(c)
    assert np.allclose(g[..., 0] + 1j * g[..., 1], exp, rtol=0, atol=1e-12)
    assert abs(n - norm) <= 1e-12 * max(1.0, norm)


def test_conjugation_symmetry():
    c = _random_case(7)
    c["ww"][c["ww"] == 0] = 1.0
    g1, n1, _ = oracle.grid_f(c)
    c2 = dict(c)
    c2["ww"] = -c["ww"]
    c2["vis"] = c["vis"] * np.array([1, -1], np.float32)
    g2, n2, _ = oracle.grid_f(c2)
    # exact in every value (untouched cells differ only as +0.0 / -0.0)
    assert np.array_equal(g2[..., 0], g1[..., 0]) and np.array_equal(g2[..., 1], -g1[..., 1])
    assert n1 == n2


def test_linearity_fp64():
    c = _random_case(8)
    rng = np.random.default_rng(0)
    v2 = rng.normal(0, 1, c["vis"].shape).astype(np.float32)
    a = 1.75
    ga, _, _ = oracle.grid_d(dict(c, vis=(a * c["vis"].astype(np.float64) + v2)))
    g1, _, _ = oracle.grid_d(c)
    g2, _, _ = oracle.grid_d(dict(c, vis=v2))
    assert errors(ga, a * g1 + g2)[0] < 1e-13


def test_hermitian_mirror():
    """(u, v, w, V) and (-u, -v, -w, conj V) grid to mirror-image conjugates
    about the centre cell (kernel lookups use |offset|)."""
    c = _random_case(9, G=128)
    c["ww"][c["ww"] == 0] = 1.0
    g1, _, _ = oracle.grid_d(c)
    c2 = dict(c, uu=-c["uu"], vv=-c["vv"], ww=-c["ww"], vis=c["vis"] * np.array([1, -1], np.float32))
    g2, _, _ = oracle.grid_d(c2)
    G = c["grid_size"]
    # cell x -> G - x about the centre G/2; row/column 0 have no mirror
    m1 = g1[1:, 1:]
    m2 = g2[1:, 1:][::-1, ::-1]
    assert np.allclose(m2[..., 0], m1[..., 0], rtol=0, atol=1e-12)
    assert np.allclose(m2[..., 1], -m1[..., 1], rtol=0, atol=1e-12)
    assert G % 2 == 0


def test_norm_is_sum_of_footprint_kernel_sums():
    c = _random_case(10, N=2000)
    sel = numpy_select(c)
    conv = c["conv_func"].astype(np.float64)
    ov = c["oversample"]
    total = 0.0
    for i in np.nonzero(sel["skip"] == 0)[0]:
        s = sel["support"][i]
        j = np.arange(-s, s + 1)
        rows = np.abs(sel["off_v"][i] + ov * j)
        cols = np.abs(sel["off_u"][i] + ov * j)
        total += float(c["weight"][i]) * conv[sel["plane"][i]][np.ix_(rows, cols)][..., 0].sum()
    _, n, _ = oracle.grid_f(c)
    assert abs(n - total) <= 1e-12 * abs(total)


def test_fsel_d_matches_d_when_selection_agrees():
    c = load_golden("random_g64")   # grid_scale is a power of two: selections agree
    assert all(np.array_equal(oracle.select_f(c)[k], oracle.select_d(c)[k])
               for k in ("grid_u", "grid_v", "off_u", "off_v", "plane"))
    g, n, s = oracle.grid_fsel_d(c)
    assert errors(g, c["grid_d"])[0] < 1e-14
    assert abs(n - float(c["norm_d"])) <= 1e-13 * abs(float(c["norm_d"]))


def test_degrid_oracle_is_adjoint():
    c = _random_case(14, N=700)
    rng = np.random.default_rng(1)
    G = c["grid_size"]
    y = rng.normal(0, 1, (G, G, 2))
    x = c["vis"].astype(np.float64)
    gx, _, _ = oracle.grid_fsel_d(c)
    aty = oracle.degrid_fsel_d(c, y)
    lhs = np.sum(gx * y)                    # Re <A x, y>
    rhs = np.sum(x * aty)                   # Re <x, A^H y>
    assert abs(lhs - rhs) <= 1e-12 * np.sqrt(np.sum(gx ** 2) * np.sum(y ** 2))


def test_tile_oracle_matches_full_grid():
    c = _random_case(15, N=800, G=128)
    g, _, _ = oracle.grid_fsel_d(c)
    tiles = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], np.int32)
    t = oracle.tiles_fsel_d(c, tiles, tile=64, threads=2)
    for k, (tx, ty) in enumerate(tiles):
        assert np.array_equal(t[k], g[ty * 64:(ty + 1) * 64, tx * 64:(tx + 1) * 64])


@pytest.mark.parametrize("name", golden_names())
def test_norm_oracle_equals_gridder_norm(name):
    """oracle_norm_f (norm without the grid) reproduces the compiled
    reference's FP32 norm and num_skipped bit for bit on every fixture, and
    oracle_grid_wproj_f's on fresh random cases (it memoises the same
    per-footprint sum)."""
    c = load_golden(name)
    n0 = float(c["norm0"]) if "grid0" in c else 0.0
    n, s = oracle.norm_f(c, norm=n0)
    if "grid0" in c:
        n, s = oracle.norm_f(c, norm=n)
    assert same_scalar(n, c["norm_f"]) and s == int(c["skipped_f"])


@pytest.mark.parametrize("seed", [21, 22])
def test_norm_oracle_random(seed):
    c = _random_case(seed, N=3000)
    _, n, s = oracle.grid_f(c)
    n2, s2 = oracle.norm_f(c)
    assert n2 == n and s2 == s
