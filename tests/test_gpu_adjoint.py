"""Degridding is the exact adjoint of gridding; autograd through it.

Bars: dot-product test <A x, y> = <x, A^H y> to 1e-12 (FP64 path) and
1e-5 (FP32 path); degridding vs the FP64 oracle adjoint to 1e-5 (FP32);
gradients vs FP64 central finite differences on 32 random entries to 1e-7
(the loss is quadratic, so central differences are exact up to rounding).
"""
import numpy as np
import pytest

from conftest import errors
from oracle import oracle
from mi355_gridding import synth

pytestmark = pytest.mark.gpu


def to_dev(torch, c, dtype):
    return {k: torch.from_numpy(np.ascontiguousarray(c[k])).to(dtype).cuda()
            for k in ("uu", "vv", "ww", "vis", "weight")}


# (support 100: footprints up to 201 wide, several grid regions in each
# direction, tables read only directly by the gridder)
@pytest.mark.parametrize("smax,N,P", [(24, 60_000, 64), (100, 6_000, 16)])
@pytest.mark.parametrize("prec,tol", [("f", 1e-5), ("d", 1e-12)])
def test_dot_product_adjoint(gpu, prec, tol, smax, N, P):
    torch, ops = gpu
    dtype = torch.float32 if prec == "f" else torch.float64
    c = synth.make(4, num_points=N, grid_size=1024, num_w_planes=P, max_support=smax)
    ks = ops.KernelStack(c["support"], c["oversample"], c["conv_size_half"], c["conv_func"],
                         dtype=dtype)
    d = to_dev(torch, c, dtype)
    G = c["grid_size"]
    gen = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn((c["num_points"], 2), dtype=dtype, device="cuda", generator=gen)
    y = torch.randn((G, G, 2), dtype=dtype, device="cuda", generator=gen)
    ax = torch.zeros((G, G, 2), dtype=dtype, device="cuda")
    norm = torch.zeros(1, dtype=torch.float64, device="cuda")
    sk = torch.zeros(1, dtype=torch.int64, device="cuda")
    ops.grid(ks, d["uu"], d["vv"], d["ww"], x, d["weight"], c["cell_size_rad"], c["w_scale"],
             G, ax, norm, sk)
    aty = ops.degrid(ks, d["uu"], d["vv"], d["ww"], d["weight"], c["cell_size_rad"],
                     c["w_scale"], G, y)
    lhs = torch.sum(ax.double() * y.double()).item()
    rhs = torch.sum(x.double() * aty.double()).item()
    scale = (torch.linalg.norm(ax.double()) * torch.linalg.norm(y.double())).item()
    assert abs(lhs - rhs) <= tol * scale, (lhs, rhs)


def test_degrid_against_oracle_adjoint(gpu):
    torch, ops = gpu
    c = synth.make(4, num_points=8_000, grid_size=512, num_w_planes=32, max_support=16)
    rng = np.random.default_rng(4)
    G = c["grid_size"]
    y = rng.normal(0, 1, (G, G, 2)).astype(np.float32)
    ref = oracle.degrid_fsel_d(c, y.astype(np.float64))
    ks = ops.KernelStack(c["support"], c["oversample"], c["conv_size_half"], c["conv_func"])
    d = to_dev(torch, c, torch.float32)
    got = ops.degrid(ks, d["uu"], d["vv"], d["ww"], d["weight"], c["cell_size_rad"],
                     c["w_scale"], G, torch.from_numpy(y).cuda()).cpu().numpy()
    assert errors(got, ref)[0] <= 1e-5
    raw = ops.degrid(ks, d["uu"], d["vv"], d["ww"], d["weight"], c["cell_size_rad"],
                     c["w_scale"], G, torch.from_numpy(y).cuda(), apply_weight=False).cpu().numpy()
    refr = oracle.degrid_fsel_d(c, y.astype(np.float64), apply_weight=False)
    assert errors(raw, refr)[0] <= 1e-5


def test_autograd_finite_differences_fp64(gpu):
    """The differentiable workload (1M visibilities, 4096^2, 128 planes, support <= 32) in FP64:
    loss = 0.5 ||grid - target||^2 + 0.5 norm^2."""
    torch, ops = gpu
    dt = torch.float64
    c = synth.make(4)
    ks = ops.KernelStack(c["support"], c["oversample"], c["conv_size_half"], c["conv_func"],
                         dtype=dt)
    d = to_dev(torch, c, dt)
    G = c["grid_size"]
    gen = torch.Generator(device="cuda").manual_seed(11)
    target = torch.randn((G, G, 2), dtype=dt, device="cuda", generator=gen)
    args = (ks, d["uu"], d["vv"], d["ww"], c["cell_size_rad"], c["w_scale"], G)

    def loss(vis, w):
        g, n = ops.grid_autograd(vis, w, *args)
        return 0.5 * torch.sum((g - target) ** 2) + 0.5 * n[0] ** 2 * 1e-8

    vis = d["vis"].clone().requires_grad_(True)
    w = d["weight"].clone().requires_grad_(True)
    L = loss(vis, w)
    gv, gw = torch.autograd.grad(L, (vis, w))
    rng = np.random.default_rng(5)
    idx = rng.choice(c["num_points"], 16, replace=False)
    h = 0.25
    with torch.no_grad():
        for i in idx:
            for comp in range(2):
                vp = d["vis"].clone(); vp[i, comp] += h
                vm = d["vis"].clone(); vm[i, comp] -= h
                fd = (loss(vp, d["weight"]) - loss(vm, d["weight"])).item() / (2 * h)
                an = gv[i, comp].item()
                assert abs(fd - an) <= 1e-7 * max(1.0, abs(an)), (i, comp, fd, an)
            wp = d["weight"].clone(); wp[i] += h
            wm = d["weight"].clone(); wm[i] -= h
            fd = (loss(d["vis"], wp) - loss(d["vis"], wm)).item() / (2 * h)
            an = gw[i].item()
            assert abs(fd - an) <= 1e-7 * max(1.0, abs(an)), (i, fd, an)


def test_autograd_fp32_matches_fp64(gpu):
    torch, ops = gpu
    c = synth.make(4, num_points=50_000, grid_size=1024, num_w_planes=64, max_support=24)
    G = c["grid_size"]
    out = {}
    for dt in (torch.float32, torch.float64):
        ks = ops.KernelStack(c["support"], c["oversample"], c["conv_size_half"],
                             c["conv_func"], dtype=dt)
        d = to_dev(torch, c, dt)
        gen = torch.Generator(device="cuda").manual_seed(2)
        target = torch.randn((G, G, 2), dtype=torch.float64, device="cuda", generator=gen).to(dt)
        vis = d["vis"].clone().requires_grad_(True)
        w = d["weight"].clone().requires_grad_(True)
        g, n = ops.grid_autograd(vis, w, ks, d["uu"], d["vv"], d["ww"], c["cell_size_rad"],
                                 c["w_scale"], G)
        L = 0.5 * torch.sum((g - target) ** 2)
        L.backward()
        out[dt] = (vis.grad.double().cpu().numpy(), w.grad.double().cpu().numpy())
    # End to end the FP32 gradient carries the FP32 forward grid's rounding
    # (~1e-6 of max|grid|) through the residual into the adjoint, where the
    # residual's cancellation amplifies it; the bar is norm-wise L2 1e-5
    # (max-norm 1e-4).  The degridding kernel itself is held to 1e-5 max-norm
    # against the FP64 oracle adjoint above.
    for k in (0, 1):
        emax, el2 = errors(out[torch.float32][k], out[torch.float64][k])
        assert el2 <= 1e-5 and emax <= 1e-4, (k, emax, el2)
