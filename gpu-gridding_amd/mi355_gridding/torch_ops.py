"""Device-resident gridding on torch tensors, and its autograd wrapper.

Torch provides device memory and the stream; every computation is the
library's HIP path (libmi355_gridder.so), called with ``tensor.data_ptr()``
on the current torch stream.  Tensors must be contiguous and on the device
the context was created for.

``GridWProj`` is a ``torch.autograd.Function`` whose forward is the gridder
and whose backward is the library's hand-written degridding kernel (the
exact adjoint), for the complex amplitudes and the weights; coordinates get
no gradient (cell and kernel selection are piecewise constant in them).
"""
import ctypes

import numpy as np
import torch

from . import _lib


def _dp(t):
    return ctypes.c_void_p(t.data_ptr())


def _check_dev(t, dtype, numel, name):
    if not t.is_cuda:
        raise ValueError("%s must be a device tensor" % name)
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if t.dtype != dtype:
        raise ValueError("%s must be %s, got %s" % (name, dtype, t.dtype))
    if numel is not None and t.numel() != numel:
        raise ValueError("%s: expected %d elements, got %d" % (name, numel, t.numel()))


class KernelStack:
    """W-kernel stack (device) plus its support table (host), as passed to
    oskar_grid_wproj_f (support, oversample, conv_size_half, conv_func)."""

    def __init__(self, support, oversample, conv_size_half, conv_func, device="cuda",
                 dtype=torch.float32):
        self.support = np.ascontiguousarray(support, dtype=np.int32)
        self.num_w_planes = int(self.support.size)
        self.oversample = int(oversample)
        self.conv_size_half = int(conv_size_half)
        conv = torch.as_tensor(np.ascontiguousarray(conv_func)).to(device=device, dtype=dtype)
        self.conv = conv.contiguous().view(-1)
        if self.conv.numel() != self.num_w_planes * self.conv_size_half ** 2 * 2:
            raise ValueError("conv_func size does not match support/conv_size_half")
        self.dtype = dtype

    def support_ptr(self):
        return ctypes.c_void_p(self.support.ctypes.data)


_ctx_cache = {}


def context(device=None):
    """Per-device Context bound to the current torch stream."""
    dev = torch.cuda.current_device() if device is None else int(device)
    ctx = _ctx_cache.get(dev)
    if ctx is None:
        ctx = _lib.Context(dev)
        _ctx_cache[dev] = ctx
    ctx.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    return ctx


def grid(ks, uu, vv, ww, vis, weight, cell_size_rad, w_scale, grid_size, grid_out,
         norm, num_skipped, ctx=None):
    """grid_out (complex [G,G] or real [G,G,2]) += gridded visibilities.

    norm: float64 tensor [1] (accumulated); num_skipped: int64 tensor [1] (set).
    Precision follows ks.dtype (float32 -> mi355_grid_wproj_f_dev, float64 ->
    mi355_grid_wproj_d_dev).
    """
    ctx = ctx or context()
    n = uu.numel()
    real = ks.dtype
    for t, name, k in ((uu, "uu", n), (vv, "vv", n), (ww, "ww", n), (weight, "weight", n),
                       (vis, "vis", 2 * n)):
        _check_dev(t, real, k, name)
    G = int(grid_size)
    g = torch.view_as_real(grid_out) if grid_out.is_complex() else grid_out
    _check_dev(g, real, 2 * G * G, "grid")
    _check_dev(norm, torch.float64, 1, "norm")
    _check_dev(num_skipped, torch.int64, 1, "num_skipped")
    lib = _lib.load()
    if real == torch.float32:
        fn, scal = lib.mi355_grid_wproj_f_dev, ctypes.c_float
    else:
        fn, scal = lib.mi355_grid_wproj_d_dev, ctypes.c_double
    _lib.check(fn(ctx.handle, ks.num_w_planes, ks.support_ptr(), ks.oversample,
                  ks.conv_size_half, _dp(ks.conv), n, _dp(uu), _dp(vv), _dp(ww), _dp(vis),
                  _dp(weight), scal(float(cell_size_rad)), scal(float(w_scale)), G,
                  _dp(num_skipped), _dp(norm), _dp(g)))
    return grid_out


def grid_fsel_d(ks, uu, vv, ww, vis, weight, cell_size_rad, w_scale, grid_size, grid_out,
                norm, num_skipped, ctx=None):
    """FP32 inputs and selection, FP64 accumulation (mi355_grid_wproj_fsel_d_dev):
    grid_out is complex128 [G,G] or float64 [G,G,2]; ks holds the FP32 stack."""
    ctx = ctx or context()
    n = uu.numel()
    if ks.dtype != torch.float32:
        raise ValueError("grid_fsel_d takes the FP32 kernel stack")
    for t, name, k in ((uu, "uu", n), (vv, "vv", n), (ww, "ww", n), (weight, "weight", n),
                       (vis, "vis", 2 * n)):
        _check_dev(t, torch.float32, k, name)
    G = int(grid_size)
    g = torch.view_as_real(grid_out) if grid_out.is_complex() else grid_out
    _check_dev(g, torch.float64, 2 * G * G, "grid")
    _check_dev(norm, torch.float64, 1, "norm")
    _check_dev(num_skipped, torch.int64, 1, "num_skipped")
    lib = _lib.load()
    _lib.check(lib.mi355_grid_wproj_fsel_d_dev(
        ctx.handle, ks.num_w_planes, ks.support_ptr(), ks.oversample, ks.conv_size_half,
        _dp(ks.conv), n, _dp(uu), _dp(vv), _dp(ww), _dp(vis), _dp(weight),
        ctypes.c_float(float(cell_size_rad)), ctypes.c_float(float(w_scale)), G,
        _dp(num_skipped), _dp(norm), _dp(g)))
    return grid_out


SELECT_FIELDS = ("grid_u", "grid_v", "off_u", "off_v", "plane", "support", "conj", "skip")


def select(ks_support, oversample, uu, vv, ww, cell_size_rad, w_scale, grid_size, ctx=None):
    """Per-visibility selection on the device: dict of int32 tensors [N]."""
    ctx = ctx or context()
    support = np.ascontiguousarray(ks_support, dtype=np.int32)
    n = uu.numel()
    real = uu.dtype
    for t, name in ((uu, "uu"), (vv, "vv"), (ww, "ww")):
        _check_dev(t, real, n, name)
    out = torch.empty((8, n), dtype=torch.int32, device=uu.device)
    lib = _lib.load()
    if real == torch.float32:
        fn, scal = lib.mi355_select_wproj_f_dev, ctypes.c_float
    else:
        fn, scal = lib.mi355_select_wproj_d_dev, ctypes.c_double
    _lib.check(fn(ctx.handle, support.size, ctypes.c_void_p(support.ctypes.data),
                  int(oversample), n, _dp(uu), _dp(vv), _dp(ww), scal(float(cell_size_rad)),
                  scal(float(w_scale)), int(grid_size), _dp(out)))
    return {k: out[f] for f, k in enumerate(SELECT_FIELDS)}


def degrid(ks, uu, vv, ww, weight, cell_size_rad, w_scale, grid_size, grid_in,
           apply_weight=True, out=None, ctx=None):
    """Adjoint of ``grid`` w.r.t. the amplitudes: [N, 2] real tensor with
    out[i] = w_i * sum conj(c'_i) * grid (w_i omitted if not apply_weight)."""
    ctx = ctx or context()
    n = uu.numel()
    real = ks.dtype
    for t, name in ((uu, "uu"), (vv, "vv"), (ww, "ww"), (weight, "weight")):
        _check_dev(t, real, n, name)
    G = int(grid_size)
    g = torch.view_as_real(grid_in) if grid_in.is_complex() else grid_in
    g = g.contiguous()
    _check_dev(g, real, 2 * G * G, "grid")
    if out is None:
        out = torch.empty((n, 2), dtype=real, device=uu.device)
    _check_dev(out, real, 2 * n, "out")
    lib = _lib.load()
    if real == torch.float32:
        fn, scal = lib.mi355_degrid_wproj_f_dev, ctypes.c_float
    else:
        fn, scal = lib.mi355_degrid_wproj_d_dev, ctypes.c_double
    _lib.check(fn(ctx.handle, ks.num_w_planes, ks.support_ptr(), ks.oversample,
                  ks.conv_size_half, _dp(ks.conv), n, _dp(uu), _dp(vv), _dp(ww), _dp(weight),
                  scal(float(cell_size_rad)), scal(float(w_scale)), G, _dp(g),
                  1 if apply_weight else 0, _dp(out)))
    return out


class GridWProj(torch.autograd.Function):
    """(vis [N,2], weight [N]) -> (grid [G,G,2], norm [1]).

    Forward grids onto a fresh zero grid.  Backward: with g = dL/dgrid and
    y_i = sum over the footprint of conj(c'_i) g (the degridding kernel),
    dL/dvis_i = w_i * y_i and dL/dw_i = Re(conj(v_i) y_i) + dL/dnorm * S_i,
    where S_i = sum Re(c_i) is what visibility i adds to norm per unit
    weight (obtained by degridding a grid of ones).
    """

    @staticmethod
    def forward(ctx, vis, weight, ks, uu, vv, ww, cell_size_rad, w_scale, grid_size):
        G = int(grid_size)
        out = torch.zeros((G, G, 2), dtype=ks.dtype, device=vis.device)
        norm = torch.zeros(1, dtype=torch.float64, device=vis.device)
        skipped = torch.zeros(1, dtype=torch.int64, device=vis.device)
        grid(ks, uu, vv, ww, vis.contiguous(), weight.contiguous(), cell_size_rad, w_scale, G,
             out, norm, skipped)
        ctx.save_for_backward(vis, weight, uu, vv, ww)
        ctx.ks = ks
        ctx.params = (cell_size_rad, w_scale, G)
        ctx.mark_non_differentiable(skipped)
        return out, norm

    @staticmethod
    def backward(ctx, grad_grid, grad_norm):
        vis, weight, uu, vv, ww = ctx.saved_tensors
        ks = ctx.ks
        cell, ws, G = ctx.params
        grad_vis = grad_w = None
        y = None
        if grad_grid is not None:
            y = degrid(ks, uu, vv, ww, weight, cell, ws, G, grad_grid.contiguous(),
                       apply_weight=False)
        if ctx.needs_input_grad[0]:
            grad_vis = (y * weight.unsqueeze(1)) if y is not None else torch.zeros_like(vis)
        if ctx.needs_input_grad[1]:
            grad_w = (vis * y).sum(dim=1) if y is not None else torch.zeros_like(weight)
            if grad_norm is not None and bool(torch.any(grad_norm != 0)):
                ones = torch.zeros((G, G, 2), dtype=ks.dtype, device=vis.device)
                ones[..., 0] = 1
                s = degrid(ks, uu, vv, ww, weight, cell, ws, G, ones, apply_weight=False)
                grad_w = grad_w + grad_norm.to(ks.dtype) * s[:, 0]
        return grad_vis, grad_w, None, None, None, None, None, None, None


def grid_autograd(vis, weight, ks, uu, vv, ww, cell_size_rad, w_scale, grid_size):
    return GridWProj.apply(vis, weight, ks, uu, vv, ww, cell_size_rad, w_scale, grid_size)


def _job_table(ks, jobs, uu, vv, ww, vis, weight):
    if ks.dtype != torch.float32:
        raise ValueError("batched gridding is FP32")
    arr = (_lib.JobF * len(jobs))()
    for j, d in enumerate(jobs):
        g = torch.view_as_real(d["grid"]) if d["grid"].is_complex() else d["grid"]
        _check_dev(g, torch.float32, 2 * int(d["grid_size"]) ** 2, "job grid")
        _check_dev(d["norm"], torch.float64, 1, "job norm")
        _check_dev(d["num_skipped"], torch.int64, 1, "job num_skipped")
        arr[j] = _lib.JobF(int(d["vis_offset"]), int(d["num_points"]),
                           float(d["cell_size_rad"]), float(d["w_scale"]), int(d["grid_size"]),
                           g.data_ptr(), d["norm"].data_ptr(), d["num_skipped"].data_ptr())
    n = uu.numel()
    for t, name, k in ((uu, "uu", n), (vv, "vv", n), (ww, "ww", n), (weight, "weight", n),
                       (vis, "vis", 2 * n)):
        _check_dev(t, torch.float32, k, name)
    return arr


def _batched_call(fn, ks, jobs, uu, vv, ww, vis, weight, ctx, *extra):
    ctx = ctx or context()
    arr = _job_table(ks, jobs, uu, vv, ww, vis, weight)
    _lib.check(fn(ctx.handle, ks.num_w_planes, ks.support_ptr(), ks.oversample,
                  ks.conv_size_half, _dp(ks.conv), len(jobs), arr, _dp(uu), _dp(vv), _dp(ww),
                  _dp(vis), _dp(weight), *extra))


def grid_batched(ks, jobs, uu, vv, ww, vis, weight, ctx=None):
    """One pipeline pass over many jobs (mi355_grid_wproj_f_batched_dev).

    jobs: list of dicts with vis_offset, num_points, cell_size_rad, w_scale,
    grid_size, grid (device tensor), norm (float64 [1]), num_skipped (int64 [1]).
    uu..weight hold all jobs' visibilities back to back.
    """
    _batched_call(_lib.load().mi355_grid_wproj_f_batched_dev, ks, jobs, uu, vv, ww, vis,
                  weight, ctx)


def grid_batched_graph(ks, jobs, uu, vv, ww, vis, weight, ctx=None):
    """The batched pass as one HIP graph (mi355_grid_wproj_f_batched_graph_dev):
    captured on the first call with these jobs, pointers and options, replayed
    on the context's stream afterwards."""
    _batched_call(_lib.load().mi355_grid_wproj_f_batched_graph_dev, ks, jobs, uu, vv, ww, vis,
                  weight, ctx)


def autotune_batched(ks, jobs, uu, vv, ww, vis, weight, ctx=None):
    """Pick the gridding kernel, work-item size, item grouping and the
    shortest staged run per support class for this workload
    (mi355_autotune_f_batched_dev); the context keeps them, and caches the
    choice per workload shape.  Returns the chosen settings' time per pass
    in ms."""
    best = ctypes.c_double(0.0)
    _batched_call(_lib.load().mi355_autotune_f_batched_dev, ks, jobs, uu, vv, ww, vis, weight,
                  ctx, ctypes.byref(best))
    return best.value


def _weights_fn(real, write):
    lib = _lib.load()
    if real == torch.float32:
        return (lib.mi355_grid_weights_write_f_dev if write else
                lib.mi355_grid_weights_read_f_dev), ctypes.c_float
    if real == torch.float64:
        return (lib.mi355_grid_weights_write_d_dev if write else
                lib.mi355_grid_weights_read_d_dev), ctypes.c_double
    raise ValueError("weights: float32 or float64 tensors")


def grid_weights_write(uu, vv, weight, cell_size_rad, grid_size, grid, num_skipped, ctx=None):
    """Uniform-weighting density: grid (real [G, G], precision of uu) += the
    weights per cell, in point order (mi355_grid_weights_write_*_dev);
    num_skipped (int64 tensor [1]) is set."""
    ctx = ctx or context()
    n = uu.numel()
    real = uu.dtype
    for t, name in ((uu, "uu"), (vv, "vv"), (weight, "weight")):
        _check_dev(t, real, n, name)
    G = int(grid_size)
    _check_dev(grid, real, G * G, "grid")
    _check_dev(num_skipped, torch.int64, 1, "num_skipped")
    fn, scal = _weights_fn(real, True)
    _lib.check(fn(ctx.handle, n, _dp(uu), _dp(vv), _dp(weight), scal(float(cell_size_rad)), G,
                  _dp(num_skipped), _dp(grid)))
    return grid


def grid_weights_read(uu, vv, weight_in, weight_out, cell_size_rad, grid_size, grid,
                      num_skipped, ctx=None):
    """weight_out = weight_in / (cell total) for points on the grid (0 where the
    total is 0; entries of points off the grid are left as they are)
    (mi355_grid_weights_read_*_dev); num_skipped (int64 tensor [1]) is set."""
    ctx = ctx or context()
    n = uu.numel()
    real = uu.dtype
    for t, name in ((uu, "uu"), (vv, "vv"), (weight_in, "weight_in"),
                    (weight_out, "weight_out")):
        _check_dev(t, real, n, name)
    G = int(grid_size)
    _check_dev(grid, real, G * G, "grid")
    _check_dev(num_skipped, torch.int64, 1, "num_skipped")
    fn, scal = _weights_fn(real, False)
    _lib.check(fn(ctx.handle, n, _dp(uu), _dp(vv), _dp(weight_in), _dp(weight_out),
                  scal(float(cell_size_rad)), G, _dp(num_skipped), _dp(grid)))
    return weight_out


def grid_to_image(grid, grid_size, image, norm=0.0, ctx=None):
    """image (real [G, G]) = centred inverse 2-D DFT of grid (complex [G, G]
    or real [G, G, 2]), real part, / norm (mi355_grid_to_image_*_dev; FP32 or
    FP64 by the grid's precision).  norm may be a float or a 1-element
    tensor (read on the host)."""
    ctx = ctx or context()
    G = int(grid_size)
    g = torch.view_as_real(grid) if grid.is_complex() else grid
    real = g.dtype
    _check_dev(g, real, 2 * G * G, "grid")
    _check_dev(image, real, G * G, "image")
    if torch.is_tensor(norm):
        norm = float(norm.item())
    lib = _lib.load()
    fn = lib.mi355_grid_to_image_f_dev if real == torch.float32 else lib.mi355_grid_to_image_d_dev
    _lib.check(fn(ctx.handle, G, _dp(g), float(norm), _dp(image)))
    return image
