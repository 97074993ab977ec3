"""Host-side mirror of the reference's gridding interface.

``grid_wproj_f`` and ``grid_wproj_d`` take the arguments of
``oskar_grid_wproj_f`` / ``oskar_grid_wproj_d`` (oskar_grid_wproj.h:68-131), in
the same order and with the same meaning, as numpy arrays, and call the
library's drop-in C entry points (mi355_grid_wproj_f / _d).  As in the
reference, ``grid`` is accumulated in place and ``norm`` is accumulated; the
C functions return num_skipped and the updated norm through pointers, which
these wrappers return as ``(num_skipped, norm)``.

``grid_weights_write_f/_d`` and ``grid_weights_read_f/_d`` do the same for
the uniform-weighting pre-pass (oskar_grid_weights.h:60-135): the weights
grid (float [G, G]) is accumulated in place, weight_out is written in place,
and num_skipped is returned.

Errors the reference would abort on (oskar_grid_wproj_gpu.cu:1833-1843,
gpu_support.hpp:143-151) raise :class:`GridderError` here.
"""
import ctypes

import numpy as np

from . import _lib


def _arr(x, dtype, n=None, name=""):
    a = np.ascontiguousarray(x, dtype=dtype)
    if n is not None and a.size != n:
        raise ValueError("%s: expected %d values, got %d" % (name, n, a.size))
    return a


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _grid_common(fn, real, num_w_planes, support, oversample, conv_size_half,
                 conv_func, num_points, uu, vv, ww, vis, weight, cell_size_rad,
                 w_scale, grid_size, norm, grid):
    P = int(num_w_planes)
    N = int(num_points)
    G = int(grid_size)
    support = _arr(support, np.int32, P, "support")
    conv = _arr(conv_func, real, None, "conv_func")
    if conv.size != P * int(conv_size_half) ** 2 * 2:
        raise ValueError("conv_func: expected num_w_planes*conv_size_half^2 complex values")
    uu = _arr(uu, real, N, "uu")
    vv = _arr(vv, real, N, "vv")
    ww = _arr(ww, real, N, "ww")
    vis = _arr(vis, real, 2 * N, "vis")
    weight = _arr(weight, real, N, "weight")
    if not isinstance(grid, np.ndarray) or not grid.flags.c_contiguous:
        raise ValueError("grid must be a C-contiguous numpy array (accumulated in place)")
    if grid.dtype == np.complex64 and real == np.float32 or \
            grid.dtype == np.complex128 and real == np.float64:
        if grid.size != G * G:
            raise ValueError("grid: expected grid_size^2 complex values")
    elif grid.dtype == real:
        if grid.size != 2 * G * G:
            raise ValueError("grid: expected grid_size^2 complex values")
    else:
        raise ValueError("grid dtype %s does not match the precision" % grid.dtype)
    skipped = ctypes.c_size_t(0)
    nrm = ctypes.c_double(float(norm))
    scal = ctypes.c_float if real == np.float32 else ctypes.c_double
    fn(P, _ptr(support), int(oversample), int(conv_size_half), _ptr(conv), N,
       _ptr(uu), _ptr(vv), _ptr(ww), _ptr(vis), _ptr(weight),
       scal(cell_size_rad), scal(w_scale), G, ctypes.byref(skipped),
       ctypes.byref(nrm), ctypes.c_void_p(grid.ctypes.data))
    _lib.check_last()
    return skipped.value, nrm.value


def grid_wproj_f(num_w_planes, support, oversample, conv_size_half, conv_func,
                 num_points, uu, vv, ww, vis, weight, cell_size_rad, w_scale,
                 grid_size, norm, grid):
    """FP32 gridding; mirrors oskar_grid_wproj_f (oskar_grid_wproj.h:113-131).

    Returns (num_skipped, norm); ``grid`` (complex64 [G, G] or float32
    [G, G, 2]) is accumulated in place.
    """
    lib = _lib.load()
    return _grid_common(lib.mi355_grid_wproj_f, np.float32, num_w_planes, support,
                        oversample, conv_size_half, conv_func, num_points, uu, vv,
                        ww, vis, weight, cell_size_rad, w_scale, grid_size, norm, grid)


def grid_wproj_d(num_w_planes, support, oversample, conv_size_half, conv_func,
                 num_points, uu, vv, ww, vis, weight, cell_size_rad, w_scale,
                 grid_size, norm, grid):
    """FP64 gridding; mirrors oskar_grid_wproj_d (oskar_grid_wproj.h:68-86)."""
    lib = _lib.load()
    return _grid_common(lib.mi355_grid_wproj_d, np.float64, num_w_planes, support,
                        oversample, conv_size_half, conv_func, num_points, uu, vv,
                        ww, vis, weight, cell_size_rad, w_scale, grid_size, norm, grid)


def set_default_option(key, value):
    """Set an option (_lib.OPT_*) on the context the drop-ins use."""
    lib = _lib.load()
    h = ctypes.c_void_p()
    _lib.check(lib.mi355_default_ctx(ctypes.byref(h)))
    _lib.check(lib.mi355_ctx_set_option(h, int(key), int(value)))


def _weights_arrays(real, num_points, uu, vv, grid_size, grid):
    N = int(num_points)
    G = int(grid_size)
    uu = _arr(uu, real, N, "uu")
    vv = _arr(vv, real, N, "vv")
    if not isinstance(grid, np.ndarray) or not grid.flags.c_contiguous or grid.dtype != real:
        raise ValueError("grid must be a C-contiguous %s numpy array" % np.dtype(real).name)
    if grid.size != G * G:
        raise ValueError("grid: expected grid_size^2 values")
    return N, G, uu, vv


def _weights_write(fn, real, num_points, uu, vv, weight, cell_size_rad, grid_size, grid):
    N, G, uu, vv = _weights_arrays(real, num_points, uu, vv, grid_size, grid)
    weight = _arr(weight, real, N, "weight")
    skipped = ctypes.c_size_t(0)
    scal = ctypes.c_float if real == np.float32 else ctypes.c_double
    fn(N, _ptr(uu), _ptr(vv), _ptr(weight), scal(cell_size_rad), G, ctypes.byref(skipped),
       _ptr(grid))
    _lib.check_last()
    return skipped.value


def _weights_read(fn, real, num_points, uu, vv, weight_in, weight_out, cell_size_rad,
                  grid_size, grid):
    N, G, uu, vv = _weights_arrays(real, num_points, uu, vv, grid_size, grid)
    weight_in = _arr(weight_in, real, N, "weight_in")
    if not isinstance(weight_out, np.ndarray) or not weight_out.flags.c_contiguous or \
            weight_out.dtype != real or weight_out.size != N:
        raise ValueError("weight_out must be a C-contiguous %s array of num_points values"
                         % np.dtype(real).name)
    skipped = ctypes.c_size_t(0)
    scal = ctypes.c_float if real == np.float32 else ctypes.c_double
    fn(N, _ptr(uu), _ptr(vv), _ptr(weight_in), _ptr(weight_out), scal(cell_size_rad), G,
       ctypes.byref(skipped), _ptr(grid))
    _lib.check_last()
    return skipped.value


def grid_weights_write_f(num_points, uu, vv, weight, cell_size_rad, grid_size, grid):
    """Mirrors oskar_grid_weights_write_f (oskar_grid_weights.h:107-111):
    grid (float32 [G, G]) += weights per cell; returns num_skipped."""
    return _weights_write(_lib.load().mi355_grid_weights_write_f, np.float32, num_points, uu,
                          vv, weight, cell_size_rad, grid_size, grid)


def grid_weights_write_d(num_points, uu, vv, weight, cell_size_rad, grid_size, grid):
    """Mirrors oskar_grid_weights_write_d (oskar_grid_weights.h:60-64)."""
    return _weights_write(_lib.load().mi355_grid_weights_write_d, np.float64, num_points, uu,
                          vv, weight, cell_size_rad, grid_size, grid)


def grid_weights_read_f(num_points, uu, vv, weight_in, weight_out, cell_size_rad, grid_size,
                        grid):
    """Mirrors oskar_grid_weights_read_f (oskar_grid_weights.h:131-135):
    weight_out[i] = weight_in[i] / grid[cell] (0 where that is 0) for points
    on the grid; returns num_skipped."""
    return _weights_read(_lib.load().mi355_grid_weights_read_f, np.float32, num_points, uu, vv,
                         weight_in, weight_out, cell_size_rad, grid_size, grid)


def grid_weights_read_d(num_points, uu, vv, weight_in, weight_out, cell_size_rad, grid_size,
                        grid):
    """Mirrors oskar_grid_weights_read_d (oskar_grid_weights.h:84-88)."""
    return _weights_read(_lib.load().mi355_grid_weights_read_d, np.float64, num_points, uu, vv,
                         weight_in, weight_out, cell_size_rad, grid_size, grid)


def grid_to_image(grid, norm=0.0):
    """Image of a gridded [G, G] complex grid (complex64 -> FP32 path,
    complex128 -> FP64), mi355_grid_to_image_f / _d: the centred inverse 2-D
    DFT, real part, divided by norm (1 when norm is 0).  Returns a real
    [G, G] array."""
    g = np.asarray(grid)
    if g.dtype == np.complex64 or g.dtype == np.float32:
        real, fn = np.float32, _lib.load().mi355_grid_to_image_f
    elif g.dtype == np.complex128 or g.dtype == np.float64:
        real, fn = np.float64, _lib.load().mi355_grid_to_image_d
    else:
        raise ValueError("grid must be complex64/complex128 (or float [G, G, 2])")
    g = np.ascontiguousarray(g)
    n = g.size if np.iscomplexobj(g) else g.size // 2
    G = int(round(n ** 0.5))
    if G * G != n:
        raise ValueError("grid must be G x G complex")
    image = np.empty((G, G), real)
    fn(G, _ptr(g), float(norm), _ptr(image))
    _lib.check_last()
    return image
