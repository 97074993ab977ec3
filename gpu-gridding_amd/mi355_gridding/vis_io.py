"""The reference's input files from Python: OSKAR binary visibility files
(.vis), W-kernel FITS cubes, and the end-to-end driver over them.

Every function calls the library's native code (include/mi355_io.h); the
arrays are numpy.  ``read_vis`` mirrors what main.cpp reads block by block
(main.cpp:120-271); ``read_kernel`` mirrors read_kernel.c:35-118;
``grid_vis_file`` runs main.cpp's whole flow on the GPU.
"""
import ctypes

import numpy as np

from . import _lib

GROUP_HEADER, GROUP_BLOCK = 11, 12
TAG = dict(num_tags_per_block=2, amp_type=5, coord_precision=6, max_times_per_block=7,
           num_times_total=8, num_channels_total=10, num_stations=11, freq_start_hz=23)
BLOCK_DIMS, BLOCK_VIS, BLOCK_UU, BLOCK_VV, BLOCK_WW = 1, 3, 4, 5, 6


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _real(prec):
    return np.float64 if prec == _lib.BIN_DOUBLE else np.float32


class BinaryFile:
    """An OSKAR binary file opened with mi355_bin_open ('r' or 'w')."""

    def __init__(self, path, mode="r"):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.mi355_bin_open(path.encode(), mode.encode(), ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            _lib.check(self._lib.mi355_bin_close(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def num_tags(self):
        return self._lib.mi355_bin_num_tags(self.handle)

    def set_query_search_start(self, start):
        _lib.check(self._lib.mi355_bin_set_query_search_start(self.handle, int(start)))

    def query(self, data_type, group, tag, index=0):
        size = ctypes.c_size_t(0)
        i = ctypes.c_int(0)
        _lib.check(self._lib.mi355_bin_query(self.handle, data_type, group, tag, int(index),
                                             ctypes.byref(size), ctypes.byref(i)))
        return i.value, size.value

    def read(self, data_type, group, tag, index, out):
        """Reads a chunk's payload into the numpy array out (bytes checked)."""
        out = np.ascontiguousarray(out)
        _lib.check(self._lib.mi355_bin_read(self.handle, data_type, group, tag, int(index),
                                            out.nbytes, _ptr(out)))
        return out

    def write(self, data_type, group, tag, index, data):
        data = np.ascontiguousarray(data)
        _lib.check(self._lib.mi355_bin_write(self.handle, data_type, group, tag, int(index),
                                             data.nbytes, _ptr(data)))

    def vis_header(self):
        h = _lib.VisHeader()
        _lib.check(self._lib.mi355_vis_read_header(self.handle, ctypes.byref(h)))
        return h

    def vis_block(self, header, block, with_vis=True):
        """(uu, vv, ww, vis) of a block; coordinates in wavelengths, vis
        complex (None without with_vis)."""
        cap = header.max_times_per_block * header.num_baselines
        cr = _real(header.coord_precision)
        ar = _real(header.amp_type & 0x0F)
        uu, vv, ww = (np.empty(cap, cr) for _ in range(3))
        vis = np.empty(2 * cap, ar) if with_vis else None
        n = ctypes.c_size_t(0)
        _lib.check(self._lib.mi355_vis_read_block(
            self.handle, ctypes.byref(header), int(block), cap, ctypes.byref(n), _ptr(uu),
            _ptr(vv), _ptr(ww), _ptr(vis) if with_vis else None))
        n = n.value
        v = vis[:2 * n].view(np.complex128 if ar == np.float64 else np.complex64) \
            if with_vis else None
        return uu[:n], vv[:n], ww[:n], v


def write_vis(path, uu, vv, ww, vis, num_stations, times_per_block, freq_start_hz,
              precision=_lib.BIN_SINGLE):
    """Writes a single-channel, single-polarisation visibility file with the
    tags main.cpp reads: per block, DIM_START_AND_SIZE, the baseline
    coordinates in metres (uu, vv, ww given in metres) and the amplitudes.
    uu, vv, ww: [num_times, num_baselines]; vis: complex of the same shape."""
    uu, vv, ww = (np.asarray(x) for x in (uu, vv, ww))
    T, B = uu.shape
    assert B == num_stations * (num_stations - 1) // 2
    r = _real(precision)
    amp_type = precision | _lib.BIN_COMPLEX
    num_blocks = (T + times_per_block - 1) // times_per_block
    tags_per_block = 5
    with BinaryFile(path, "w") as f:
        ints = dict(num_tags_per_block=tags_per_block, amp_type=amp_type,
                    coord_precision=precision, max_times_per_block=times_per_block,
                    num_times_total=T, num_channels_total=1, num_stations=num_stations)
        for k, v in ints.items():
            f.write(_lib.BIN_INT, GROUP_HEADER, TAG[k], 0, np.array([v], np.int32))
        f.write(_lib.BIN_DOUBLE, GROUP_HEADER, TAG["freq_start_hz"], 0,
                np.array([freq_start_hz], np.float64))
        for b in range(num_blocks):
            t0, t1 = b * times_per_block, min(T, (b + 1) * times_per_block)
            f.write(_lib.BIN_INT, GROUP_BLOCK, BLOCK_DIMS, b,
                    np.array([t0, 0, t1 - t0, 1, B, num_stations], np.int32))
            f.write(amp_type, GROUP_BLOCK, BLOCK_VIS, b,
                    np.ascontiguousarray(vis[t0:t1]).astype(
                        np.complex128 if r == np.float64 else np.complex64))
            for tag, a in ((BLOCK_UU, uu), (BLOCK_VV, vv), (BLOCK_WW, ww)):
                f.write(precision, GROUP_BLOCK, tag, b, a[t0:t1].astype(r))
    # header chunks come first, so block b's tags start at chunk 8 + 5 b;
    # main.cpp starts its search for block b at b * num_tags_per_block, which
    # is at or before that


def read_kernel(path, precision=_lib.BIN_SINGLE, with_support=True):
    """(data [P, csh, csh], support or None, info dict), as read_kernel.c."""
    lib = _lib.load()
    info = _lib.KernelInfo()
    d = ctypes.c_void_p()
    s = ctypes.c_void_p()
    _lib.check(lib.mi355_fits_read_kernel(path.encode(), int(precision), ctypes.byref(info),
                                          ctypes.byref(d), ctypes.byref(s) if with_support
                                          else None))
    try:
        r = _real(precision)
        n = info.conv_size_half ** 2 * info.num_w_planes
        data = np.ctypeslib.as_array(ctypes.cast(d, ctypes.POINTER(
            ctypes.c_double if r == np.float64 else ctypes.c_float)), (n,)).copy()
        sup = np.ctypeslib.as_array(ctypes.cast(s, ctypes.POINTER(ctypes.c_int)),
                                    (info.num_w_planes,)).copy() if with_support else None
    finally:
        lib.mi355_io_free(d)
        if with_support:
            lib.mi355_io_free(s)
    return data.reshape(info.num_w_planes, info.conv_size_half, info.conv_size_half), sup, \
        info.as_dict()


def write_kernel(path, data, support, oversample, grid_size, cellsize_rad, w_scale):
    """Writes a kernel cube file read_kernel.c accepts; data [P, csh, csh]
    real (float32 -> FLOAT_IMG, float64 -> DOUBLE_IMG)."""
    data = np.ascontiguousarray(data)
    prec = _lib.BIN_DOUBLE if data.dtype == np.float64 else _lib.BIN_SINGLE
    data = data.astype(_real(prec))
    P, c1, c0 = data.shape
    info = _lib.KernelInfo(c0, P, int(oversample), int(grid_size), float(cellsize_rad),
                           float(w_scale))
    sup = None if support is None else np.ascontiguousarray(support, np.int32)
    _lib.check(_lib.load().mi355_fits_write_kernel(path.encode(), prec, ctypes.byref(info),
                                                   _ptr(data), _ptr(sup) if sup is not None
                                                   else None))


def write_image(path, image):
    image = np.ascontiguousarray(image)
    prec = _lib.BIN_DOUBLE if image.dtype == np.float64 else _lib.BIN_SINGLE
    image = image.astype(_real(prec))
    h, w = image.shape
    _lib.check(_lib.load().mi355_fits_write_image(path.encode(), prec, w, h, _ptr(image)))


def grid_vis_file(root, check=False, want_grid=True, image_path=None, ctx=None):
    """main.cpp's flow on the GPU (mi355_grid_vis_file): returns (grid or
    None, result dict).  ctx: an _lib.Context (one is created if None)."""
    lib = _lib.load()
    own = ctx is None
    ctx = ctx or _lib.Context(0)
    try:
        res = _lib.DriverResult()
        grid = None
        if want_grid:
            # size known only after reading the kernel header
            _, _, info = read_kernel(root + "_KERNELS_REAL.fits", with_support=False)
            with BinaryFile(root + ".vis") as f:
                prec = f.vis_header().amp_type & 0x0F
            G = info["grid_size"]
            grid = np.zeros((G, G), np.complex128 if prec == _lib.BIN_DOUBLE else np.complex64)
        _lib.check(lib.mi355_grid_vis_file(
            ctx.handle, root.encode(), int(bool(check)), _ptr(grid) if grid is not None else None,
            grid.size if grid is not None else 0,
            image_path.encode() if image_path else None, ctypes.byref(res)))
        return grid, res.as_dict()
    finally:
        if own:
            ctx.close()
