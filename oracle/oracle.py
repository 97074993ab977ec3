"""ctypes access to the CPU checkers -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, as the checker; the product library never does.

  liboracle.so   C restatement of oskar_grid_wproj_f/_d (grid_oracle.c)
  _ref/libref.so the reference's own oskar_grid_wproj.c compiled unmodified
                 (built only where the reference tree exists; optional)
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

P = ctypes.c_void_p
SZ = ctypes.c_size_t
I = ctypes.c_int
F = ctypes.c_float
D = ctypes.c_double

_GRID_F = [SZ, P, I, I, P, SZ, P, P, P, P, P, F, F, I, P, P, P]
_GRID_D = [SZ, P, I, I, P, SZ, P, P, P, P, P, D, D, I, P, P, P]

_W_WRITE_F = [SZ, P, P, P, F, I, P, P]
_W_WRITE_D = [SZ, P, P, P, D, I, P, P]
_W_READ_F = [SZ, P, P, P, P, F, I, P, P]
_W_READ_D = [SZ, P, P, P, P, D, I, P, P]

_oracle = None
_ref = None
_ref_io = None


def _weights_sigs(lib, prefix):
    for suffix, sig in (("write_f", _W_WRITE_F), ("write_d", _W_WRITE_D),
                        ("read_f", _W_READ_F), ("read_d", _W_READ_D)):
        fn = getattr(lib, prefix + suffix)
        fn.argtypes = sig
        fn.restype = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError("oracle not built: make -C oracle")
        lib = ctypes.CDLL(ORACLE_PATH)
        lib.oracle_grid_wproj_f.argtypes = _GRID_F
        lib.oracle_grid_wproj_d.argtypes = _GRID_D
        lib.oracle_grid_wproj_fsel_d.argtypes = _GRID_F
        lib.oracle_select_f.argtypes = [SZ, P, P, P, SZ, P, I, F, F, I] + [P] * 8
        lib.oracle_select_d.argtypes = [SZ, P, P, P, SZ, P, I, D, D, I] + [P] * 8
        lib.oracle_degrid_wproj_fsel_d.argtypes = [SZ, P, I, I, P, SZ, P, P, P, P, F, F, I,
                                                   P, I, P]
        lib.oracle_grid_wproj_fsel_d_tiles.argtypes = [SZ, P, I, I, P, SZ, P, P, P, P, P, F, F,
                                                       I, I, P, SZ, I, P]
        lib.oracle_norm_f.argtypes = [SZ, P, I, I, P, SZ, P, P, P, P, F, F, I, P, P]
        for fn in (lib.oracle_grid_wproj_f, lib.oracle_grid_wproj_d, lib.oracle_grid_wproj_fsel_d,
                   lib.oracle_select_f, lib.oracle_select_d, lib.oracle_degrid_wproj_fsel_d,
                   lib.oracle_grid_wproj_fsel_d_tiles, lib.oracle_norm_f):
            fn.restype = None
        _weights_sigs(lib, "oracle_grid_weights_")
        lib.oracle_grid_to_image_d.argtypes = [I, P, D, P]
        lib.oracle_grid_to_image_d.restype = I
        _oracle = lib
    return _oracle


def ref():
    """The compiled reference, or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        lib = ctypes.CDLL(REF_PATH)
        lib.oskar_grid_wproj_f.argtypes = _GRID_F
        lib.oskar_grid_wproj_d.argtypes = _GRID_D
        lib.oskar_grid_wproj_f.restype = None
        lib.oskar_grid_wproj_d.restype = None
        lib.check_value_float.argtypes = [F, F]
        lib.check_value_float.restype = I
        _weights_sigs(lib, "oskar_grid_weights_")
        _ref = lib
    return _ref


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _inputs(case, real):
    c = case
    return dict(
        support=np.ascontiguousarray(c["support"], np.int32),
        conv=np.ascontiguousarray(c["conv_func"], real),
        uu=np.ascontiguousarray(c["uu"], real),
        vv=np.ascontiguousarray(c["vv"], real),
        ww=np.ascontiguousarray(c["ww"], real),
        vis=np.ascontiguousarray(c["vis"], real),
        weight=np.ascontiguousarray(c["weight"], real),
    )


def _call(fn, case, real, grid_real, scal, grid=None, norm=0.0):
    x = _inputs(case, real)
    G = int(case["grid_size"])
    N = int(case["num_points"])
    if grid is None:
        grid = np.zeros((G, G, 2), grid_real)
    skipped = ctypes.c_size_t(0)
    nrm = ctypes.c_double(norm)
    fn(int(case["num_w_planes"]), _p(x["support"]), int(case["oversample"]),
       int(case["conv_size_half"]), _p(x["conv"]), N, _p(x["uu"]), _p(x["vv"]), _p(x["ww"]),
       _p(x["vis"]), _p(x["weight"]), scal(float(case["cell_size_rad"])),
       scal(float(case["w_scale"])), G, ctypes.byref(skipped), ctypes.byref(nrm), _p(grid))
    return grid, nrm.value, skipped.value


def grid_f(case, use_ref=False, grid=None, norm=0.0):
    """oskar_grid_wproj_f semantics -> (grid float32 [G,G,2], norm, num_skipped)."""
    lib = ref() if use_ref else oracle()
    fn = lib.oskar_grid_wproj_f if use_ref else lib.oracle_grid_wproj_f
    return _call(fn, case, np.float32, np.float32, ctypes.c_float, grid, norm)


def grid_d(case, use_ref=False, grid=None, norm=0.0):
    """oskar_grid_wproj_d semantics on the case's values promoted to double."""
    lib = ref() if use_ref else oracle()
    fn = lib.oskar_grid_wproj_d if use_ref else lib.oracle_grid_wproj_d
    return _call(fn, case, np.float64, np.float64, ctypes.c_double, grid, norm)


def grid_fsel_d(case, grid=None, norm=0.0):
    """FP32 selection, FP64 accumulation (the FP64 truth for FP32 gridding)."""
    lib = oracle()
    x = _inputs(case, np.float32)
    G = int(case["grid_size"])
    if grid is None:
        grid = np.zeros((G, G, 2), np.float64)
    skipped = ctypes.c_size_t(0)
    nrm = ctypes.c_double(norm)
    lib.oracle_grid_wproj_fsel_d(
        int(case["num_w_planes"]), _p(x["support"]), int(case["oversample"]),
        int(case["conv_size_half"]), _p(x["conv"]), int(case["num_points"]), _p(x["uu"]),
        _p(x["vv"]), _p(x["ww"]), _p(x["vis"]), _p(x["weight"]),
        ctypes.c_float(float(case["cell_size_rad"])), ctypes.c_float(float(case["w_scale"])),
        G, ctypes.byref(skipped), ctypes.byref(nrm), _p(grid))
    return grid, nrm.value, skipped.value


def norm_f(case, norm=0.0):
    """(norm, num_skipped) of oskar_grid_wproj_f on case, without the grid
    (bit-identical to grid_f's)."""
    lib = oracle()
    x = _inputs(case, np.float32)
    skipped = ctypes.c_size_t(0)
    nrm = ctypes.c_double(norm)
    lib.oracle_norm_f(int(case["num_w_planes"]), _p(x["support"]), int(case["oversample"]),
                      int(case["conv_size_half"]), _p(x["conv"]), int(case["num_points"]),
                      _p(x["uu"]), _p(x["vv"]), _p(x["ww"]), _p(x["weight"]),
                      ctypes.c_float(float(case["cell_size_rad"])),
                      ctypes.c_float(float(case["w_scale"])), int(case["grid_size"]),
                      ctypes.byref(skipped), ctypes.byref(nrm))
    return nrm.value, skipped.value


def select_f(case):
    """Per-visibility selection of oskar_grid_wproj_f: dict of int32 arrays."""
    lib = oracle()
    x = _inputs(case, np.float32)
    N = int(case["num_points"])
    out = {k: np.zeros(N, np.int32) for k in
           ("grid_u", "grid_v", "off_u", "off_v", "plane", "support", "conj", "skip")}
    lib.oracle_select_f(N, _p(x["uu"]), _p(x["vv"]), _p(x["ww"]), int(case["num_w_planes"]),
                        _p(x["support"]), int(case["oversample"]),
                        ctypes.c_float(float(case["cell_size_rad"])),
                        ctypes.c_float(float(case["w_scale"])), int(case["grid_size"]),
                        *[_p(out[k]) for k in ("grid_u", "grid_v", "off_u", "off_v", "plane",
                                               "support", "conj", "skip")])
    return out


def select_d(case):
    lib = oracle()
    x = _inputs(case, np.float64)
    N = int(case["num_points"])
    out = {k: np.zeros(N, np.int32) for k in
           ("grid_u", "grid_v", "off_u", "off_v", "plane", "support", "conj", "skip")}
    lib.oracle_select_d(N, _p(x["uu"]), _p(x["vv"]), _p(x["ww"]), int(case["num_w_planes"]),
                        _p(x["support"]), int(case["oversample"]),
                        ctypes.c_double(float(case["cell_size_rad"])),
                        ctypes.c_double(float(case["w_scale"])), int(case["grid_size"]),
                        *[_p(out[k]) for k in ("grid_u", "grid_v", "off_u", "off_v", "plane",
                                               "support", "conj", "skip")])
    return out


def degrid_fsel_d(case, grid, apply_weight=True):
    """FP64 adjoint of grid_fsel_d: out [N, 2] float64."""
    lib = oracle()
    x = _inputs(case, np.float32)
    N = int(case["num_points"])
    g = np.ascontiguousarray(grid, np.float64)
    out = np.zeros((N, 2), np.float64)
    lib.oracle_degrid_wproj_fsel_d(
        int(case["num_w_planes"]), _p(x["support"]), int(case["oversample"]),
        int(case["conv_size_half"]), _p(x["conv"]), N, _p(x["uu"]), _p(x["vv"]), _p(x["ww"]),
        _p(x["weight"]), ctypes.c_float(float(case["cell_size_rad"])),
        ctypes.c_float(float(case["w_scale"])), int(case["grid_size"]), _p(g),
        1 if apply_weight else 0, _p(out))
    return out


def tiles_fsel_d(case, tiles_xy, tile=64, threads=8):
    """FP64 truth on selected tiles: [n, tile, tile, 2] float64."""
    lib = oracle()
    x = _inputs(case, np.float32)
    t = np.ascontiguousarray(tiles_xy, np.int32).reshape(-1, 2)
    out = np.zeros((t.shape[0], tile, tile, 2), np.float64)
    lib.oracle_grid_wproj_fsel_d_tiles(
        int(case["num_w_planes"]), _p(x["support"]), int(case["oversample"]),
        int(case["conv_size_half"]), _p(x["conv"]), int(case["num_points"]), _p(x["uu"]),
        _p(x["vv"]), _p(x["ww"]), _p(x["vis"]), _p(x["weight"]),
        ctypes.c_float(float(case["cell_size_rad"])), ctypes.c_float(float(case["w_scale"])),
        int(case["grid_size"]), int(tile), _p(t), t.shape[0], int(threads), _p(out))
    return out


def weights_write(uu, vv, weight, cell_size_rad, grid_size, grid, use_ref=False):
    """oskar_grid_weights_write_f / _d semantics (precision from uu's dtype):
    grid ([G, G] real) += weights per cell in visibility order; returns
    num_skipped."""
    real = np.float32 if np.asarray(uu).dtype == np.float32 else np.float64
    lib = ref() if use_ref else oracle()
    if lib is None:
        raise RuntimeError("compiled reference not built (make -C oracle ref)")
    fn = (lib.oskar_grid_weights_write_f if use_ref else lib.oracle_grid_weights_write_f) \
        if real == np.float32 else \
        (lib.oskar_grid_weights_write_d if use_ref else lib.oracle_grid_weights_write_d)
    uu, vv, weight = (np.ascontiguousarray(x, real) for x in (uu, vv, weight))
    assert grid.dtype == real and grid.flags.c_contiguous
    sk = ctypes.c_size_t(0)
    scal = ctypes.c_float if real == np.float32 else ctypes.c_double
    fn(uu.size, _p(uu), _p(vv), _p(weight), scal(float(cell_size_rad)), int(grid_size),
       ctypes.byref(sk), _p(grid))
    return sk.value


def weights_read(uu, vv, weight_in, weight_out, cell_size_rad, grid_size, grid, use_ref=False):
    """oskar_grid_weights_read_f / _d semantics: weight_out[i] = weight_in[i]
    / grid[cell] (0 where the cell total is 0) for points inside the grid;
    returns num_skipped."""
    real = np.float32 if np.asarray(uu).dtype == np.float32 else np.float64
    lib = ref() if use_ref else oracle()
    if lib is None:
        raise RuntimeError("compiled reference not built (make -C oracle ref)")
    fn = (lib.oskar_grid_weights_read_f if use_ref else lib.oracle_grid_weights_read_f) \
        if real == np.float32 else \
        (lib.oskar_grid_weights_read_d if use_ref else lib.oracle_grid_weights_read_d)
    uu, vv, weight_in = (np.ascontiguousarray(x, real) for x in (uu, vv, weight_in))
    grid = np.ascontiguousarray(grid, real)
    assert weight_out.dtype == real and weight_out.flags.c_contiguous
    sk = ctypes.c_size_t(0)
    scal = ctypes.c_float if real == np.float32 else ctypes.c_double
    fn(uu.size, _p(uu), _p(vv), _p(weight_in), _p(weight_out), scal(float(cell_size_rad)),
       int(grid_size), ctypes.byref(sk), _p(grid))
    return sk.value


def grid_to_image(grid, norm=0.0):
    """Centred inverse DFT of a [G, G] complex grid (any complex or [G, G, 2]
    real layout), real part scaled by 1/norm, in FP64 by direct summation
    (oracle_grid_to_image_d).  Returns float64 [G, G]."""
    g = np.asarray(grid)
    if np.iscomplexobj(g):
        g = g.astype(np.complex128).view(np.float64)
    g = np.ascontiguousarray(g, np.float64)
    G = int(round((g.size // 2) ** 0.5))
    assert g.size == 2 * G * G, "grid must be G x G complex"
    out = np.empty((G, G), np.float64)
    if oracle().oracle_grid_to_image_d(G, _p(g), float(norm), _p(out)):
        raise ValueError("grid size must be even")
    return out


def ref_io():
    """The reference's input-file code compiled from its sources
    (oracle/_ref/libref_io.so: OSKAR binary library, cfitsio 3.41,
    utils/read_vis.c and read_kernel.c), or None where it was not built."""
    global _ref_io
    if _ref_io is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libref_io.so")
        if not os.path.exists(path):
            return None
        lib = ctypes.CDLL(path)
        UB, IP = ctypes.c_ubyte, ctypes.POINTER(I)
        sigs = {
            "oskar_binary_create": (P, [ctypes.c_char_p, ctypes.c_char, IP]),
            "oskar_binary_free": (None, [P]),
            "oskar_binary_num_tags": (I, [P]),
            "oskar_binary_write": (None, [P, UB, UB, UB, I, SZ, P, IP]),
            "oskar_binary_write_ext": (None, [P, UB, ctypes.c_char_p, ctypes.c_char_p, I, SZ,
                                              P, IP]),
            "oskar_binary_read": (None, [P, UB, UB, UB, I, SZ, P, IP]),
            "oskar_binary_query": (I, [P, UB, UB, UB, I, ctypes.POINTER(SZ), IP]),
            "oskar_binary_set_query_search_start": (None, [P, I, IP]),
            "read_coords": (None, [P, I, I, SZ, P, P, P, IP]),
            "scale_coords": (None, [D, I, SZ, P, P, P]),
            "read_kernel": (P, [ctypes.c_char_p, I, IP, IP, ctypes.POINTER(P), IP, IP,
                                ctypes.POINTER(D), ctypes.POINTER(D), IP]),
            # cfitsio (fits_create_file, _create_img, _write_key, _write_pix,
            # _create_tbl, _write_col, _close_file)
            "ffinit": (I, [ctypes.POINTER(P), ctypes.c_char_p, IP]),
            "ffcrim": (I, [P, I, I, ctypes.POINTER(ctypes.c_long), IP]),
            "ffpky": (I, [P, I, ctypes.c_char_p, P, ctypes.c_char_p, IP]),
            "ffppx": (I, [P, I, ctypes.POINTER(ctypes.c_long), ctypes.c_longlong, P, IP]),
            "ffcrtb": (I, [P, I, ctypes.c_longlong, I, ctypes.POINTER(ctypes.c_char_p),
                           ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p),
                           ctypes.c_char_p, IP]),
            "ffpcl": (I, [P, I, I, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong, P,
                          IP]),
            "ffclos": (I, [P, IP]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _ref_io = lib
    return _ref_io


# cfitsio codes (fitsio.h:219-270)
TSHORT, TINT, TFLOAT, TDOUBLE = 21, 31, 42, 82
SHORT_IMG, FLOAT_IMG, DOUBLE_IMG, BINARY_TBL = 16, -32, -64, 2
