"""ctypes bindings of libmi355_gridder.so (the C ABI in include/mi355_gridder.h
and include/mi355_io.h).

The library is loaded from the package's own ``lib/`` directory, where the
build puts it; there is no fallback implementation, so a missing or stale
library raises instead of computing anything elsewhere.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(os.path.dirname(_HERE), "lib")
GRIDDER_PATH = os.path.join(LIB_DIR, "libmi355_gridder.so")
SYNTH_PATH = os.path.join(LIB_DIR, "libmi355_synth.so")

# Status codes (include/mi355_gridder.h).
OK = 0
ERR_INVALID_ARG = 1
ERR_KERNEL_TOO_SMALL = 2
ERR_NO_DEVICE = 3
ERR_HIP = 4
ERR_OUT_OF_MEMORY = 5
ERR_TOO_LARGE = 6
ERR_IO = 7
ERR_FORMAT = 8

# OSKAR binary payload types (include/mi355_io.h).
BIN_CHAR, BIN_INT, BIN_SINGLE, BIN_DOUBLE, BIN_COMPLEX, BIN_MATRIX = 1, 2, 4, 8, 0x20, 0x40

OPT_VISITS_PER_ITEM = 1
OPT_COMPENSATED = 2
OPT_PHASE_TIMING = 3
OPT_CACHE_KERNELS = 4
OPT_GRID_KERNEL = 5
OPT_MIN_STAGE_RUN = 6
OPT_MIN_STAGE_RUN_W33 = 9
OPT_MIN_STAGE_RUN_W65 = 10
OPT_MIN_STAGE_RUN_W97 = 11
OPT_MIN_STAGE_RUN_W111 = 12
OPT_AUTOTUNE_CACHE = 13
OPT_ITEM_GROUP = 7
OPT_MAX_WORKSPACE = 8


class GridderError(RuntimeError):
    """A non-zero status from the gridder library."""

    def __init__(self, status, message):
        super().__init__("mi355 gridder status %d: %s" % (status, message))
        self.status = status


class Stats(ctypes.Structure):
    _fields_ = [
        ("ms_prepare", ctypes.c_double),
        ("ms_select", ctypes.c_double),
        ("ms_bucket", ctypes.c_double),
        ("ms_grid", ctypes.c_double),
        ("ms_merge", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("num_points", ctypes.c_ulonglong),
        ("num_skipped", ctypes.c_ulonglong),
        ("num_visits", ctypes.c_ulonglong),
        ("num_items", ctypes.c_ulonglong),
        ("num_taps", ctypes.c_ulonglong),
        ("num_merge_tiles", ctypes.c_ulonglong),
        ("visits_per_item", ctypes.c_ulonglong),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class JobF(ctypes.Structure):
    _fields_ = [
        ("vis_offset", ctypes.c_size_t),
        ("num_points", ctypes.c_size_t),
        ("cell_size_rad", ctypes.c_float),
        ("w_scale", ctypes.c_float),
        ("grid_size", ctypes.c_int),
        ("grid", ctypes.c_void_p),
        ("norm", ctypes.c_void_p),
        ("num_skipped", ctypes.c_void_p),
    ]


class VisHeader(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "num_tags_per_block", "amp_type", "coord_precision", "max_times_per_block",
        "num_times_total", "num_channels_total", "num_stations")] + [
        ("freq_start_hz", ctypes.c_double), ("num_baselines", ctypes.c_int),
        ("num_blocks", ctypes.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class KernelInfo(ctypes.Structure):
    _fields_ = [("conv_size_half", ctypes.c_int), ("num_w_planes", ctypes.c_int),
                ("oversample", ctypes.c_int), ("grid_size", ctypes.c_int),
                ("cellsize_rad", ctypes.c_double), ("w_scale", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class DriverResult(ctypes.Structure):
    _fields_ = [("num_blocks", ctypes.c_int), ("grid_size", ctypes.c_int),
                ("num_w_planes", ctypes.c_int), ("precision", ctypes.c_int),
                ("num_points", ctypes.c_ulonglong), ("num_skipped_total", ctypes.c_ulonglong),
                ("num_skipped_last", ctypes.c_ulonglong), ("norm", ctypes.c_double),
                ("seconds_read", ctypes.c_double), ("seconds_weights", ctypes.c_double),
                ("seconds_grid", ctypes.c_double), ("check_max_rel", ctypes.c_double),
                ("check_l2_rel", ctypes.c_double), ("seconds_wall", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


P = ctypes.c_void_p
SZ = ctypes.c_size_t
I = ctypes.c_int
F = ctypes.c_float
D = ctypes.c_double

# name -> (restype, argtypes); every function the header declares.
SIGNATURES = {
    "mi355_grid_wproj_d": (None, [SZ, P, I, I, P, SZ, P, P, P, P, P, D, D, I, P, P, P]),
    "mi355_grid_wproj_f": (None, [SZ, P, I, I, P, SZ, P, P, P, P, P, F, F, I, P, P, P]),
    "mi355_grid_last_error": (I, []),
    "mi355_grid_last_error_message": (ctypes.c_char_p, []),
    "mi355_grid_error_string": (ctypes.c_char_p, [I]),
    "mi355_ctx_create": (I, [I, ctypes.POINTER(P)]),
    "mi355_ctx_destroy": (I, [P]),
    "mi355_default_ctx": (I, [ctypes.POINTER(P)]),
    "mi355_ctx_set_stream": (I, [P, P]),
    "mi355_ctx_get_stream": (I, [P, ctypes.POINTER(P)]),
    "mi355_ctx_set_option": (I, [P, I, ctypes.c_longlong]),
    "mi355_ctx_get_option": (I, [P, I, ctypes.POINTER(ctypes.c_longlong)]),
    "mi355_ctx_get_stats": (I, [P, ctypes.POINTER(Stats)]),
    "mi355_ctx_get_buckets": (I, [P, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(I), P, P,
                                  P, ctypes.c_ulonglong]),
    "mi355_grid_wproj_f_dev": (I, [P, SZ, P, I, I, P, SZ, P, P, P, P, P, F, F, I, P, P, P]),
    "mi355_grid_wproj_f_dev_noalloc": (I, [P, SZ, P, I, I, P, SZ, P, P, P, P, P, F, F, I, P, P, P]),
    "mi355_grid_wproj_d_dev": (I, [P, SZ, P, I, I, P, SZ, P, P, P, P, P, D, D, I, P, P, P]),
    "mi355_grid_wproj_fsel_d_dev": (I, [P, SZ, P, I, I, P, SZ, P, P, P, P, P, F, F, I, P, P, P]),
    "mi355_select_wproj_f_dev": (I, [P, SZ, P, I, SZ, P, P, P, F, F, I, P]),
    "mi355_select_wproj_d_dev": (I, [P, SZ, P, I, SZ, P, P, P, D, D, I, P]),
    "mi355_degrid_wproj_f_dev": (I, [P, SZ, P, I, I, P, SZ, P, P, P, P, F, F, I, P, I, P]),
    "mi355_degrid_wproj_d_dev": (I, [P, SZ, P, I, I, P, SZ, P, P, P, P, D, D, I, P, I, P]),
    "mi355_grid_wproj_f_batched_dev": (I, [P, SZ, P, I, I, P, I, ctypes.POINTER(JobF), P, P, P, P, P]),
    "mi355_grid_wproj_f_batched_graph_dev": (I, [P, SZ, P, I, I, P, I, ctypes.POINTER(JobF),
                                                 P, P, P, P, P]),
    "mi355_autotune_f_batched_dev": (I, [P, SZ, P, I, I, P, I, ctypes.POINTER(JobF), P, P, P, P, P,
                                         ctypes.POINTER(D)]),
    "mi355_grid_weights_write_d": (None, [SZ, P, P, P, D, I, P, P]),
    "mi355_grid_weights_write_f": (None, [SZ, P, P, P, F, I, P, P]),
    "mi355_grid_weights_read_d": (None, [SZ, P, P, P, P, D, I, P, P]),
    "mi355_grid_weights_read_f": (None, [SZ, P, P, P, P, F, I, P, P]),
    "mi355_grid_weights_write_f_dev": (I, [P, SZ, P, P, P, F, I, P, P]),
    "mi355_grid_weights_write_d_dev": (I, [P, SZ, P, P, P, D, I, P, P]),
    "mi355_grid_weights_read_f_dev": (I, [P, SZ, P, P, P, P, F, I, P, P]),
    "mi355_grid_weights_read_d_dev": (I, [P, SZ, P, P, P, P, D, I, P, P]),
    "mi355_grid_to_image_f": (None, [I, P, D, P]),
    "mi355_grid_to_image_d": (None, [I, P, D, P]),
    "mi355_grid_to_image_f_dev": (I, [P, I, P, D, P]),
    "mi355_grid_to_image_d_dev": (I, [P, I, P, D, P]),
    # mi355_io.h
    "mi355_bin_open": (I, [ctypes.c_char_p, ctypes.c_char, ctypes.POINTER(P)]),
    "mi355_bin_close": (I, [P]),
    "mi355_bin_num_tags": (I, [P]),
    "mi355_bin_set_query_search_start": (I, [P, I]),
    "mi355_bin_query": (I, [P, ctypes.c_ubyte, ctypes.c_ubyte, ctypes.c_ubyte, I,
                            ctypes.POINTER(SZ), ctypes.POINTER(I)]),
    "mi355_bin_read": (I, [P, ctypes.c_ubyte, ctypes.c_ubyte, ctypes.c_ubyte, I, SZ, P]),
    "mi355_bin_write": (I, [P, ctypes.c_ubyte, ctypes.c_ubyte, ctypes.c_ubyte, I, SZ, P]),
    "mi355_vis_read_header": (I, [P, ctypes.POINTER(VisHeader)]),
    "mi355_vis_read_block": (I, [P, ctypes.POINTER(VisHeader), I, SZ, ctypes.POINTER(SZ),
                                 P, P, P, P]),
    "mi355_fits_read_kernel": (I, [ctypes.c_char_p, I, ctypes.POINTER(KernelInfo),
                                   ctypes.POINTER(P), ctypes.POINTER(P)]),
    "mi355_fits_write_kernel": (I, [ctypes.c_char_p, I, ctypes.POINTER(KernelInfo), P, P]),
    "mi355_fits_write_image": (I, [ctypes.c_char_p, I, I, I, P]),
    "mi355_io_free": (None, [P]),
    "mi355_grid_vis_file": (I, [P, ctypes.c_char_p, I, P, SZ, ctypes.c_char_p,
                                ctypes.POINTER(DriverResult)]),
}

_lib = None
_synth = None


def load():
    """Load (once) and return the gridder library with prototypes set."""
    global _lib
    if _lib is None:
        if not os.path.exists(GRIDDER_PATH):
            raise GridderError(-1, "%s not built; run __graft_entry__.build()" % GRIDDER_PATH)
        lib = ctypes.CDLL(GRIDDER_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(status):
    if status != OK:
        lib = load()
        msg = lib.mi355_grid_last_error_message()
        raise GridderError(status, msg.decode() if msg else "")
    return status


def check_last():
    """Raise if the last void drop-in call on this thread failed."""
    return check(load().mi355_grid_last_error())


def load_synth():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise GridderError(-1, "%s not built; run __graft_entry__.build()" % SYNTH_PATH)
        lib = ctypes.CDLL(SYNTH_PATH)
        lib.mi355_synth_support.restype = I
        lib.mi355_synth_support.argtypes = [I, I, P]
        lib.mi355_synth_kernels.restype = I
        lib.mi355_synth_kernels.argtypes = [I, I, I, I, P, P]
        lib.mi355_synth_cell_size.restype = D
        lib.mi355_synth_cell_size.argtypes = [I, I]
        lib.mi355_synth_vis.restype = I
        lib.mi355_synth_vis.argtypes = [ctypes.c_ulonglong, I, SZ, SZ, SZ, I, P, P, P, P, P,
                                        ctypes.POINTER(D)]
        _synth = lib
    return _synth


class Context:
    """Owns an mi355_ctx (persistent device workspace + stream binding)."""

    def __init__(self, device=0):
        lib = load()
        h = P()
        check(lib.mi355_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self._lib = lib

    def set_stream(self, stream_handle):
        check(self._lib.mi355_ctx_set_stream(self.handle, P(stream_handle)))

    def set_option(self, key, value):
        check(self._lib.mi355_ctx_set_option(self.handle, key, int(value)))

    def get_option(self, key):
        v = ctypes.c_longlong(0)
        check(self._lib.mi355_ctx_get_option(self.handle, key, ctypes.byref(v)))
        return v.value

    def stats(self):
        s = Stats()
        check(self._lib.mi355_ctx_get_stats(self.handle, ctypes.byref(s)))
        return s.as_dict()

    def close(self):
        if self.handle:
            self._lib.mi355_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
