"""The library's readers and writers of the reference's input files
(include/mi355_io.h), against the reference's own code compiled from its
sources (oracle/_ref/libref_io.so: the OSKAR binary library, cfitsio 3.41,
utils/read_vis.c and read_kernel.c).

1. OSKAR binary (.vis): files written by the reference library read
   natively, files written natively read by the reference library (which
   checks every CRC), bit for bit, including the coordinate scaling of
   read_vis.c (scale_coords) and the query-search-start semantics main.cpp
   relies on; corrupted payloads are refused by both.
2. FITS kernel cubes: files written natively read by the reference's
   read_kernel (cfitsio), and files written by cfitsio read natively, with
   the same dimensions, keywords, data and supports, for FLOAT_IMG,
   DOUBLE_IMG and 16-bit integer images and 32- and 16-bit support columns.
Round trips through the native code alone run where the reference build is
absent; the rest skip there.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle
from mi355_gridding import _lib, vis_io

C = 299792458.0


def _ref():
    r = oracle.ref_io()
    if r is None:
        pytest.skip("oracle/_ref/libref_io.so not built (needs the reference sources)")
    return r


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _case(rng, S=7, T=8, prec=np.float32):
    B = S * (S - 1) // 2
    uu, vv, ww = (rng.normal(0, 300, (T, B)) for _ in range(3))
    vis = rng.normal(size=(T, B)) + 1j * rng.normal(size=(T, B))
    return S, B, T, uu, vv, ww, vis


# ---------------------------------------------------------------------------
# OSKAR binary
# ---------------------------------------------------------------------------
def _ref_write_vis(r, path, S, T, tpb, freq, uu, vv, ww, vis, prec):
    """The same tags as vis_io.write_vis, through oskar_binary_write."""
    st = ctypes.c_int(0)
    h = r.oskar_binary_create(path.encode(), b"w", ctypes.byref(st))
    assert st.value == 0
    B = uu.shape[1]
    real = np.float64 if prec == _lib.BIN_DOUBLE else np.float32
    amp = prec | _lib.BIN_COMPLEX

    def w(t, g, tag, idx, a):
        a = np.ascontiguousarray(a)
        r.oskar_binary_write(h, t, g, tag, idx, a.nbytes, _p(a), ctypes.byref(st))
        assert st.value == 0

    for k, v in dict(num_tags_per_block=5, amp_type=amp, coord_precision=prec,
                     max_times_per_block=tpb, num_times_total=T, num_channels_total=1,
                     num_stations=S).items():
        w(_lib.BIN_INT, 11, vis_io.TAG[k], 0, np.array([v], np.int32))
    w(_lib.BIN_DOUBLE, 11, 23, 0, np.array([freq]))
    for b in range((T + tpb - 1) // tpb):
        t0, t1 = b * tpb, min(T, (b + 1) * tpb)
        w(_lib.BIN_INT, 12, 1, b, np.array([t0, 0, t1 - t0, 1, B, S], np.int32))
        w(amp, 12, 3, b, vis[t0:t1].astype(np.complex128 if real == np.float64 else np.complex64))
        for tag, a in ((4, uu), (5, vv), (6, ww)):
            w(prec, 12, tag, b, a[t0:t1].astype(real))
    r.oskar_binary_free(h)


def _ref_read_block(r, path, b, tpb, prec):
    """main.cpp's reads of block b (main.cpp:213-227, :252-271) through the
    reference's oskar_binary_read, read_coords and scale_coords."""
    st = ctypes.c_int(0)
    h = r.oskar_binary_create(path.encode(), b"r", ctypes.byref(st))
    assert st.value == 0
    hdr = {}
    for k in ("num_tags_per_block", "num_stations", "amp_type"):
        v = np.zeros(1, np.int32)
        r.oskar_binary_read(h, _lib.BIN_INT, 11, vis_io.TAG[k], 0, 4, _p(v), ctypes.byref(st))
        hdr[k] = int(v[0])
    freq = np.zeros(1)
    r.oskar_binary_read(h, _lib.BIN_DOUBLE, 11, 23, 0, 8, _p(freq), ctypes.byref(st))
    r.oskar_binary_set_query_search_start(h, b * hdr["num_tags_per_block"], ctypes.byref(st))
    dims = np.zeros(6, np.int32)
    r.oskar_binary_read(h, _lib.BIN_INT, 12, 1, b, 24, _p(dims), ctypes.byref(st))
    S = hdr["num_stations"]
    n = int(dims[2]) * S * (S - 1) // 2
    real = np.float64 if prec == _lib.BIN_DOUBLE else np.float32
    uu, vv, ww = (np.zeros(n, real) for _ in range(3))
    r.read_coords(h, b, prec, n, _p(uu), _p(vv), _p(ww), ctypes.byref(st))
    r.scale_coords(float(freq[0]), prec, n, _p(uu), _p(vv), _p(ww))
    vis = np.zeros(n, np.complex128 if real == np.float64 else np.complex64)
    r.oskar_binary_read(h, hdr["amp_type"], 12, 3, b, vis.nbytes, _p(vis), ctypes.byref(st))
    status = st.value
    r.oskar_binary_free(h)
    return status, uu, vv, ww, vis


def _same(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.dtype == b.dtype and a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.mark.parametrize("prec", [_lib.BIN_SINGLE, _lib.BIN_DOUBLE])
def test_reference_written_vis_read_natively(tmp_path, prec):
    r = _ref()
    rng = np.random.default_rng(prec)
    S, B, T, uu, vv, ww, vis = _case(rng)
    path = str(tmp_path / "ref.vis")
    _ref_write_vis(r, path, S, T, 3, 1.42e8, uu, vv, ww, vis, prec)
    with vis_io.BinaryFile(path) as f:
        h = f.vis_header()
        assert (h.num_stations, h.num_baselines, h.num_blocks, h.max_times_per_block,
                h.num_times_total, h.freq_start_hz) == (S, B, 3, 3, T, 1.42e8)
        for b in range(h.num_blocks):
            got = f.vis_block(h, b)
            st, *want = _ref_read_block(r, path, b, 3, prec)
            assert st == 0
            for g, w in zip(got, want):
                assert _same(g, w)


@pytest.mark.parametrize("prec", [_lib.BIN_SINGLE, _lib.BIN_DOUBLE])
def test_native_written_vis_read_by_reference(tmp_path, prec):
    r = _ref()
    rng = np.random.default_rng(10 + prec)
    S, B, T, uu, vv, ww, vis = _case(rng, S=9, T=10)
    path = str(tmp_path / "native.vis")
    vis_io.write_vis(path, uu, vv, ww, vis, S, 4, 1.0e8, precision=prec)
    with vis_io.BinaryFile(path) as f:
        h = f.vis_header()
        for b in range(h.num_blocks):
            got = f.vis_block(h, b)
            st, *want = _ref_read_block(r, path, b, 4, prec)
            assert st == 0                                  # every CRC checked
            for g, w in zip(got, want):
                assert _same(g, w)
    # the raw payload written is the input (metres), not a rounded copy
    real = np.float64 if prec == _lib.BIN_DOUBLE else np.float32
    with vis_io.BinaryFile(path) as f:
        out = f.read(prec, 12, 4, 1, np.zeros(4 * B, real))
    assert _same(out, uu[4:8].astype(real).ravel())


def test_corrupt_payload_refused_by_both(tmp_path):
    r = _ref()
    rng = np.random.default_rng(3)
    S, B, T, uu, vv, ww, vis = _case(rng)
    path = str(tmp_path / "c.vis")
    vis_io.write_vis(path, uu, vv, ww, vis, S, 8, 1e8)
    with vis_io.BinaryFile(path) as f:
        i, size = f.query(_lib.BIN_SINGLE, 12, 5, 0)
    raw = bytearray(open(path, "rb").read())
    # payload of chunk i: locate by re-reading the index through the file
    # layout (64-byte header, 20-byte tags, payload, 4-byte CRC)
    off = 64
    for _ in range(i):
        blk = int.from_bytes(raw[off + 12:off + 20], "little")
        off += 20 + blk
    raw[off + 20 + 5] ^= 0x40
    open(path, "wb").write(bytes(raw))
    with vis_io.BinaryFile(path) as f:
        h = f.vis_header()
        with pytest.raises(_lib.GridderError) as e:
            f.vis_block(h, 0)
        assert e.value.status == _lib.ERR_FORMAT and "CRC" in str(e.value)
    st, *_ = _ref_read_block(r, path, 0, 8, _lib.BIN_SINGLE)
    assert st != 0


def test_query_semantics_match_reference(tmp_path):
    """Repeated (group, tag, index) keys, extended tags and search starts:
    the same chunk is found, and a search past the last match fails."""
    r = _ref()
    path = str(tmp_path / "q.vis")
    st = ctypes.c_int(0)
    h = r.oskar_binary_create(path.encode(), b"w", ctypes.byref(st))
    seq = [(2, 12, 1, 0), (4, 12, 1, 0), (2, 12, 1, 1), ("ext", 0), (2, 12, 1, 0), (8, 3, 7, 5)]
    for k, e in enumerate(seq):
        a = np.arange(k + 1, dtype=np.int32 if k % 2 == 0 else np.float32)
        if e[0] == "ext":
            a = np.arange(3, dtype=np.int32)
            r.oskar_binary_write_ext(h, 2, b"grp", b"tagname", 0, a.nbytes, _p(a),
                                     ctypes.byref(st))
        else:
            t = e[0]
            a = np.arange(k + 1, dtype={2: np.int32, 4: np.float32, 8: np.float64}[t])
            r.oskar_binary_write(h, t, e[1], e[2], e[3], a.nbytes, _p(a), ctypes.byref(st))
    r.oskar_binary_free(h)
    assert st.value == 0
    st = ctypes.c_int(0)
    rh = r.oskar_binary_create(path.encode(), b"r", ctypes.byref(st))
    with vis_io.BinaryFile(path) as f:
        assert f.num_tags() == r.oskar_binary_num_tags(rh) == len(seq)
        for start in range(len(seq)):
            r.oskar_binary_set_query_search_start(rh, start, ctypes.byref(st))
            f.set_query_search_start(start)
            for q in [(2, 12, 1, 0), (0, 12, 1, 0), (4, 12, 1, 0), (2, 12, 1, 1), (8, 3, 7, 5),
                      (0, 3, 7, 5)]:
                st = ctypes.c_int(0)
                size = ctypes.c_size_t(0)
                want = r.oskar_binary_query(rh, q[0], q[1], q[2], q[3], ctypes.byref(size),
                                            ctypes.byref(st))
                if st.value:
                    with pytest.raises(_lib.GridderError):
                        f.query(*q)
                else:
                    assert f.query(*q) == (want, size.value)
    r.oskar_binary_free(rh)


def test_native_round_trip_without_reference(tmp_path):
    rng = np.random.default_rng(5)
    S, B, T, uu, vv, ww, vis = _case(rng, T=5)
    path = str(tmp_path / "n.vis")
    vis_io.write_vis(path, uu, vv, ww, vis, S, 2, C, precision=_lib.BIN_DOUBLE)
    with vis_io.BinaryFile(path) as f:
        h = f.vis_header()
        assert h.num_blocks == 3
        got = [f.vis_block(h, b) for b in range(3)]
    # freq = c: the scale is exactly 1
    assert np.array_equal(np.concatenate([g[0] for g in got]), uu.ravel())
    assert np.array_equal(np.concatenate([g[3] for g in got]), vis.ravel())


def test_bad_files_refused(tmp_path):
    p = tmp_path / "x.vis"
    p.write_bytes(b"NOTOSKAR" + bytes(100))
    with pytest.raises(_lib.GridderError) as e:
        vis_io.BinaryFile(str(p))
    assert e.value.status == _lib.ERR_FORMAT
    with pytest.raises(_lib.GridderError) as e:
        vis_io.BinaryFile(str(tmp_path / "missing.vis"))
    assert e.value.status == _lib.ERR_IO


# ---------------------------------------------------------------------------
# FITS kernel cubes
# ---------------------------------------------------------------------------
def _ref_read_kernel(r, path, prec, support=True):
    csh, P, ov, gs, st = (ctypes.c_int(0) for _ in range(5))
    cell, ws = ctypes.c_double(0), ctypes.c_double(0)
    sp = ctypes.c_void_p()
    d = r.read_kernel(path.encode(), prec, ctypes.byref(csh), ctypes.byref(P),
                      ctypes.byref(sp) if support else None, ctypes.byref(ov), ctypes.byref(gs),
                      ctypes.byref(cell), ctypes.byref(ws), ctypes.byref(st))
    n = csh.value ** 2 * P.value
    real = np.float64 if prec == _lib.BIN_DOUBLE else np.float32
    data = np.ctypeslib.as_array(ctypes.cast(d, ctypes.POINTER(
        ctypes.c_double if real == np.float64 else ctypes.c_float)), (n,)).copy() if d else None
    sup = np.ctypeslib.as_array(ctypes.cast(sp, ctypes.POINTER(ctypes.c_int)),
                                (P.value,)).copy() if support and sp.value else None
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    libc.free(d)
    if support:
        libc.free(sp)
    info = dict(conv_size_half=csh.value, num_w_planes=P.value, oversample=ov.value,
                grid_size=gs.value, cellsize_rad=cell.value, w_scale=ws.value)
    return st.value, data, sup, info


def _ref_write_kernel(r, path, data, support, info, bitpix, tform=b"1J"):
    st = ctypes.c_int(0)
    f = ctypes.c_void_p()
    r.ffinit(ctypes.byref(f), path.encode(), ctypes.byref(st))
    P, c1, c0 = data.shape
    naxes = (ctypes.c_long * 3)(c0, c1, P)
    r.ffcrim(f, bitpix, 3, naxes, ctypes.byref(st))
    for k in ("OVERSAMP", "GRIDSIZE"):
        v = ctypes.c_int(info[k])
        r.ffpky(f, oracle.TINT, k.encode(), ctypes.byref(v), None, ctypes.byref(st))
    for k in ("CELLSIZE", "W_SCALE"):
        v = ctypes.c_double(info[k])
        r.ffpky(f, oracle.TDOUBLE, k.encode(), ctypes.byref(v), None, ctypes.byref(st))
    first = (ctypes.c_long * 3)(1, 1, 1)
    d = np.ascontiguousarray(data, np.float64)
    r.ffppx(f, oracle.TDOUBLE, first, d.size, _p(d), ctypes.byref(st))
    ttype = (ctypes.c_char_p * 1)(b"SUPPORT")
    tf = (ctypes.c_char_p * 1)(tform)
    tunit = (ctypes.c_char_p * 1)(b"")
    r.ffcrtb(f, oracle.BINARY_TBL, P, 1, ttype, tf, tunit, b"W_KERNELS", ctypes.byref(st))
    s = np.ascontiguousarray(support, np.int32)
    r.ffpcl(f, oracle.TINT, 1, 1, 1, P, _p(s), ctypes.byref(st))
    r.ffclos(f, ctypes.byref(st))
    assert st.value == 0, st.value


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("prec", [_lib.BIN_SINGLE, _lib.BIN_DOUBLE])
def test_native_kernel_read_by_reference(tmp_path, dtype, prec):
    r = _ref()
    rng = np.random.default_rng(1)
    data = rng.normal(size=(5, 12, 12)).astype(dtype)
    data[0, 0, :3] = [np.nan, np.inf, -0.0]
    sup = np.array([3, 4, 5, 6, 7], np.int32)
    path = str(tmp_path / "k.fits")
    vis_io.write_kernel(path, data, sup, 4, 4096, 3.0303030303030303e-06, 0.123456789012345)
    st, d, s, info = _ref_read_kernel(r, path, prec)
    assert st == 0
    nd, ns, ninfo = vis_io.read_kernel(path, prec)
    assert info == ninfo
    assert info["cellsize_rad"] == 3.0303030303030303e-06 and info["w_scale"] == 0.123456789012345
    assert _same(nd.ravel(), d) and np.array_equal(ns, s) and np.array_equal(s, sup)


@pytest.mark.parametrize("bitpix,tform", [(-32, b"1J"), (-64, b"1I"), (16, b"1J")])
@pytest.mark.parametrize("prec", [_lib.BIN_SINGLE, _lib.BIN_DOUBLE])
def test_cfitsio_kernel_read_natively(tmp_path, bitpix, tform, prec):
    r = _ref()
    rng = np.random.default_rng(2)
    data = rng.normal(size=(3, 9, 9)) * (1000 if bitpix == 16 else 1)
    if bitpix == 16:
        data = np.round(data)
    sup = np.array([2, 9, 33], np.int32)
    info = dict(OVERSAMP=4, GRIDSIZE=2048, CELLSIZE=1.1e-5, W_SCALE=2.5e-3)
    path = str(tmp_path / "c.fits")
    _ref_write_kernel(r, path, data, sup, info, bitpix, tform)
    st, d, s, rinfo = _ref_read_kernel(r, path, prec)
    assert st == 0
    nd, ns, ninfo = vis_io.read_kernel(path, prec)
    assert rinfo == ninfo
    assert _same(nd.ravel(), d) and np.array_equal(ns, s)


def test_kernel_round_trip_and_errors(tmp_path):
    data = np.arange(2 * 4 * 4, dtype=np.float32).reshape(2, 4, 4)
    path = str(tmp_path / "k.fits")
    vis_io.write_kernel(path, data, [1, 2], 4, 64, 0.01, 1.0)
    d, s, info = vis_io.read_kernel(path, _lib.BIN_DOUBLE)
    assert np.array_equal(d, data.astype(np.float64)) and list(s) == [1, 2]
    vis_io.write_kernel(path, data, None, 4, 64, 0.01, 1.0)     # no support table
    with pytest.raises(_lib.GridderError) as e:
        vis_io.read_kernel(path)
    assert "W_KERNELS" in str(e.value)
    vis_io.write_image(path, np.zeros((4, 4), np.float32))      # 2-D: not a kernel cube
    with pytest.raises(_lib.GridderError) as e:
        vis_io.read_kernel(path)
    assert "3D" in str(e.value)


def test_kernel_support_column_wider_than_row_refused(tmp_path):
    """A W_KERNELS table whose TFORM1 repeat count times the element width
    exceeds the row width (NAXIS1) is refused before any element is read."""
    data = np.arange(2 * 4 * 4, dtype=np.float32).reshape(2, 4, 4)
    path = str(tmp_path / "k.fits")
    vis_io.write_kernel(path, data, [1, 2], 4, 64, 0.01, 1.0)
    raw = open(path, "rb").read()
    assert raw.count(b"'1J") == 1
    open(path, "wb").write(raw.replace(b"'1J", b"'3K"))   # 24-byte column in 4-byte rows
    with pytest.raises(_lib.GridderError) as e:
        vis_io.read_kernel(path)
    assert e.value.status == _lib.ERR_FORMAT and "does not fit" in str(e.value)
