"""CPU tests of the host stack I/O (motioncorr1exp_amd/libmchost.so) against
the reference's own readers and writer (src-1125/mrc.cpp, mrcz.cpp,
mrct.cpp built into oracle/_ref/libref_io.so by oracle/ref.mk) on the same
files: MRC modes 0/1/2/6 (mode 0 signed), extended headers, gzip MRC and
multi-page 8/16-bit TIFF, plus byte-identical MRC output files."""
import ctypes
import gzip
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# MC_SANITIZED=1: the AddressSanitizer/UBSan build (tests/test_sanitize.py)
HOST = os.path.join(ROOT, "motioncorr1exp_amd", "_san" if os.environ.get("MC_SANITIZED") else "",
                    "libmchost.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libref_io.so")


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HOST):
        pytest.skip("libmchost.so not built")
    L = ctypes.CDLL(HOST)
    vp, i = ctypes.c_void_p, ctypes.c_int
    L.mch_sniff.argtypes = [ctypes.c_char_p]
    L.mch_info.argtypes = [ctypes.c_char_p, vp]
    L.mch_read_float.argtypes = [ctypes.c_char_p, i, vp]
    L.mch_read_float.restype = ctypes.c_longlong
    L.mch_write_float.argtypes = [ctypes.c_char_p, vp, i, i, i, i]
    L.mch_header_stats.argtypes = [ctypes.c_char_p, vp]
    return L


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF):
        pytest.skip("reference I/O not built (oracle/ref.mk needs /root/reference)")
    ctypes.CDLL("/opt/conda/lib/libtiff.so.5", mode=ctypes.RTLD_GLOBAL)
    L = ctypes.CDLL(REF)
    vp, i = ctypes.c_void_p, ctypes.c_int
    L.ref_info.argtypes = [i, ctypes.c_char_p, vp]
    L.ref_read_float.argtypes = [i, ctypes.c_char_p, i, vp]
    L.ref_write_float.argtypes = [ctypes.c_char_p, vp, i, i, i]
    return L


def _mrc_bytes(data, mode, nsymbt=0, stamp=b"MAP "):
    nz, ny, nx = data.shape
    h = np.zeros(256, np.int32)
    h[0:4] = (nx, ny, nz, mode)
    h[23] = nsymbt
    hb = bytearray(h.tobytes())
    hb[208:212] = stamp
    dt = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16}[mode]
    return bytes(hb) + b"\0" * nsymbt + data.astype(dt).tobytes()


def _host_read(L, fn, shape):
    nz, ny, nx = shape
    out = np.zeros(shape, np.float32)
    for n in range(nz):
        r = L.mch_read_float(fn.encode(), n, out[n].ctypes.data)
        assert r > 0
    return out


def _ref_read(L, kind, fn, shape):
    out = np.zeros(shape, np.float32)
    assert L.ref_read_float(kind, fn.encode(), shape[0], out.ctypes.data) == 1
    return out


@pytest.mark.parametrize("mode", [0, 1, 2, 6])
def test_mrc_modes_match_reference(host, ref, tmp_path, mode):
    rng = np.random.default_rng(mode)
    lo, hi = {0: (-128, 128), 1: (-32768, 32768), 2: (-5, 5), 6: (0, 65536)}[mode]
    data = rng.integers(lo, hi, size=(3, 17, 23)) if mode != 2 else rng.uniform(lo, hi, (3, 17, 23))
    fn = str(tmp_path / ("m%d.mrc" % mode))
    open(fn, "wb").write(_mrc_bytes(data, mode, nsymbt=80))
    a = _host_read(host, fn, data.shape)
    b = _ref_read(ref, 1, fn, data.shape)
    assert np.array_equal(a, b)
    dims = np.zeros(4, np.int32)
    host.mch_info(fn.encode(), dims.ctypes.data)
    assert list(dims) == [23, 17, 3, mode]
    if mode == 0:
        assert a.min() < 0  # bytes read as signed (reference mrc.cpp:338-348)


def test_mrc_writer_is_byte_identical(host, ref, tmp_path):
    data = np.random.default_rng(1).standard_normal((2, 31, 29)).astype(np.float32)
    fa, fb = str(tmp_path / "a.mrc"), str(tmp_path / "b.mrc")
    assert host.mch_write_float(fa.encode(), data.ctypes.data, 29, 31, 2, 0)
    assert ref.ref_write_float(fb.encode(), data.ctypes.data, 29, 31, 2)
    assert open(fa, "rb").read() == open(fb, "rb").read()
    mmm = np.zeros(3, np.float32)
    host.mch_header_stats(fa.encode(), mmm.ctypes.data)
    assert mmm[0] == data.min() and mmm[1] == data.max()
    assert abs(mmm[2] - data.astype(np.float64).mean()) < 1e-6


def test_gzip_mrc_matches_reference(host, ref, tmp_path):
    data = np.random.default_rng(2).integers(0, 100, size=(4, 12, 16))
    fn = str(tmp_path / "s.mrc.gz")
    with gzip.open(fn, "wb") as f:
        f.write(_mrc_bytes(data, 1))
    assert host.mch_sniff(fn.encode()) == 2  # gzip sniffed before the name test
    a = _host_read(host, fn, data.shape)
    b = _ref_read(ref, 2, fn, data.shape)
    assert np.array_equal(a, b) and np.array_equal(a, data.astype(np.float32))


def test_gzip_writer_roundtrip(host, tmp_path):
    data = np.random.default_rng(3).standard_normal((2, 8, 8)).astype(np.float32)
    fn = str(tmp_path / "w.mrc.gz")
    assert host.mch_write_float(fn.encode(), data.ctypes.data, 8, 8, 2, 1)
    assert np.array_equal(_host_read(host, fn, data.shape), data)


@pytest.mark.parametrize("bits", [8, 16])
def test_tiff_stack_matches_reference(host, ref, tmp_path, bits):
    from PIL import Image

    rng = np.random.default_rng(bits)
    data = rng.integers(0, 200 if bits == 8 else 60000, size=(3, 20, 24))
    mode = "L" if bits == 8 else "I;16"
    ims = [Image.fromarray(d.astype(np.uint8 if bits == 8 else np.uint16), mode=mode) for d in data]
    fn = str(tmp_path / ("t%d.tif" % bits))
    ims[0].save(fn, save_all=True, append_images=ims[1:])
    assert host.mch_sniff(fn.encode()) == 3
    dims = np.zeros(4, np.int32)
    assert host.mch_info(fn.encode(), dims.ctypes.data)
    assert list(dims) == [24, 20, 3, 7199]
    a = _host_read(host, fn, data.shape)
    b = _ref_read(ref, 3, fn, data.shape)
    assert np.array_equal(a, b) and np.array_equal(a, data.astype(np.float32))


def test_sniffing(host, tmp_path):
    d = np.zeros((1, 4, 4))
    f1 = str(tmp_path / "x.mrc")
    open(f1, "wb").write(_mrc_bytes(d, 2))
    f2 = str(tmp_path / "noname.dat")
    open(f2, "wb").write(_mrc_bytes(d, 2, stamp=b"\0\0\0\0"))
    f3 = str(tmp_path / "old.mrc")
    open(f3, "wb").write(_mrc_bytes(d, 2, stamp=b"\0\0\0\0"))
    assert host.mch_sniff(f1.encode()) == 1
    assert host.mch_sniff(f2.encode()) == 0   # no stamp, no "mrc" in the name
    assert host.mch_sniff(f3.encode()) == 1   # old MRC accepted by name
    assert host.mch_sniff(str(tmp_path / "missing.mrc").encode()) == 0
