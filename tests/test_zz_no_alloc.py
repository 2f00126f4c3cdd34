"""GPU test of mc_find_shift_noalloc, the FindShift entry point that makes no
device allocation (its box copy and peak record live in a fixed device array
of the library).  Called through the C ABI, it must give exactly what
mc_find_shift gives on the same correlation map: shift, cc and box, bit for
bit, on two calls in a row with the second one smaller (the array keeps the
first call's contents), and it must reject a box beyond MC_NOALLOC_MAX_BOX
with an error and still work afterwards."""
import ctypes

import numpy as np
import pytest

import motioncorr1exp_amd as mc

pytestmark = pytest.mark.gpu

MAX_BOX = 256  # MC_NOALLOC_MAX_BOX in include/mc_abi.h


def _map(n, seed):
    """An n x (n+2) padded C2R-layout map: a Gaussian peak at a random
    sub-pixel offset from the centre over noise."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:n, 0:n]
    x0, y0 = n / 2 + rng.uniform(-9, 9), n / 2 + rng.uniform(-9, 9)
    img = 50.0 * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / 6.0) + rng.normal(0, 1.0, (n, n))
    buf = np.zeros((n, n + 2), np.float32)
    buf[:, :n] = img
    return buf


def _call(fn, dptr, n, box, wnoise):
    hbox = np.full((box, box), np.nan, np.float32)
    sx, sy, cc = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    mc.check(fn(dptr, n, hbox.ctypes.data, box, ctypes.byref(sx), ctypes.byref(sy),
                ctypes.byref(cc), wnoise))
    return sx.value, sy.value, cc.value, hbox


@pytest.mark.parametrize("cases", [[(512, 96, -1, 1), (256, 64, -1, 2)],
                                   [(1024, 256, 2, 3), (384, 48, 0, 4)]])
def test_noalloc_matches_find_shift(cases):
    L = mc.lib()
    for n, box, wnoise, seed in cases:
        d = mc.DeviceArray(_map(n, seed))
        ref = _call(L.mc_find_shift, d.ptr, n, box, wnoise)
        got = _call(L.mc_find_shift_noalloc, d.ptr, n, box, wnoise)
        print("n %d box %d wnoise %d: shift (%g, %g) cc %g" % (n, box, wnoise, *ref[:3]), flush=True)
        assert got[:3] == ref[:3]
        assert np.array_equal(got[3], ref[3])
        d.free()


def test_noalloc_rejects_large_box_then_works():
    L = mc.lib()
    n = 1024
    d = mc.DeviceArray(_map(n, 5))
    hbox = np.zeros((MAX_BOX + 2, MAX_BOX + 2), np.float32)
    sx, sy, cc = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    rc = L.mc_find_shift_noalloc(d.ptr, n, hbox.ctypes.data, MAX_BOX + 2, ctypes.byref(sx),
                                 ctypes.byref(sy), ctypes.byref(cc), -1)
    assert rc == -1
    assert b"MC_NOALLOC_MAX_BOX" in L.mc_last_error()
    ref = _call(L.mc_find_shift, d.ptr, n, 128, -1)
    got = _call(L.mc_find_shift_noalloc, d.ptr, n, 128, -1)
    assert got[:3] == ref[:3] and np.array_equal(got[3], ref[3])
    d.free()
