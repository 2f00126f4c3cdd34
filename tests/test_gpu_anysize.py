"""GPU parity at transform sizes without a Stockham plan (chirp-z path,
mc_czt.hip; the forward stage-1 transform at the 3710 and 4092 canvases in
two levels, 35 x 53 / 70 x 53 and 66 x 31 / 132 x 31): the reference crops to min(nx, ny) rounded to even and plans
cuFFT at that size (func.cpp:80-103, DFAlign.cu:508-515, cufunc.cu:202-215),
so a 3838 x 3710 K2 frame runs at n = 3710 = 2*5*7*53, a K3 frame binned 2x
at 4092 -> 2046 = 2*3*11*31, and a non-integer -bin gives sizes such as 640.

- level 0: GPUFFT2d / GPUIFFT2d of these sizes against numpy (FP64);
- whole movies through mc_align_movie against the FP64 oracle on the same
  frames: spectra within 1e-6 of their maximum (the chirp-z transform in
  FP32 is within about 3e-7 on random data), box maps within 1e-5 of each
  map's maximum, pair and frame shifts within 0.01 px (components beyond
  counted, printed and required to be 0), aligned sum within 1e-4 relative;
- the drop-in CLI at its default options (no -pad, quick-look images on) on
  a stack whose square crop has no Stockham plan."""
import os
import re
import subprocess

import numpy as np
import pytest

import motioncorr1exp_amd as mc
from motioncorr1exp_amd.synth import make_movie
from _oracle import Oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "motioncorr1exp_amd", "bin", "dosefgpu_driftcorr")


@pytest.mark.parametrize("n", [66, 640, 1000, 1854, 2046, 3710, 4092])
def test_level0_fft_any_even_size(n):
    rng = np.random.default_rng(n)
    img = rng.standard_normal((n, n)).astype(np.float32)
    buf = np.zeros((n, n + 2), np.float32)
    buf[:, :n] = img
    d = mc.DeviceArray(buf)
    L = mc.lib()
    plan, iplan = mc.ctypes.c_void_p(), mc.ctypes.c_void_p()
    mc.check(L.mc_fft_plan(mc.ctypes.byref(plan), n))
    mc.check(L.mc_ifft_plan(mc.ctypes.byref(iplan), n))
    mc.check(L.mc_fft2d(d.ptr, plan))
    mc.check(L.mc_sync())
    spec = d.to_host(np.float32, (n, n + 2)).view(np.complex64)
    ref = np.fft.rfft2(img.astype(np.float64))
    err = np.abs(spec - ref).max() / np.abs(ref).max()
    mc.check(L.mc_ifft2d(d.ptr, iplan))
    mc.check(L.mc_sync())
    back = d.to_host(np.float32, (n, n + 2))[:, :n] / (n * n)
    berr = np.abs(back - img).max() / np.abs(img).max()
    print("n %d: R2C error / max %.3g, round trip error / max %.3g" % (n, err, berr), flush=True)
    assert err <= 1e-6 and berr <= 2e-6
    L.mc_fft_destroy(plan)
    L.mc_fft_destroy(iplan)
    d.free()


@pytest.mark.parametrize("n,box", [(3710, 32), (3710, 64), (3710, 96), (3710, 128), (3710, 160),
                                   (4092, 64), (4092, 96), (4092, 128)])
def test_box_maps_at_3710_match_oracle(n, box):
    """The pair column pass at the default crops of K2 (n = 3710) and K3 (n =
    4092) frames runs decimated for boxes up to 128: at 3710, 53 inverse
    70-point FFTs (radix 2, 5, 7) and a 53-term combine split 27 + 26 over
    two lanes per row; at 4092, 31 inverse 132-point FFTs (radix 4, 3, 11)
    and a 31-term combine.  Larger boxes run the chirp-z pass.  Box maps of
    random spectra against the FP64 oracle's cross power and pruned inverse
    (cufunc.cu:343-370, DFAlign.cu:897, FindShift's read-back
    cufunc.cu:396-424)."""
    F = 4
    rng = np.random.default_rng(n + box)
    spec = np.fft.rfft2(rng.standard_normal((F, n, n))).astype(np.complex64)
    ctx = mc.Context(0)
    ctx.set_spectra(spec.view(np.float32))
    pairs = np.array([(0, 1), (0, 3), (1, 2), (2, 3)], np.int32)
    _, _, boxes = ctx.pair_shifts(pairs, 150.0, box, want_boxes=True)
    ctx.close()
    o = Oracle("f64")
    worst = 0.0
    for q, (a, b) in enumerate(pairs):
        ref = o.pruned_box(o.cross_power(spec[a], spec[b], n, 150.0), box)
        worst = max(worst, float(np.abs(boxes[q] - ref).max() / np.abs(ref).max()))
    print("n %d box %d: max error / max %.3g" % (n, box, worst), flush=True)
    assert worst < 1e-5


def _off(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b))
    d = d[np.isfinite(d)]
    return int(np.sum(d > 0.01)), float(d.max())


# (name, nx, ny, frames, oracle/GPU geometry, expected n, seed)
CASES = [
    ("k2_3710", 3838, 3710, 8, dict(fod=1, iter=1), 3710, 501),
    ("k3_bin2_2046", 4200, 4092, 6, dict(fod=1, iter=1, bin=2.0), 2046, 502),
    ("bin1.5_640", 1000, 960, 8, dict(fod=2, iter=2, bin=1.5), 640, 503),
    ("crop1854", 1900, 1854, 8, dict(fod=2, iter=1), 1854, 504),
    # canvas 4200 binned 2x: packed rows of 2100 through the 8192-point
    # persistent chirp-z row kernel, columns through the 16384-point one
    ("crop4200_bin2_2100", 4300, 4200, 4, dict(fod=1, iter=1, bin=2.0), 2100, 505),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_any_size_movie_matches_oracle(case):
    name, nx, ny, F, geo, n, seed = case
    frames, drift = make_movie(nx, ny, F, dose=1.5, step=0.8, seed=seed)
    print("movie %s ready" % name, flush=True)
    binf = geo.get("bin", 1.0)
    gp = mc.default_params(F, FrameDistOffset=geo["fod"], iter=geo["iter"], bin=binf,
                           CCPeakSearchDim=96)
    ctx = mc.Context(0)
    d = mc.DeviceArray(frames)
    ctx.load_frames(d, frames.shape, frames.dtype, gp)
    spec = ctx.get_spectra(F, n).view(np.complex64)
    res = ctx.align_movie(d, frames.shape, frames.dtype, gp,
                          want=("boxmaps", "pair_b", "frame_shift", "sum_image"))
    d.free()
    ctx.close()
    assert res["report"]["n"] == n
    o = Oracle("f64").align_movie(frames, dict(bin=binf, fod=geo["fod"], iter=geo["iter"],
                                               box=96, full_c2r=0, nthreads=16),
                                  want_spectra=True)
    assert o["n"] == n
    ref = o["spectra"].view(np.complex128)
    serr = np.abs(spec - ref).max() / np.abs(ref).max()
    bm, obm = res["boxmaps"], o["boxmaps"]
    scale = np.abs(obm).reshape(len(obm), -1).max(1)[:, None, None]
    berr = float((np.abs(bm - obm) / scale).max())
    kp, dp = _off(res["pair_b"], o["pair_b"])
    kf, df = _off(res["frame_shift"], o["frame_shift"])
    rel = np.linalg.norm(res["sum_image"] - o["sum_image"]) / np.linalg.norm(o["sum_image"])
    print("%s (n %d): spectra error / max %.3g, box maps %.3g, pair shifts beyond 0.01 px %d "
          "(max %.3g), frame shifts beyond %d (max %.3g), sum rel L2 %.3g"
          % (name, n, serr, berr, kp, dp, kf, df, rel), flush=True)
    assert serr < 1e-6 and berr < 1e-5
    assert kp == 0 and kf == 0
    assert rel < 1e-4


def _mrc_bytes(data):
    nz, ny, nx = data.shape
    h = np.zeros(256, np.int32)
    h[0:4] = (nx, ny, nz, 2)
    hb = bytearray(h.tobytes())
    hb[208:212] = b"MAP "
    return bytes(hb) + data.astype(np.float32).tobytes()


def test_cli_default_options_non_smooth_crop(tmp_path):
    """dosefgpu_driftcorr with only the stack name: square crop 690 =
    2*3*5*23, quick-look images on (-dsp 1).  Final shifts from the log and
    the aligned sum against the FP64 oracle."""
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    frames, _ = make_movie(700, 690, 10, dose=1.5, step=0.7, seed=690)
    open(tmp_path / "s.mrc", "wb").write(_mrc_bytes(frames))
    r = subprocess.run([CLI, "s.mrc"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    log = open(tmp_path / "s_Log.txt").read()
    shifts = np.array([[float(a), float(b)] for a, b in
                       re.findall(r"\.\.\.\.\.\.Shift of Frame #\d+ : +(\S+) +(\S+)", log)])
    o = Oracle("f64").align_movie(frames, dict(nthreads=16), want_boxes=False)
    ref = np.vstack([[0, 0], np.cumsum(o["final_shift"], axis=0)])
    err = np.abs(shifts - ref)
    print("CLI at n 690: frame shifts max difference %.3g px" % err.max(), flush=True)
    assert err.max() <= 0.01 + 5e-5
    raw = open(tmp_path / "s_SumCorr.mrc", "rb").read()
    s = np.frombuffer(raw[1024:], np.float32).reshape(690, 690)
    rel = np.linalg.norm(s - o["sum_image"]) / np.linalg.norm(o["sum_image"])
    assert rel < 1e-4, rel
    for suffix in ("_RawFFT", "_CorrSum", "_CorrFFT"):
        assert any(suffix in f for f in os.listdir(tmp_path / "dosef_quick")), suffix
