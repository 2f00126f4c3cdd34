"""GPU tests of the dark/gain/stretch corrections applied on the device in
the load of the row FFT (mc_corrections; reference DFAlign.cu:614-766, done
there on the host).  The same corrections computed with numpy in float32
with the reference's operation order, and uploaded as corrected float
frames (frames2 for the second section), must give bit-identical spectra,
pair shifts and sums.  The CLI with -fdk/-fnm/-scx/-scy is checked against
the same host-corrected alignment."""
import os
import subprocess

import numpy as np
import pytest

import motioncorr1exp_amd as mc
from motioncorr1exp_amd.synth import make_movie

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "motioncorr1exp_amd", "bin", "dosefgpu_driftcorr")


def host_correct(raw, dark, gain, dark2, gain2, xs, ys):
    """DFAlign.cu:718-766 in float32: section-1 and section-2 corrected
    frames, then the nearest-neighbour stretch (rows/columns >= 1 take
    int(k * scale + 0.5) with the product in float32, the rounding addition
    in float64)."""
    f = raw.astype(np.float32)

    def corr(d, g):
        if d is not None and g is not None:
            return (f - d) * g
        if g is not None:
            return f * g
        if d is not None:
            return f - d
        return f.copy()

    b1, b2 = corr(dark, gain), corr(dark2, gain2)
    if xs < 1 or ys < 1:
        nz, ny, nx = raw.shape
        sx = (np.arange(nx, dtype=np.float32) * np.float32(xs)).astype(np.float64) + 0.5
        sy = (np.arange(ny, dtype=np.float32) * np.float32(ys)).astype(np.float64) + 0.5
        ix, iy = sx.astype(np.int64), sy.astype(np.int64)
        ix[0], iy[0] = 0, 0
        for b in (b1, b2):
            src = b[:, iy][:, :, ix]
            src[:, 0, :] = b[:, 0, :]
            src[:, :, 0] = b[:, :, 0]
            b[...] = src
    return b1.astype(np.float32), b2.astype(np.float32)


@pytest.fixture(scope="module")
def ctx():
    c = mc.Context(0)
    yield c
    c.close()


def _refs(ny, nx, seed, two):
    rng = np.random.default_rng(seed)
    dark = rng.normal(1.5, 0.3, (ny, nx)).astype(np.float32)
    gain = rng.uniform(0.85, 1.15, (ny, nx)).astype(np.float32)
    dark2 = rng.normal(1.2, 0.3, (ny, nx)).astype(np.float32) if two else None
    gain2 = rng.uniform(0.9, 1.1, (ny, nx)).astype(np.float32) if two else None
    return dark, gain, dark2, gain2


@pytest.mark.parametrize("dtype,two,stretch", [(np.uint16, True, True), (np.int8, False, False),
                                               (np.float32, True, False), (np.uint16, False, True)])
def test_device_corrections_match_host(ctx, dtype, two, stretch):
    F = 6
    frames, _ = make_movie(544, 520, F, dose=3.0, step=0.8, seed=71)
    raw = np.clip(frames, 0, 127).astype(dtype)
    nz, ny, nx = raw.shape
    dark, gain, dark2, gain2 = _refs(ny, nx, 5, two)
    xs, ys = (1 / 1.03, 1 / 1.015) if stretch else (1.0, 1.0)
    b1, b2 = host_correct(raw, dark, gain, dark2, gain2, xs, ys)
    prm = mc.default_params(F, FrameDistOffset=1, CCPeakSearchDim=64, crop_nsam=512)
    want = ("pair_b", "frame_shift", "sum_image", "boxmaps")
    h = ctx.align_movie(mc.DeviceArray(b1), b1.shape, np.float32, prm,
                        dev_frames2=mc.DeviceArray(b2), want=want)
    refs = [mc.DeviceArray(a) if a is not None else None for a in (dark, gain, dark2, gain2)]
    corr = mc.corrections(*refs, xscale=xs, yscale=ys, second=True)
    d = ctx.align_movie(mc.DeviceArray(raw), raw.shape, dtype, prm, want=want, corr=corr)
    assert np.array_equal(d["boxmaps"], h["boxmaps"])
    assert np.array_equal(d["pair_b"], h["pair_b"])
    assert np.array_equal(d["frame_shift"], h["frame_shift"])
    assert np.array_equal(d["sum_image"], h["sum_image"])
    # section-1 spectra through the stage entry point as well
    ctx.load_frames(mc.DeviceArray(raw), raw.shape, dtype, prm, corr=corr)
    s_dev = ctx.get_spectra(F, 512)
    ctx.load_frames(mc.DeviceArray(b1), b1.shape, np.float32, prm)
    assert np.array_equal(s_dev, ctx.get_spectra(F, 512))


@pytest.mark.parametrize("nx,ny,crop", [(2500, 2480, 2470), (3730, 3720, 3710)])
def test_device_corrections_match_host_chirp_z(ctx, nx, ny, crop):
    """The same at square crops without a Stockham plan.  At 2470 = 2 5 13 19
    stage 1 runs the persistent chirp-z row kernel (packed length 1235,
    4096-point convolution) and column kernel (8192-point); at 3710 = 70 x 53
    the two-level row (35 x 53) and column (70 x 53) kernels.  The
    corrections are applied in their row loads."""
    F = 3
    frames, _ = make_movie(nx, ny, F, dose=2.0, step=0.8, seed=72)
    raw = np.clip(frames, 0, 127).astype(np.uint16)
    nz, ny, nx = raw.shape
    dark, gain, dark2, gain2 = _refs(ny, nx, 6, True)
    xs, ys = 1 / 1.02, 1 / 1.01
    b1, b2 = host_correct(raw, dark, gain, dark2, gain2, xs, ys)
    prm = mc.default_params(F, FrameDistOffset=1, CCPeakSearchDim=64, crop_nsam=crop)
    want = ("pair_b", "frame_shift", "sum_image")
    h = ctx.align_movie(mc.DeviceArray(b1), b1.shape, np.float32, prm,
                        dev_frames2=mc.DeviceArray(b2), want=want)
    refs = [mc.DeviceArray(a) for a in (dark, gain, dark2, gain2)]
    corr = mc.corrections(*refs, xscale=xs, yscale=ys, second=True)
    d = ctx.align_movie(mc.DeviceArray(raw), raw.shape, np.uint16, prm, want=want, corr=corr)
    assert np.array_equal(d["pair_b"], h["pair_b"])
    assert np.array_equal(d["frame_shift"], h["frame_shift"])
    assert np.array_equal(d["sum_image"], h["sum_image"])
    ctx.load_frames(mc.DeviceArray(raw), raw.shape, np.uint16, prm, corr=corr)
    s_dev = ctx.get_spectra(F, crop)
    ctx.load_frames(mc.DeviceArray(b1), b1.shape, np.float32, prm)
    assert np.array_equal(s_dev, ctx.get_spectra(F, crop))


def test_corrections_reject_bad_arguments(ctx):
    frames, _ = make_movie(256, 256, 4, dose=2.0, step=0.5, seed=9)
    d = mc.DeviceArray(frames)
    prm = mc.default_params(4, CCPeakSearchDim=64)
    with pytest.raises(mc.MCError, match="stretch factors"):
        ctx.align_movie(d, frames.shape, frames.dtype, prm, corr=mc.corrections(xscale=1.5))
    with pytest.raises(mc.MCError, match="exclusive"):
        ctx.align_movie(d, frames.shape, frames.dtype, prm, dev_frames2=d,
                        corr=mc.corrections(second=True))


def _mrc(path, data, mode):
    nz, ny, nx = data.shape
    h = np.zeros(256, np.int32)
    h[0:4] = (nx, ny, nz, mode)
    hb = bytearray(h.tobytes())
    hb[208:212] = b"MAP "
    dt = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16}[mode]
    open(path, "wb").write(bytes(hb) + data.astype(dt).tobytes())


def _read_mrc(fn):
    raw = open(fn, "rb").read()
    h = np.frombuffer(raw[:1024], np.int32)
    nx, ny, nz = h[:3]
    return np.frombuffer(raw[1024:], np.float32).reshape(nz, ny, nx)


def test_cli_dark_gain_stretch(tmp_path, ctx):
    """-fdk (two sections), -fnm (one section), -scx/-scy: the CLI uploads the
    raw mode-1 stack and corrects on the device; its sum equals the
    host-corrected alignment bit for bit."""
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    F = 8
    frames, _ = make_movie(544, 520, F, dose=3.0, step=0.8, seed=72)
    raw = np.clip(frames, 0, 300).astype(np.int16)
    nz, ny, nx = raw.shape
    rng = np.random.default_rng(8)
    dark = rng.normal(1.5, 0.3, (2, ny, nx)).astype(np.float32)
    gain = rng.uniform(0.85, 1.15, (1, ny, nx)).astype(np.float32)
    _mrc(tmp_path / "stack.mrc", raw, 1)
    _mrc(tmp_path / "dark.mrc", dark, 2)
    _mrc(tmp_path / "gain.mrc", gain, 2)
    r = subprocess.run([CLI, "stack.mrc", "-fdk", "dark.mrc", "-fnm", "gain.mrc", "-scx", "1.02",
                        "-scy", "1.01", "-fod", "2", "-pbx", "64", "-dsp", "0", "-crd", "512"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    log = open(tmp_path / "stack_Log.txt").read()
    assert "Correcting dark/gain..." in log
    # reference semantics: gain2 = 1 (single section), dark2 = second section
    # DFAlign.cu:503-504: float xscale = 1.0 / para.xscale (double quotient)
    xs = np.float32(1.0 / np.float64(np.float32(1.02)))
    ys = np.float32(1.0 / np.float64(np.float32(1.01)))
    b1, b2 = host_correct(raw, dark[0], gain[0], dark[1], np.ones((ny, nx), np.float32),
                          float(xs), float(ys))
    prm = mc.default_params(F, FrameDistOffset=2, CCPeakSearchDim=64, crop_nsam=512)
    h = ctx.align_movie(mc.DeviceArray(b1), b1.shape, np.float32, prm,
                        dev_frames2=mc.DeviceArray(b2), want=("sum_image",))
    s = _read_mrc(tmp_path / "stack_SumCorr.mrc")[0]
    assert np.array_equal(s, h["sum_image"])
