"""GPU tests of the C++ drop-in: dosefgpu_driftcorr on synthetic stacks in
MRC (mode 0 bytes and float), gzip-MRC and TIFF, compared with the FP64
oracle: final per-frame shifts parsed from the log ("Shift of Frame" lines,
reference DFAlign.cu:1086), the corrected sum and the unaligned sum files."""
import gzip
import os
import re
import subprocess

import numpy as np
import pytest

from _oracle import Oracle
from motioncorr1exp_amd.synth import make_movie

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "motioncorr1exp_amd", "bin", "dosefgpu_driftcorr")


def _mrc_bytes(data, mode):
    nz, ny, nx = data.shape
    h = np.zeros(256, np.int32)
    h[0:4] = (nx, ny, nz, mode)
    hb = bytearray(h.tobytes())
    hb[208:212] = b"MAP "
    dt = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16}[mode]
    return bytes(hb) + data.astype(dt).tobytes()


def _read_mrc(fn):
    raw = open(fn, "rb").read()
    h = np.frombuffer(raw[:1024], np.int32)
    nx, ny, nz, mode = h[:4]
    assert mode == 2
    return np.frombuffer(raw[1024:], np.float32).reshape(nz, ny, nx)


def _run(tmp_path, fname, args):
    r = subprocess.run([CLI, fname] + args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def _final_shifts(log):
    return np.array([[float(a), float(b)] for a, b in
                     re.findall(r"\.\.\.\.\.\.Shift of Frame #\d+ : +(\S+) +(\S+)", log)])


@pytest.fixture(scope="module")
def movie():
    frames, drift = make_movie(544, 512, 10, dose=1.0, step=0.7, seed=1001)
    return frames


def test_cli_mrc_float(tmp_path, movie):
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    open(tmp_path / "stack.mrc", "wb").write(_mrc_bytes(movie, 2))
    out = _run(tmp_path, "stack.mrc", ["-fod", "4", "-bft", "150", "-pbx", "96", "-dsp", "0",
                                      "-srs", "1", "-itr", "2"])
    log = open(tmp_path / "stack_Log.txt").read()
    assert "******Parameter List******" in log and "Done." in log
    assert "Calculate relative drift between frames" in out
    o = Oracle("f64").align_movie(movie, dict(fod=4, iter=2, nthreads=8), want_boxes=False)
    shifts = _final_shifts(log)
    ref = np.vstack([[0, 0], np.cumsum(o["final_shift"], axis=0)])
    assert shifts.shape == ref.shape
    # the log prints shifts with %8.4f: 0.01 px plus half a unit of the last digit
    err = np.abs(shifts - ref)
    print("CLI frame shifts vs oracle: max %.3g px, > 0.01 px: %d" % (err.max(), int((err > 0.01).sum())))
    assert err.max() <= 0.01 + 5e-5
    s = _read_mrc(tmp_path / "stack_SumCorr.mrc")[0]
    assert s.shape == (512, 512)
    rel = np.linalg.norm(s - o["sum_image"]) / np.linalg.norm(o["sum_image"])
    assert rel < 1e-4, rel
    raw = _read_mrc(tmp_path / "stack_SumRaw.mrc")[0]
    rel = np.linalg.norm(raw - o["rawsum_image"]) / np.linalg.norm(o["rawsum_image"])
    assert rel < 1e-5, rel


def test_cli_mode0_bytes_and_gzip(tmp_path, movie):
    """int8 counts decoded on the device; the same stack gzip-compressed."""
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    q = np.clip(movie, 0, 127).astype(np.int8)
    open(tmp_path / "b.mrc", "wb").write(_mrc_bytes(q, 0))
    with gzip.open(tmp_path / "g.mrc.gz", "wb") as f:
        f.write(_mrc_bytes(q, 0))
    _run(tmp_path, "b.mrc", ["-fod", "4", "-dsp", "0", "-atm", "0"])
    _run(tmp_path, "g.mrc.gz", ["-fod", "4", "-dsp", "0", "-atm", "0"])
    a = _read_mrc(tmp_path / "b_SumCorr.mrc")
    b = _read_mrc(tmp_path / "g.mrc_SumCorr.mrc")
    assert np.array_equal(a, b)
    o = Oracle("f64").align_movie(q.astype(np.float32), dict(fod=4, atm=0, nthreads=8), want_boxes=False)
    rel = np.linalg.norm(a[0] - o["sum_image"]) / np.linalg.norm(o["sum_image"])
    assert rel < 1e-4, rel


def test_cli_tiff_bin2_and_outputs(tmp_path, movie):
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    from PIL import Image

    big, _ = make_movie(512, 512, 8, dose=2.0, step=1.0, seed=5)
    ims = [Image.fromarray(np.clip(f, 0, 255).astype(np.uint8), mode="L") for f in big]
    ims[0].save(tmp_path / "t.tif", save_all=True, append_images=ims[1:])
    out = _run(tmp_path, "t.tif", ["-bin", "2", "-fod", "3", "-dsp", "0", "-ssc", "1", "-scc", "1",
                                   "-pbx", "64", "-nes", "6", "-2es", "4"])
    assert "Auto-detected file format: TIFF" in out
    s = _read_mrc(tmp_path / "t_binned_SumCorr.mrc")
    assert s.shape == (1, 256, 256)
    s2 = _read_mrc(tmp_path / "t_binned_SumCorr2.mrc")
    st = _read_mrc(tmp_path / "t_binned_Corrected.mrc")
    assert st.shape == (7, 256, 256)
    # the corrected stack frames add up to the sum; the first frames to the second sum
    assert np.abs(st.sum(0) - s[0]).max() < 1e-3 * np.abs(s).max()
    assert np.abs(st[:5].sum(0) - s2[0]).max() < 1e-3 * np.abs(s2).max()
    cc = _read_mrc(tmp_path / "t_binned_CCMap.mrc")
    assert cc.shape[1:] == (64, 64)
    frames = np.stack([np.asarray(im, np.float32) for im in ims])
    o = Oracle("f64").align_movie(frames, dict(bin=2.0, fod=3, box=64, nEndSum=6, nEndSum2=4,
                                               nthreads=8), want_boxes=True)
    assert cc.shape[0] == len(o["pairs"])
    rel = np.linalg.norm(s[0] - o["sum_image"]) / np.linalg.norm(o["sum_image"])
    assert rel < 1e-4, rel


def test_cli_errors(tmp_path):
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    r = subprocess.run([CLI, "missing.mrc"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1 and "Not a stack or wrong file" in r.stdout
    d = np.zeros((4, 64, 64), np.float32)
    open(tmp_path / "s.mrc", "wb").write(_mrc_bytes(d, 2))
    r = subprocess.run([CLI, "s.mrc", "-fod"], cwd=tmp_path, capture_output=True, text=True)
    assert "Wrong number of arguments" in r.stdout
    r = subprocess.run([CLI, "s.mrc", "-xyz", "1"], cwd=tmp_path, capture_output=True, text=True)
    assert "Undefined option: -xyz" in r.stdout
    r = subprocess.run([CLI, "s.mrc", "-fod", "5"], cwd=tmp_path, capture_output=True, text=True)
    assert "Frame distance offset is too big" in r.stdout
    r = subprocess.run([CLI, "s.mrc", "-scx", "0.5"], cwd=tmp_path, capture_output=True, text=True)
    assert "Stretching factor along X-axis needs to be >=1.0." in r.stdout


def test_cli_new_form_flags(tmp_path):
    """-pad, -prc 1 (FP16 hi/lo pair stage) and -grd/-grs (sharpness
    refinement) reach the library: the log's parameter block lists them and
    the final shifts equal those of mc_align_movie with the same parameters
    (to the printed precision)."""
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    import motioncorr1exp_amd as mc
    frames, _ = make_movie(1000, 990, 8, dose=2.0, step=0.7, seed=4040)
    open(tmp_path / "f.mrc", "wb").write(_mrc_bytes(frames, 2))
    _run(tmp_path, "f.mrc", ["-pad", "1024", "-prc", "1", "-grd", "5", "-grs", "0.04",
                             "-dsp", "0", "-fod", "1"])
    log = open(tmp_path / "f_Log.txt").read()
    assert "  -pad  1024\n" in log and "  -prc  1\n" in log
    assert "  -grd  5\n  -grs  0.040000\n" in log
    shifts = _final_shifts(log)
    ctx = mc.Context(0)
    d = mc.DeviceArray(frames)
    p = mc.default_params(8, pad=1024, precision=1, grad_steps=5, grad_step_px=0.04,
                          FrameDistOffset=1)
    r = ctx.align_movie(d, frames.shape, frames.dtype, p, want=("final_shift",))
    ctx.close()
    ref = np.vstack([[0, 0], np.cumsum(r["final_shift"], axis=0)])
    err = np.abs(shifts - ref).max()
    print("CLI -prc 1 -grd 5 vs library: max difference %.3g px" % err)
    assert err <= 5e-5 + 1e-9
    r = subprocess.run([CLI, "f.mrc", "-prc", "3"], cwd=tmp_path, capture_output=True, text=True)
    assert "Precision form (-prc) must be 0 or 1." in r.stdout

