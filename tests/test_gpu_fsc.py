"""GPU tests of the FSC table (-fsc; reference DFAlign.cu:212-326 PlotFSC and
its inputs DFAlign.cu:808-822, 1143-1266, 1314-1318): the shell kernel
through mc_fsc_table against the FP64 oracle on the same float32 half
spectra, one-element spectra whose shell and side are fixed by the float
arithmetic of the reference, the per-movie table of mc_align_movie against
the oracle's whole-movie table, batch against single, and the table printed
by the command-line driver."""
import os
import re
import subprocess

import numpy as np
import pytest

import motioncorr1exp_amd as mc
from _oracle import Oracle
from motioncorr1exp_amd.synth import make_movie

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "motioncorr1exp_amd", "bin", "dosefgpu_driftcorr")


@pytest.fixture(scope="module")
def ctx():
    c = mc.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def o():
    return Oracle("f64")


def _half(n, seed):
    rng = np.random.default_rng(seed)
    h = (n, n // 2 + 1)
    return (rng.standard_normal(h) + 1j * rng.standard_normal(h)).astype(np.complex64)


def _dev(spec):
    return mc.DeviceArray(np.ascontiguousarray(spec, np.complex64))


@pytest.mark.parametrize("n", [768, 1024, 4096])
@pytest.mark.parametrize("d", [(0.3, -1.7), (1.0, 0.0), (0.02, 0.05)])
def test_fsc_table_matches_oracle(ctx, o, n, d):
    """Same float32 inputs: only the FP64 summation order differs."""
    a = _half(n, 7 + n)
    b = (0.6 * a + 0.4 * _half(n, 8 + n)).astype(np.complex64)
    da, db = _dev(a), _dev(b)
    got = ctx.fsc_table(da, db, n, d)
    ref = o.fsc(a.astype(np.complex128), b.astype(np.complex128), d)
    print("n", n, "d", d, "max |diff|", np.abs(got - ref).max(), flush=True)
    assert got.shape == ref.shape
    assert np.array_equal(got == 0, ref == 0)
    assert np.abs(got - ref).max() < 1e-12
    da.free()
    db.free()


@pytest.mark.parametrize("kx,y,d,shell,col", [
    (100, -30, (1.0, 0.0), 52, 0),
    (100, -30, (0.0, 5.0), 52, 1),
    (100, -30, (0.05, 0.0), 52, 0),
    (100, 30, (-3.0, 4.0), 52, 1),
    (0, 1, (1.0, 0.0), 0, 1),
    (512, 0, (1.0, 0.0), 256, 0),
    (0, -512, (0.0, 1.0), 256, 0),
    (3, 3, (1.0, 0.0), 2, 0),
    (7, 7, (1.0, 0.0), 4, 1),
    (0, 0, (1.0, 0.0), -1, -1),
    (400, 400, (1.0, 0.0), -1, -1),
])
def test_fsc_single_element(ctx, kx, y, d, shell, col):
    """One nonzero element: shell int(sqrtf(kx^2+y^2)/2) at n = 1024 and the
    side from the float angle test (DFAlign.cu:243-249); DC and |k| >= 513
    are skipped."""
    n = 1024
    a = np.zeros((n, n // 2 + 1), np.complex64)
    a[y % n, kx] = 3.0 - 2.0j
    da = _dev(a)
    got = ctx.fsc_table(da, da, n, d)
    expect = np.zeros_like(got)
    if shell >= 0:
        expect[shell, col] = 1.0
    assert np.array_equal(got, expect)
    da.free()


def _movie_params(F, **kw):
    return mc.default_params(F, **kw)


@pytest.mark.parametrize("case", [
    dict(F=8, fod=2, atm=1, nss=0, nes=0, itr=0),
    dict(F=9, fod=1, atm=0, nss=1, nes=7, itr=2),
])
def test_movie_fsc_matches_oracle(ctx, o, case):
    F = case["F"]
    frames, _ = make_movie(520, 512, F, dose=2.0, step=0.9, seed=2024 + F)
    p = _movie_params(F, FrameDistOffset=case["fod"], bAlignToMid=case["atm"],
                      nStartSum=case["nss"], nEndSum=case["nes"], iter=case["itr"],
                      CCPeakSearchDim=64)
    dev = mc.DeviceArray(frames)
    r = ctx.align_movie(dev, frames.shape, frames.dtype, p, want=("fsc", "frame_shift"))
    dev.free()
    ro = o.align_movie(frames, dict(fod=case["fod"], atm=case["atm"], nStartSum=p["nStartSum"],
                                    nEndSum=p["nEndSum"], iter=case["itr"], box=64, nthreads=8),
                       want_boxes=False)
    fs_err = np.nanmax(np.abs(r["frame_shift"] - ro["frame_shift"]))
    print(case, "frame shift diff", fs_err, "fsc diff", np.abs(r["fsc"] - ro["fsc"]).max(),
          flush=True)
    assert fs_err < 1e-6
    assert r["fsc"].shape == ro["fsc"].shape == (o.fsc_bins(512)[0], 4)
    assert np.array_equal(r["fsc"] == 0, ro["fsc"] == 0)
    # FP32 stage-1 spectra against FP64 ones
    assert np.abs(r["fsc"] - ro["fsc"]).max() < 2e-5


def test_batch_fsc_equals_single(ctx):
    movies, singles = [], []
    for k in range(3):
        frames, _ = make_movie(512, 512, 8, dose=1.5, step=0.8, seed=300 + k)
        p = _movie_params(8, FrameDistOffset=2, CCPeakSearchDim=64, iter=1 + k % 2)
        dev = mc.DeviceArray(frames)
        singles.append(ctx.align_movie(dev, frames.shape, frames.dtype, p, want=("fsc",))["fsc"])
        movies.append(dict(dev=dev, shape=frames.shape, dtype=frames.dtype, params=p,
                           want=("fsc",)))
    res, _ = ctx.align_batch(movies)
    for k in range(3):
        assert np.array_equal(res[k]["fsc"], singles[k])
        movies[k]["dev"].free()


def _mrc_float(data):
    nz, ny, nx = data.shape
    h = np.zeros(256, np.int32)
    h[0:4] = (nx, ny, nz, 2)
    hb = bytearray(h.tobytes())
    hb[208:212] = b"MAP "
    return bytes(hb) + data.astype(np.float32).tobytes()


def test_cli_prints_fsc_table(ctx, tmp_path):
    """-fsc 1: the table follows the "Save Sum to:" line and precedes
    "Done." (DFAlign.cu:1311-1321) with the rows of mc_align_movie."""
    if not os.path.exists(CLI):
        pytest.skip("CLI not built")
    frames, _ = make_movie(544, 512, 10, dose=1.0, step=0.7, seed=1001)
    open(tmp_path / "stack.mrc", "wb").write(_mrc_float(frames))
    r = subprocess.run([CLI, "stack.mrc", "-fod", "4", "-dsp", "0", "-fsc", "1"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    log = open(tmp_path / "stack_Log.txt").read()
    head = "\nFSC parallel(D) and perpendicular(U) to drift direction:\n" \
           "   Nq%       D_Raw       U_Raw       D_Corr       U_Corr\n"
    i = log.index(head)
    assert log.rindex("Save Sum to:") < i < log.index("Done.")
    rows = re.findall(r"^ *(\S+) +(\S+) +(\S+) +(\S+) +(\S+)$", log[i + len(head):log.index("Done.")],
                      re.M)
    table = np.array(rows, float)
    nb = Oracle("f64").fsc_bins(512)[0]
    assert table.shape == (nb, 5)
    assert np.abs(table[:, 0] - np.arange(nb) * 100.0 / nb).max() <= 0.005 + 1e-9
    p = _movie_params(10, FrameDistOffset=4)
    dev = mc.DeviceArray(frames)
    lib = ctx.align_movie(dev, frames.shape, frames.dtype, p, want=("fsc",))["fsc"]
    dev.free()
    assert np.abs(table[:, 1:] - lib).max() <= 5.01e-5
