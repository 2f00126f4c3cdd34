"""GPU tests of the new forms.  Differentiation: the adjoint of the
phase-shift-and-sum sharpness objective (mc_sharpness) against the FP64
oracle gradient on the same spectra and against central finite differences
(h = 1e-3 px), and the gradient refinement inside mc_align_movie
(grad_steps > 0).  Precision: the MFMA FP16 hi/lo form of the pair stage
against the FP64 oracle (box maps) and against the FP32 form (shifts within
0.01 px RMS, sums within 1e-4)."""
import json

import numpy as np
import pytest

import motioncorr1exp_amd as mc
from motioncorr1exp_amd.synth import make_movie
from _oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = mc.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def loaded(ctx):
    F, n = 6, 256
    frames, drift = make_movie(n, n, F, dose=3.0, step=0.6, seed=55)
    d = mc.DeviceArray(frames)
    ctx.load_frames(d, frames.shape, frames.dtype, mc.default_params(F, FrameDistOffset=1))
    spec = ctx.get_spectra(F, n)
    return F, n, frames, drift, spec, d


def test_gradient_matches_fp64_oracle(ctx, loaded):
    F, n, frames, drift, spec, _ = loaded
    rng = np.random.default_rng(1)
    s = (drift - drift[0]) + rng.normal(0, 0.4, size=(F, 2))
    J, g = ctx.sharpness(s, 150.0)
    oJ, og = Oracle("f64").sharpness(spec.astype(np.float64), n, 150.0, s)
    assert abs(J - oJ) <= 2e-6 * abs(oJ)
    assert np.abs(g - og).max() <= 1e-4 * np.abs(og).max()


def test_gradient_matches_central_differences(ctx, loaded):
    F, n, frames, drift, spec, _ = loaded
    rng = np.random.default_rng(2)
    s = (drift - drift[0]) + rng.normal(0, 0.4, size=(F, 2))
    J, g = ctx.sharpness(s, 150.0)
    h = 1e-3
    for f in range(F):
        for ax in range(2):
            sp, sm = s.copy(), s.copy()
            sp[f, ax] += h
            sm[f, ax] -= h
            fd = (ctx.sharpness(sp, 150.0)[0] - ctx.sharpness(sm, 150.0)[0]) / (2 * h)
            assert abs(fd - g[f, ax]) <= 2e-3 * np.abs(g).max() + 2e-4 * abs(J) / 1e3


def test_gradient_is_translation_invariant(ctx, loaded):
    F, n, frames, drift, spec, _ = loaded
    s = np.random.default_rng(3).normal(0, 1.0, size=(F, 2))
    J0, g = ctx.sharpness(s, 150.0)
    J1, _ = ctx.sharpness(s + np.array([0.37, -1.21]), 150.0)
    assert abs(J1 - J0) <= 1e-5 * abs(J0)
    assert np.abs(g.sum(0)).max() <= 1e-4 * np.abs(g).max()


def test_refinement_increases_sharpness(ctx):
    F, n = 8, 256
    frames, drift = make_movie(n, n, F, dose=4.0, step=0.4, seed=66)
    d = mc.DeviceArray(frames)
    base = mc.default_params(F, FrameDistOffset=2, CCPeakSearchDim=64, bAlignToMid=0)
    r0 = ctx.align_movie(d, frames.shape, frames.dtype, base, want=("frame_shift",))
    r1 = ctx.align_movie(d, frames.shape, frames.dtype, dict(base, grad_steps=20, grad_step_px=0.05),
                         want=("frame_shift",))
    rep = r1["report"]
    assert rep["sharp_j1"] >= rep["sharp_j0"]
    truth = drift - drift[0]
    e0 = r0["frame_shift"] - r0["frame_shift"][0] - truth
    e1 = r1["frame_shift"] - r1["frame_shift"][0] - truth
    # refinement leaves the 1/32 grid and does not move away from the truth
    assert np.sqrt((e1 ** 2).mean()) <= np.sqrt((e0 ** 2).mean()) + 0.01


# --------------------------------------------------- MFMA FP16 hi/lo form
def _forms(ctx, dev, shape, dtype, prm, want):
    r0 = ctx.align_movie(dev, shape, dtype, dict(prm, precision=0), want=want)
    r1 = ctx.align_movie(dev, shape, dtype, dict(prm, precision=1), want=want)
    return r0, r1


def test_mfma_form_box_maps_match_fp64_oracle(ctx):
    """Box maps of the hi/lo form against the FP64 oracle: within 1e-5 of
    each map's maximum, the bar of the FP32 form."""
    F, n = 6, 512
    frames, drift = make_movie(n, n, F, dose=2.0, step=0.7, seed=91)
    d = mc.DeviceArray(frames)
    prm = mc.default_params(F, FrameDistOffset=1, CCPeakSearchDim=64, precision=1)
    r = ctx.align_movie(d, frames.shape, frames.dtype, prm, want=("boxmaps", "pair_b"))
    o = Oracle("f64").align_movie(frames, dict(bin=1.0, fod=1, bfactor=150.0, box=64, full_c2r=1,
                                               nthreads=8))
    bm, obm = r["boxmaps"], o["boxmaps"]
    scale = np.abs(obm).reshape(len(obm), -1).max(1)[:, None, None]
    assert (np.abs(bm - obm) / scale).max() < 1e-5
    db = np.abs(r["pair_b"] - o["pair_b"])
    print("hi/lo form vs oracle: pair components beyond 0.01 px: %d" % int((db > 0.01).sum()))
    assert int((db > 0.01).sum()) == 0


@pytest.mark.parametrize("n,F,box", [(1024, 8, 96), (4096, 10, 96)])
def test_mfma_form_matches_fp32_form(ctx, n, F, box):
    """Per-frame shifts of the hi/lo form within 0.01 px RMS of the FP32 form
    and sums within 1e-4 relative L2, up to the 4096 canvas."""
    from motioncorr1exp_amd.synth import make_movie_gpu

    dev, drift = make_movie_gpu(n, n, F, dose=1.0, step=0.7, seed=n + F, dtype=np.float32)
    prm = mc.default_params(F, FrameDistOffset=2, CCPeakSearchDim=box, iter=2)
    want = ("frame_shift", "sum_image", "pair_b")
    r0, r1 = _forms(ctx, dev, (F, n, n), np.float32, prm, want)
    d = r1["frame_shift"] - r0["frame_shift"]
    assert np.sqrt(np.mean(d ** 2)) <= 0.01
    db = np.abs(r1["pair_b"] - r0["pair_b"])
    print("n %d: hi/lo vs FP32 frame shift RMS %.3g px, pair components beyond 0.01 px: %d"
          % (n, np.sqrt(np.mean(d ** 2)), int((db > 0.01).sum())))
    assert int((db > 0.01).sum()) == 0
    rel = np.linalg.norm(r1["sum_image"] - r0["sum_image"]) / np.linalg.norm(r0["sum_image"])
    assert rel < 1e-4, rel
    dev.free()


def test_mfma_form_rejects_unsupported_geometry(ctx):
    frames, _ = make_movie(256, 256, 4, dose=2.0, step=0.5, seed=5)
    d = mc.DeviceArray(frames)
    with pytest.raises(mc.MCError, match="precision form 1"):
        ctx.align_movie(d, frames.shape, frames.dtype,
                        mc.default_params(4, CCPeakSearchDim=64, precision=1))


# ---------------------------------------------------------------- autotuner
def test_autotune_picks_valid_shapes_and_keeps_results(ctx, tmp_path):
    """mc_autotune times every candidate, keeps the fastest, and the tuned
    path gives the same shifts as the default shapes within 0.01 px (it may
    add the box maps in another order; frames beyond 0.01 px are counted and
    must be 0);
    the cache file round-trips through mc_tune_load on a fresh context."""
    F, n = 8, 1024
    frames, drift = make_movie(n, n, F, dose=2.0, step=0.7, seed=123)
    d = mc.DeviceArray(frames)
    prm = mc.default_params(F, FrameDistOffset=2, CCPeakSearchDim=96)
    c0 = mc.Context(0)
    ref = c0.align_movie(d, frames.shape, frames.dtype, prm, want=("frame_shift", "sum_image"))
    c0.close()
    cache = str(tmp_path / "tune.jsonl")
    rep = ctx.autotune(d, frames.shape, frames.dtype, prm, cache_path=cache)
    shape, cols_p = rep["r2c_variant"] & 7, rep["r2c_variant"] >> 3
    assert shape in (1, 2) and cols_p in (0, 1) and rep["chunk_frames"] in (1, 2, 4, 8)
    assert rep["pair_cols_variant"] in (0, 1, 2, 3) and rep["pair_rows_variant"] in (0, 1, 2, 3)
    # radix plan x LDS padding of the column pass, timed on this 1024 canvas
    assert 0 <= rep["cols_plan"] < 6 and (cols_p or all(v > 0 for v in rep["ms_cols_plan"]))
    assert cols_p or rep["r2c_variant"] >> 4 == rep["cols_plan"]
    best = min(v for blk in rep["ms_stage1"] for row in blk for v in row if v > 0)
    if cols_p:
        assert rep["ms_stage1"][cols_p][shape][[1, 2, 4, 8].index(rep["chunk_frames"])] == best
    tuned = ctx.align_movie(d, frames.shape, frames.dtype, prm, want=("frame_shift", "sum_image"))
    dd = np.abs(tuned["frame_shift"] - ref["frame_shift"])
    print("tuned vs default shapes: frame shift max difference %.3g px" % dd.max())
    assert int((dd > 0.01).sum()) == 0
    c1 = mc.Context(0)
    assert c1.tune_load(cache) == 1
    again = c1.align_movie(d, frames.shape, frames.dtype, prm, want=("frame_shift",))
    assert np.array_equal(again["frame_shift"], tuned["frame_shift"])
    c1.close()


@pytest.mark.parametrize("n", [1024, 4096])
def test_column_plans_match_oracle(ctx, tmp_path, n):
    """Every radix plan x LDS padding of the stage-1 column pass (the
    autotuner's candidates) gives spectra within 5e-7 of the FP64 oracle's.
    Each variant is forced through a tuning-table line (mc_tune_load)."""
    F = 3
    frames, _ = make_movie(n, n, F, dose=1.0, step=0.7, seed=n + 7)
    d = mc.DeviceArray(frames)
    prm = mc.default_params(F, FrameDistOffset=0, CCPeakSearchDim=96)
    ref = Oracle("f64").align_movie(frames, dict(fod=0, box=96, nthreads=16),
                                    want_spectra=True, want_boxes=False)["spectra"].view(np.complex128)
    cache = str(tmp_path / "tune.jsonl")
    c0 = mc.Context(0)
    c0.autotune(d, frames.shape, frames.dtype, prm, cache_path=cache)
    c0.close()
    line = json.loads(open(cache).read().splitlines()[-1])
    for cpl in range(6):
        line["r2c"] = 1 | (cpl << 4)  # 4 rows x 256 threads, per-column pass variant cpl
        fn = str(tmp_path / ("plan%d.jsonl" % cpl))
        open(fn, "w").write(json.dumps(line) + "\n")
        c = mc.Context(0)
        assert c.tune_load(fn) == 1
        c.load_frames(d, frames.shape, frames.dtype, prm)
        spec = c.get_spectra(F, n).view(np.complex64)
        c.close()
        err = np.abs(spec - ref).max() / np.abs(ref).max()
        print("n %d column variant %d (plan %d, padding %s): spectra error / max %.3g"
              % (n, cpl, cpl // 2, "none" if cpl % 2 else "1/16", err), flush=True)
        assert err < 5e-7
    d.free()

