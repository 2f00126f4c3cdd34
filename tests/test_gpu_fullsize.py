"""GPU parity at full benchmark sizes.  First, 4096x4096x24 uint8 Poisson counts
(mean 1.0 over the random-field phantom, random-walk drift with N(0, 1.0 px)
steps), 2x Fourier-crop binning, all 276 frame pairs, B-factor 150, peak box
96 and 3 align-to-sum rounds.  The C ABI's results are compared with the FP64
oracle run on the same frames (its pruned box inverse, which equals the
reference's full inverse plus read-out): pair shifts and final shifts to the
0.01 px bar (components beyond it counted, printed, required to be 0), the aligned sum to
1e-4 relative, and the recovered drift against the known drift.  Then the
3838x3710x40 headline movie in the plain FP32 form and in the MFMA FP16
hi/lo precision form, the differentiation form at the same size, and the
8192 canvas (below)."""
import numpy as np
import pytest

import motioncorr1exp_amd as mc
from _oracle import Oracle
from motioncorr1exp_amd.synth import make_movie

pytestmark = pytest.mark.gpu


def test_4096x4096x24_uint8_bin2_itr3_matches_oracle():
    F = 24
    frames, drift = make_movie(4096, 4096, F, dose=1.0, step=1.0, seed=2401, dtype=np.uint8)
    print("movie ready", flush=True)
    ctx = mc.Context(0)
    p = mc.default_params(F, bin=2.0, FrameDistOffset=11, iter=3, CCPeakSearchDim=96)
    dev = mc.DeviceArray(frames)
    r = ctx.align_movie(dev, frames.shape, frames.dtype, p,
                        want=("pair_b", "frame_shift", "sum_image"))
    dev.free()
    ctx.close()
    assert r["report"]["n"] == 2048 and r["report"]["npairs"] == 276
    print("gpu done; oracle ...", flush=True)
    o = Oracle("f64").align_movie(frames.astype(np.float32),
                                  dict(bin=2.0, fod=11, iter=3, box=96, full_c2r=0, nthreads=16),
                                  want_boxes=False)
    db = np.abs(r["pair_b"] - o["pair_b"])
    flips = int((db > 1e-6).sum())
    fs = r["frame_shift"] - o["frame_shift"]
    rms = float(np.sqrt(np.mean(fs ** 2)))
    rel = float(np.linalg.norm(r["sum_image"] - o["sum_image"]) / np.linalg.norm(o["sum_image"]))
    # known drift: frame shifts are t_f - t_ref (reference frame nz/2+1) in
    # binned pixels
    truth = (drift - drift[F // 2 + 1]) / 2.0
    vs_truth = float(np.sqrt(np.mean((r["frame_shift"] - truth) ** 2)))
    print("pair-shift flips", flips, "max", db.max(), "frame-shift RMS", rms, "sum rel", rel,
          "vs truth RMS", vs_truth, flush=True)
    assert flips == 0 and rms < 0.01
    assert rel < 1e-4
    assert vs_truth < 0.1


def test_3838x3710x40_plain_and_precision_forms_match_oracle():
    """The headline workload: 3838x3710x40 float32 counts (dose 0.8, random
    walk N(0, 0.7 px)), padded to 4096, -fod 2 (253 pairs), B-factor 150,
    box 96, -itr 0, run in the plain FP32 form and in the MFMA FP16 hi/lo
    form (precision=1).  Both against the FP64 oracle on the same frames:
    frame shifts within 0.01 px (RMS and every frame), aligned sum within
    1e-4 relative L2; pair-shift and frame-shift differences beyond 0.01 px
    are counted, printed and must be 0."""
    F, n = 40, 4096
    frames, drift = make_movie(3838, 3710, F, dose=0.8, step=0.7, seed=1003)
    print("movie ready", flush=True)
    ctx = mc.Context(0)
    dev = mc.DeviceArray(frames)
    want = ("pair_b", "frame_shift", "sum_image")
    p = mc.default_params(F, pad=n, FrameDistOffset=2, CCPeakSearchDim=96)
    res = {form: ctx.align_movie(dev, frames.shape, frames.dtype, dict(p, precision=form),
                                 want=want) for form in (0, 1)}
    dev.free()
    ctx.close()
    print("gpu done; oracle ...", flush=True)
    o = Oracle("f64").align_movie(frames, dict(pad=n, fod=2, box=96, iter=0, full_c2r=0,
                                               nthreads=16), want_boxes=False)
    assert res[0]["report"]["npairs"] == 253
    for form, r in res.items():
        db = np.abs(r["pair_b"] - o["pair_b"])
        fs = np.abs(r["frame_shift"] - o["frame_shift"])
        rms = float(np.sqrt(np.mean(fs ** 2)))
        rel = float(np.linalg.norm(r["sum_image"] - o["sum_image"]) / np.linalg.norm(o["sum_image"]))
        print("precision %d: pair components > 0.01 px off: %d (max %.3g), frame shift RMS %.3g "
              "max %.3g, frames > 0.01 px: %d, sum rel L2 %.3g" %
              (form, int((db > 0.01).sum()), db.max(), rms, fs.max(), int((fs > 0.01).sum()), rel),
              flush=True)
        assert int((db > 0.01).sum()) == 0 and int((fs > 0.01).sum()) == 0
        assert rms <= 0.01 and rel < 1e-4
    d = np.abs(res[1]["frame_shift"] - res[0]["frame_shift"])
    print("precision form vs plain form: frame shift RMS %.3g px" % np.sqrt(np.mean(d ** 2)))
    assert np.sqrt(np.mean(d ** 2)) <= 0.01


def test_3838x3710x40_differentiation_form():
    """The differentiation form at its full size: 3838x3710x40 float32 frames
    (padded to 4096) with sub-pixel random-walk drift N(0, 0.4 px) plus a
    quadratic slow drift.  The adjoint gradient of the sharpness on the
    resident spectra is checked against central differences (h = 1e-3 px) and
    against the FP64 oracle gradient on the same spectra; 20 refinement steps
    raise the sharpness and do not move the shifts away from the known drift."""
    F, n = 40, 4096
    frames, drift = make_movie(3838, 3710, F, dose=0.8, step=0.4, seed=2402, quad_sigma=0.02)
    print("movie ready", flush=True)
    ctx = mc.Context(0)
    p = mc.default_params(F, pad=n, FrameDistOffset=2, CCPeakSearchDim=96)
    dev = mc.DeviceArray(frames)
    r0 = ctx.align_movie(dev, frames.shape, frames.dtype, p, want=("frame_shift",))
    r1 = ctx.align_movie(dev, frames.shape, frames.dtype, dict(p, grad_steps=20, grad_step_px=0.05),
                         want=("frame_shift",))
    rep = r1["report"]
    truth = drift - drift[F // 2 + 1]
    e0 = float(np.sqrt(np.mean((r0["frame_shift"] - truth) ** 2)))
    e1 = float(np.sqrt(np.mean((r1["frame_shift"] - truth) ** 2)))
    print("J0", rep["sharp_j0"], "J1", rep["sharp_j1"], "accepted", rep["grad_accepted"],
          "vs truth RMS before", e0, "after", e1, flush=True)
    assert rep["sharp_j1"] >= rep["sharp_j0"]
    assert e1 <= e0 + 0.01
    # gradient checks on the resident spectra, away from the optimum
    ctx.load_frames(dev, frames.shape, frames.dtype, p)
    s = r0["frame_shift"] + np.random.default_rng(7).normal(0, 0.4, size=(F, 2))
    J, g = ctx.sharpness(s, 150.0)
    h = 1e-3
    for f, ax in [(0, 0), (1, 1), (5, 1), (10, 0), (20, 1), (25, 0), (30, 0), (39, 1)]:
        sp, sm = s.copy(), s.copy()
        sp[f, ax] += h
        sm[f, ax] -= h
        fd = (ctx.sharpness(sp, 150.0)[0] - ctx.sharpness(sm, 150.0)[0]) / (2 * h)
        print("frame", f, "axis", ax, "adjoint", g[f, ax], "central difference", fd, flush=True)
        assert abs(fd - g[f, ax]) <= 2e-3 * np.abs(g).max()
    spec = ctx.get_spectra(F, n)
    dev.free()
    ctx.close()
    oJ, og = Oracle("f64").sharpness(spec.astype(np.float64), n, 150.0, s)
    print("J rel err", abs(J - oJ) / abs(oJ), "gradient max err / max",
          np.abs(g - og).max() / np.abs(og).max(), flush=True)
    assert abs(J - oJ) <= 2e-6 * abs(oJ)
    assert np.abs(g - og).max() <= 1e-4 * np.abs(og).max()


def test_8192_canvas_matches_oracle():
    """The largest canvas: 8192x8192x5 float32, fod 1, one align-to-sum round,
    so every n = 8192 kernel runs: stage 1, the column pass in kx segments,
    the row pass, the peak search, the batched and final sums and the C2R of
    the sum image.  Pair shifts and frame shifts against the FP64 oracle on
    the same frames within 0.01 px, the aligned sum to 1e-4 relative."""
    F, n = 5, 8192
    frames, _ = make_movie(n, n, F, dose=1.0, step=0.8, seed=2403)
    print("movie ready", flush=True)
    ctx = mc.Context(0)
    p = mc.default_params(F, FrameDistOffset=1, iter=1, CCPeakSearchDim=96)
    dev = mc.DeviceArray(frames)
    r = ctx.align_movie(dev, frames.shape, frames.dtype, p,
                        want=("pair_b", "frame_shift", "sum_image"))
    dev.free()
    ctx.close()
    assert r["report"]["n"] == n
    print("gpu done; oracle ...", flush=True)
    o = Oracle("f64").align_movie(frames, dict(fod=1, iter=1, box=96, full_c2r=0, nthreads=16),
                                  want_boxes=False)
    db = np.abs(r["pair_b"] - o["pair_b"])
    fs = np.abs(r["frame_shift"] - o["frame_shift"])
    rel = float(np.linalg.norm(r["sum_image"] - o["sum_image"]) / np.linalg.norm(o["sum_image"]))
    print("pair-shift max diff", db.max(), "frame-shift max diff", fs.max(), "sum rel", rel, flush=True)
    assert int((db > 0.01).sum()) == 0 and fs.max() <= 0.01
    assert rel < 1e-4
