"""GPU tests of the differentiation form: the adjoint of the phase-shift-and-
sum sharpness objective (mc_sharpness) against the FP64 oracle gradient on
the same spectra and against central finite differences (h = 1e-3 px), and
the gradient refinement inside mc_align_movie (grad_steps > 0)."""
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
