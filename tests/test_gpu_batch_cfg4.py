"""GPU test of the batching form on movies shaped like the 64-movie batch
workload: random-field phantoms (motioncorr1exp_amd.synth.make_movie), sides
from 1024 to 4096, 8 to 50 frames, dose 0.5-1.5 e/px/frame, 8-bit counts,
each padded to the next supported transform size, -fod 2 -bft 150 -pbx 96
-itr 3.  Six movies with six different transform sizes go through one
mc_align_batch call.  Every movie must get bit for bit what mc_align_movie
gives it alone (pair shifts, box maps, round measurements, final shifts,
sum), and every movie is checked against the FP64 oracle: per-frame shifts
within 0.01 px and aligned sum within 1e-4 relative L2.  The number of
frames off by more than 0.01 px (a 1/32-grid flip of the peak pick on a
near tie would show here) is printed per movie and must be 0."""
import numpy as np
import pytest

import motioncorr1exp_amd as mc
from motioncorr1exp_amd.synth import make_movie
from _oracle import Oracle

pytestmark = pytest.mark.gpu

# (nx, ny, frames, dose, canvas, seed)
SPECS = [(1024, 1024, 50, 0.6, 1024, 401), (1300, 1250, 9, 1.4, 1536, 402),
         (2048, 2048, 23, 1.0, 2048, 403), (2500, 2400, 14, 0.8, 2560, 404),
         (3000, 3000, 11, 1.2, 3072, 405), (4000, 3900, 18, 0.5, 4096, 406)]
WANT = ("frame_shift", "pair_b", "pair_cc", "iter_meas", "rmsd", "sum_image", "boxmaps",
        "final_shift")


@pytest.fixture(scope="module")
def ctx():
    c = mc.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def movies():
    out = []
    for nx, ny, F, dose, pad, seed in SPECS:
        frames, drift = make_movie(nx, ny, F, dose=dose, step=0.7, seed=seed, dtype=np.uint8)
        prm = mc.default_params(F, FrameDistOffset=2, CCPeakSearchDim=96, iter=3, pad=pad)
        out.append(dict(frames=frames, drift=drift, params=prm, pad=pad,
                        dev=mc.DeviceArray(frames)))
        print("movie %dx%dx%d dose %.1f canvas %d" % (nx, ny, F, dose, pad), flush=True)
    return out


@pytest.fixture(scope="module")
def batch(ctx, movies):
    res, br = ctx.align_batch([dict(dev=m["dev"], shape=m["frames"].shape,
                                    dtype=m["frames"].dtype, params=m["params"], want=WANT)
                               for m in movies])
    return res, br


def test_batch_bit_identical_to_single(ctx, movies, batch):
    res, br = batch
    # one group and one sub-batch: the column pass of all six sizes runs in
    # the batched plan kernel, one graph replay per round
    assert br["groups"] == 1 and br["sub_batches"] == 1
    for m, r in zip(movies, res):
        s = ctx.align_movie(m["dev"], m["frames"].shape, m["frames"].dtype, m["params"],
                            want=WANT)
        for k in WANT:
            assert np.array_equal(r[k], s[k]), (m["pad"], k)
        for k in ("niter", "npairs", "ngood", "n", "nframe", "converged"):
            assert r["report"][k] == s["report"][k], (m["pad"], k)
        assert r["report"]["rmsd"] == s["report"]["rmsd"]
        print("canvas %d: %d pairs, %d rounds, batch == single bitwise" %
              (m["pad"], r["report"]["npairs"], r["report"]["niter"]), flush=True)


@pytest.mark.parametrize("which", range(len(SPECS)))
def test_batch_movie_matches_oracle(movies, batch, which):
    m, r = movies[which], batch[0][which]
    o = Oracle("f64").align_movie(m["frames"].astype(np.float32),
                                  dict(bin=1.0, fod=2, bfactor=150.0, box=96, pad=m["pad"],
                                       iter=3, full_c2r=0, nthreads=16), want_boxes=False)
    d = np.abs(r["frame_shift"] - o["frame_shift"])
    flips = int((d > 0.01).sum())
    rms = float(np.sqrt(np.mean(d ** 2)))
    pair_flips = int((np.abs(r["pair_b"] - o["pair_b"]) > 1e-6).sum())
    rel = np.linalg.norm(r["sum_image"] - o["sum_image"]) / np.linalg.norm(o["sum_image"])
    print("canvas %d: frame shift RMS %.3g px max %.3g px, frames > 0.01 px: %d, pair "
          "components differing: %d, sum rel L2 %.3g" % (m["pad"], rms, d.max(), flips,
                                                         pair_flips, rel), flush=True)
    assert flips == 0 and rms <= 0.01
    assert rel < 1e-4
