"""GPU tests of the batched form (mc_align_batch): movies of uneven size and
frame count in one call must give each movie exactly what mc_align_movie
gives it (the batched column-pass kernel runs the same per-item body as the
single-movie kernel; the other kernels are shared), with the align-to-sum
rounds replayed from a captured HIP graph or launched eagerly
(MC_NO_GRAPH=1), in one sub-batch or split by a small memory budget.  One
movie of the batch is also checked against the FP64 oracle (shifts within
0.01 px, sum within 1e-4 relative L2).  tests/test_gpu_batch_cfg4.py covers
movies of the batch workload's sizes."""
import os

import numpy as np
import pytest

import motioncorr1exp_amd as mc
from motioncorr1exp_amd.synth import make_movie
from _oracle import Oracle

pytestmark = pytest.mark.gpu

WANT = ("frame_shift", "pair_b", "iter_meas", "rmsd", "sum_image", "frame_mean", "boxmaps")


@pytest.fixture(scope="module")
def ctx():
    c = mc.Context(0)
    yield c
    c.close()


def _movies():
    # (side x, side y, frames, pad, iter, fod, seed): two transform sizes,
    # uneven frame counts, a non-power-of-two side padded to 512
    specs = [(256, 256, 6, 0, 2, 1, 11), (400, 376, 9, 512, 2, 2, 12),
             (256, 256, 8, 0, 0, 2, 13), (512, 512, 5, 0, 3, 1, 14),
             (256, 256, 7, 0, 1, 2, 15)]
    out = []
    for nx, ny, F, pad, itr, fod, seed in specs:
        frames, drift = make_movie(nx, ny, F, dose=2.0, step=0.8, seed=seed)
        prm = mc.default_params(F, FrameDistOffset=fod, CCPeakSearchDim=64, iter=itr, pad=pad)
        out.append(dict(frames=frames, drift=drift, params=prm))
    return out


@pytest.fixture(scope="module")
def movies():
    ms = _movies()
    for m in ms:
        m["dev"] = mc.DeviceArray(m["frames"])
    return ms


@pytest.fixture(scope="module")
def singles(ctx, movies):
    return [ctx.align_movie(m["dev"], m["frames"].shape, m["frames"].dtype, m["params"], want=WANT)
            for m in movies]


def _batch(ctx, movies):
    return ctx.align_batch([dict(dev=m["dev"], shape=m["frames"].shape, dtype=m["frames"].dtype,
                                 params=m["params"], want=WANT) for m in movies])


def _same(a, b):
    """Batch and single runs use the same kernels with the same per-pair
    arithmetic and summation order, so every output is bit-identical."""
    for k in WANT:
        assert np.array_equal(a[k], b[k]), k
    for k in ("niter", "npairs", "ngood", "n", "nframe"):
        assert a["report"][k] == b["report"][k], k
    assert a["report"]["rmsd"] == b["report"]["rmsd"]


def test_batch_matches_single_movies(ctx, movies, singles):
    res, br = _batch(ctx, movies)
    # one group across both transform sizes (same box, B-factor, form)
    assert br["groups"] == 1 and br["sub_batches"] == 1
    # one graph replay per round for every iterating movie of both sizes
    # (max iter 3), a new capture whenever the set of iterating movies changes
    assert br["graph_replays"] == 3
    assert br["graph_captures"] >= 1
    for a, b in zip(res, singles):
        _same(a, b)


def test_batch_eager_rounds_match_graph(ctx, movies, singles):
    os.environ["MC_NO_GRAPH"] = "1"
    try:
        res, br = _batch(ctx, movies)
    finally:
        del os.environ["MC_NO_GRAPH"]
    assert br["graph_replays"] == 0 and br["graph_captures"] == 0
    for a, b in zip(res, singles):
        _same(a, b)


def test_batch_split_by_memory_budget(ctx, movies, singles):
    os.environ["MC_BATCH_MEM_MB"] = "1"  # one movie per sub-batch
    try:
        res, br = _batch(ctx, movies)
    finally:
        del os.environ["MC_BATCH_MEM_MB"]
    assert br["sub_batches"] == len(movies)
    for a, b in zip(res, singles):
        _same(a, b)


def test_batch_movie_matches_oracle(ctx, movies):
    res, _ = _batch(ctx, movies)
    m, r = movies[1], res[1]
    prm = m["params"]
    o = Oracle("f64").align_movie(m["frames"], dict(bin=1.0, fod=prm["FrameDistOffset"],
                                                    bfactor=150.0, box=64, pad=512,
                                                    iter=prm["iter"], full_c2r=1, nthreads=8),
                                  want_boxes=False)
    d = np.abs(r["frame_shift"] - o["frame_shift"])
    flips = int((d > 0.01).sum())
    print("frame shift max error %.3g px, frames > 0.01 px: %d" % (d.max(), flips), flush=True)
    assert flips == 0
    rel = np.linalg.norm(r["sum_image"] - o["sum_image"]) / np.linalg.norm(o["sum_image"])
    assert rel < 1e-4, rel


def test_batch_rejects_bad_item(ctx, movies):
    bad = dict(movies[0]["params"], CCPeakSearchDim=8)
    items = [dict(dev=m["dev"], shape=m["frames"].shape, dtype=m["frames"].dtype,
                  params=m["params"]) for m in movies[:2]]
    items.append(dict(dev=movies[0]["dev"], shape=movies[0]["frames"].shape,
                      dtype=movies[0]["frames"].dtype, params=bad))
    with pytest.raises(mc.MCError, match="batch item 2"):
        ctx.align_batch(items)
