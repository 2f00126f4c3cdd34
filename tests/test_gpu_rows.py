"""GPU tests of the pair stage's box maps over the launch shapes of the row
pass (the C2R row transform restricted to the box, reference DFAlign.cu:897
with FindShift's read-back, cufunc.cu:396-424): pair counts from 1 to 300
(at n = 2048 the kx range of every pair is split over 2 workgroups whose
partial maps are added in order, whatever the pair count) and boxes of one,
two and three 32-row blocks, whole and partial.  Spectra of random real
images are set directly, so the only difference from the FP64 oracle's
cross power (cufunc.cu:343-370) and pruned box inverse is the FP32
arithmetic."""
import numpy as np
import pytest

import motioncorr1exp_amd as mc
from _oracle import Oracle

pytestmark = pytest.mark.gpu
N, F = 2048, 25


@pytest.fixture(scope="module")
def loaded():
    rng = np.random.default_rng(4242)
    spec = np.fft.rfft2(rng.standard_normal((F, N, N))).astype(np.complex64)
    ctx = mc.Context(0)
    ctx.set_spectra(spec.view(np.float32))
    yield ctx, spec
    ctx.close()


@pytest.fixture(scope="module")
def o():
    return Oracle("f64")


def _pairs(npairs):
    allp = np.array([(a, b) for a in range(F) for b in range(a + 1, F)], np.int32)
    order = np.random.default_rng(7).permutation(len(allp))
    return allp[order[:npairs]]


@pytest.mark.parametrize("npairs", [1, 7, 40, 129, 300])
@pytest.mark.parametrize("box", [32, 48, 64, 80, 96])
def test_box_maps_match_oracle(loaded, o, npairs, box):
    ctx, spec = loaded
    pairs = _pairs(npairs)
    _, _, boxes = ctx.pair_shifts(pairs, 150.0, box, want_boxes=True)
    check = sorted({0, npairs - 1, npairs // 2, npairs // 3})
    worst = 0.0
    for q in check:
        a, b = pairs[q]
        ref = o.pruned_box(o.cross_power(spec[a], spec[b], N, 150.0), box)
        worst = max(worst, float(np.abs(boxes[q] - ref).max() / np.abs(ref).max()))
    print("pairs", npairs, "box", box, "max error / max", worst, flush=True)
    assert worst < 1e-5


@pytest.mark.parametrize("box", [64, 96])
def test_map_independent_of_launch_pair_count(loaded, box):
    """The same pair alone, among 7, 40 and 300 gives a bit-identical map and
    shift: the row pass splits a pair's kx range by n alone, so the FP32
    summation order does not change with the launch (a single movie, a batch
    or an align-to-sum round)."""
    ctx, _ = loaded
    pairs = _pairs(300)
    b_all, cc_all, many = ctx.pair_shifts(pairs, 150.0, box, want_boxes=True)
    for cnt in (1, 7, 40):
        for q in (0, 150, 299):
            sub = pairs[q:q + cnt] if q + cnt <= 300 else pairs[q - cnt + 1:q + 1]
            at = 0 if q + cnt <= 300 else cnt - 1
            b, cc, one = ctx.pair_shifts(sub, 150.0, box, want_boxes=True)
            assert np.array_equal(one[at], many[q]), (cnt, q)
            assert np.array_equal(b[at], b_all[q]) and cc[at] == cc_all[q], (cnt, q)


@pytest.mark.parametrize("npairs", [1, 40, 96, 129])
def test_peak_search_slices_match_oracle(loaded, o, npairs):
    """With fewer pairs than half the CUs, the peak search splits the 512 map
    rows of each pair over 2 or 4 workgroups and combines their first maxima
    (FindShift, cufunc.cu:396-497).  On 256 CUs, 1 and 40 pairs take 4
    slices, 96 take 2 and 129 one.  Box maps with one Gaussian peak at a
    random sub-pixel position over weak noise: shifts identical to the FP64
    oracle's search of the same FP32 maps, cc within 1e-9."""
    ctx, _ = loaded
    rng = np.random.default_rng(900 + npairs)
    y, x = np.mgrid[0:96, 0:96]
    maps = np.empty((npairs, 96, 96), np.float32)
    for q in range(npairs):
        x0, y0 = rng.uniform(20, 76, 2)
        maps[q] = 1e3 * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / 8.0) + rng.normal(0, 1.0, (96, 96))
    b, cc = ctx.find_peaks(maps, N)
    for q in range(npairs):
        ox, oy, occ = o.find_shift_box(maps[q].astype(np.float64), N)
        assert (b[q, 0], b[q, 1]) == (ox, oy), q
        assert abs(cc[q] - occ) <= 1e-9 * abs(occ), q
