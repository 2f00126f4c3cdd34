"""GPU test of the level-0 seam (include/mc_abi.h, the cufunc.h entry points)
used the way the reference's worker uses it: the control flow of
CDFAlign::ThreadFunc_cuAlign (DFAlign.cu:565-1313) replayed call for call
through level-0 entry points only -- per frame R2C, Fourier-crop binning
with GPUMemBinD2D + D2D and the even/odd raw sums with GPUAdd
(:768-823); the raw sum's log-modulus display and C2R x 1/size
(:842-857); per pair two uploads, GPUShiftCC, C2R and FindShift (:884-905);
the drift model (mc_lsq / mc_threshold, :914-948); align-to-sum rounds with
GPUShift + GPUAdd sums and GPUShiftCC / C2R / FindShift against the sum
(:968-1052); the final shifted sum into alternating accumulators with the
display crop GPUMemBinD2H (:1114-1292).  The per-frame shifts and the
aligned sum are compared with the FP64 oracle (0.01 px, 1e-4 relative L2);
GPUFFTLogModulus and GPUMemBinD2H are checked directly against numpy on the
same device spectrum."""
import ctypes

import numpy as np
import pytest

import motioncorr1exp_amd as mc
from motioncorr1exp_amd.synth import make_movie
from _oracle import Oracle

pytestmark = pytest.mark.gpu

NU, BIN, F, FOD, BOX, ITER, BFT, FFTSCALE, DISPDIM = 1024, 2.0, 8, 2, 96, 2, 150.0, 1e-4, 400


class Seam:
    """Thin helpers over the level-0 entry points (sizes in floats)."""

    def __init__(self):
        self.L = mc.lib()

    def alloc(self, nfloats):
        p = ctypes.c_void_p()
        mc.check(self.L.mc_mem_alloc(ctypes.byref(p), 4 * nfloats))
        return p.value

    def h2d(self, dst, host):
        h = np.ascontiguousarray(host, np.float32)
        mc.check(self.L.mc_mem_h2d(dst, h.ctypes.data, h.nbytes))

    def d2h(self, src, shape):
        out = np.empty(shape, np.float32)
        mc.check(self.L.mc_mem_d2h(out.ctypes.data, src, out.nbytes))
        return out


def _replay(frames):
    s = Seam()
    L = s.L
    mc.check(L.mc_init_gpu(0))
    n = int(NU / BIN) // 2 * 2  # DFAlign.cu:515
    sizeb, sizebu = n * (n + 2), NU * (NU + 2)
    dsum, dsumcorr, dtmp = s.alloc(sizeb), s.alloc(sizeb), s.alloc(sizeb)
    fft_plan = ctypes.c_void_p()
    mc.check(L.mc_fft_plan(ctypes.byref(fft_plan), NU))
    dfft = s.alloc(sizebu)
    hpos = (mc.MaskElem * (n * (n // 2 + 1)))()
    mc.check(L.mc_make_pos_list(hpos, n, BFT))
    dpos = s.alloc(3 * n * (n // 2 + 1))
    mc.check(L.mc_mem_h2d(dpos, ctypes.addressof(hpos), ctypes.sizeof(hpos)))
    mc.check(L.mc_mem_zero(dsum, 4 * sizeb))
    mc.check(L.mc_mem_zero(dsumcorr, 4 * sizeb))
    # stage 1 (:768-823): crop2fft, R2C, Fourier-crop binning, raw half sums
    hbuf = np.empty((F, n, n + 2), np.float32)
    even = True
    for j in range(F):
        buf = np.zeros((NU, NU + 2), np.float32)
        buf[:, :NU] = frames[j, :NU, :NU]
        s.h2d(dfft, buf)
        mc.check(L.mc_fft2d(dfft, fft_plan))
        mc.check(L.mc_mem_bin_d2d(dtmp, dfft, n, NU))
        mc.check(L.mc_mem_d2d(dfft, dtmp, 4 * sizeb))
        mc.check(L.mc_add(dsum if even else dsumcorr, dfft, sizeb))
        even = not even
        hbuf[j] = s.d2h(dfft, (n, n + 2))
    mc.check(L.mc_add(dsum, dsumcorr, sizeb))
    mc.check(L.mc_sync())
    mc.check(L.mc_mem_free(dfft))
    mc.check(L.mc_fft_destroy(fft_plan))
    dfft = s.alloc(sizeb)
    mc.check(L.mc_mem_zero(dsumcorr, 4 * sizeb))
    ifft_plan = ctypes.c_void_p()
    mc.check(L.mc_ifft_plan(ctypes.byref(ifft_plan), n))
    # raw-sum display and image (:842-857)
    rawspec = s.d2h(dsum, (n, n + 2)).view(np.complex64)
    mc.check(L.mc_fft_log_modulus(dfft, dsum, n, FFTSCALE))
    mc.check(L.mc_sync())
    logmod = s.d2h(dfft, (n, n // 2 + 1))
    mc.check(L.mc_ifft2d(dsum, ifft_plan))
    mc.check(L.mc_multiply_num(dsum, 1.0 / (n * n), sizeb))
    rawsum = s.d2h(dsum, (n, n + 2))[:, :n]
    # pair shifts (:879-905)
    pairs = np.zeros((F * F, 2), np.int32)
    npairs = L.mc_pair_list(F, FOD, pairs.ctypes.data, F * F)
    pairs = pairs[:npairs]
    b = np.zeros((npairs, 2))
    hbox = np.zeros((BOX, BOX), np.float32)
    sx, sy, cc = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()

    def find_shift(dmap):
        mc.check(L.mc_find_shift(dmap, n, hbox.ctypes.data, BOX, ctypes.byref(sx),
                                 ctypes.byref(sy), ctypes.byref(cc), -1))
        return sx.value, sy.value

    for q, (a, c) in enumerate(pairs):
        s.h2d(dsum, hbuf[a])
        s.h2d(dfft, hbuf[c])
        mc.check(L.mc_shift_cc(dfft, dsum, dpos, 0.0, 0.0, n))
        mc.check(L.mc_ifft2d(dfft, ifft_plan))
        b[q] = find_shift(dfft)
    # drift model (:914-948)
    shift = np.zeros((F - 1, 2))
    mc.check(L.mc_lsq(pairs.ctypes.data, npairs, F, b.ctypes.data, shift.ctypes.data))
    pos = np.vstack([np.zeros(2), np.cumsum(shift, 0)])
    ki = np.hypot(*(pos[pairs[:, 1]] - pos[pairs[:, 0]] - b).T)
    good = np.zeros(npairs, np.int32)
    ng = L.mc_threshold(pairs.ctypes.data, npairs, F, ki.ctypes.data, 1.0, good.ctypes.data)
    gp = np.ascontiguousarray(pairs[good[:ng]])
    gb = np.ascontiguousarray(b[good[:ng]])
    mc.check(L.mc_lsq(gp.ctypes.data, ng, F, gb.ctypes.data, shift.ctypes.data))
    # align-to-sum rounds (:953-1052); the oracle's rmsd recursion
    old = np.vstack([np.zeros(2), np.cumsum(shift, 0)])
    for it in range(ITER):
        s.h2d(dsum, hbuf[0])
        mc.check(L.mc_shift(dsum, dpos, -old[0, 0], -old[0, 1], n))
        for k in range(1, F):
            s.h2d(dfft, hbuf[k])
            mc.check(L.mc_shift(dfft, dpos, -old[k, 0], -old[k, 1], n))
            mc.check(L.mc_add(dsum, dfft, sizeb))
        new = np.zeros_like(old)
        for k in range(F):
            s.h2d(dfft, hbuf[k])
            mc.check(L.mc_shift_cc(dfft, dsum, dpos, 0.0, 0.0, n))
            mc.check(L.mc_ifft2d(dfft, ifft_plan))
            new[k] = find_shift(dfft)
        old = new - new[0]
    shift = old[1:] - old[:-1]
    # final sum (:1114-1292): -atm 1, frames 0..F-1, alternating accumulators
    mc.check(L.mc_mem_zero(dsum, 4 * sizeb))
    mc.check(L.mc_mem_zero(dsumcorr, 4 * sizeb))
    ref_frame = F // 2 + 1
    mid = shift[:ref_frame].sum(0)
    total = np.zeros(2)
    applied = np.zeros((F, 2))
    s.h2d(dsumcorr, hbuf[0])
    applied[0] = -total + mid
    mc.check(L.mc_shift(dsumcorr, dpos, applied[0, 0], applied[0, 1], n))
    even = False
    for j in range(1, F):
        total = total + shift[j - 1]
        applied[j] = -total + mid
        s.h2d(dfft, hbuf[j])
        mc.check(L.mc_shift(dfft, dpos, applied[j, 0], applied[j, 1], n))
        mc.check(L.mc_add(dsumcorr if even else dsum, dfft, sizeb))
        even = not even
    mc.check(L.mc_add(dsumcorr, dsum, sizeb))
    mc.check(L.mc_sync())
    sumspec = s.d2h(dsumcorr, (n, n + 2)).view(np.complex64)
    hdisp = np.zeros((DISPDIM, DISPDIM + 2), np.float32)
    mc.check(L.mc_mem_bin_d2h(hdisp.ctypes.data, dsumcorr, DISPDIM, n))
    mc.check(L.mc_ifft2d(dsumcorr, ifft_plan))
    mc.check(L.mc_multiply_num(dsumcorr, 1.0 / (n * n), sizeb))
    sumimg = s.d2h(dsumcorr, (n, n + 2))[:, :n]
    mc.check(L.mc_fft_destroy(ifft_plan))
    for p in (dsum, dsumcorr, dtmp, dfft, dpos):
        mc.check(L.mc_mem_free(p))
    return dict(n=n, pair_b=b, frame_shift=applied, sum_image=sumimg, rawsum=rawsum,
                rawspec=rawspec, logmod=logmod, sumspec=sumspec, hdisp=hdisp)


@pytest.fixture(scope="module")
def replay():
    frames, _ = make_movie(NU, NU, F, dose=1.5, step=1.0, seed=777)
    r = _replay(frames)
    o = Oracle("f64").align_movie(frames, dict(bin=BIN, fod=FOD, box=BOX, iter=ITER,
                                               bfactor=BFT, nthreads=8), want_boxes=False)
    return r, o


def test_replay_shifts_and_sum_match_oracle(replay):
    r, o = replay
    assert r["n"] == o["n"] == NU // 2
    db = np.abs(r["pair_b"] - o["pair_b"])
    d = np.abs(r["frame_shift"] - o["frame_shift"])
    rel = np.linalg.norm(r["sum_image"] - o["sum_image"]) / np.linalg.norm(o["sum_image"])
    rraw = np.linalg.norm(r["rawsum"] - o["rawsum_image"]) / np.linalg.norm(o["rawsum_image"])
    print("level-0 replay: pair components differing %d (max %.3g px), frame shift max error "
          "%.3g px, sum rel L2 %.3g, raw sum rel L2 %.3g" %
          (int((db > 1e-6).sum()), db.max(), d.max(), rel, rraw), flush=True)
    assert int((d > 0.01).sum()) == 0
    assert rel < 1e-4 and rraw < 1e-5


def test_log_modulus_and_display_crop(replay):
    r, _ = replay
    n = r["n"]
    spec = r["rawspec"]
    ref = np.log(1 + np.abs(spec.astype(np.complex128)) * FFTSCALE)
    assert np.abs(r["logmod"] - ref).max() <= 1e-6 * max(1.0, np.abs(ref).max())
    # GPUMemBinD2H (cufunc.cu:147-160): rows 0..199 and the last 200 rows,
    # columns 0..200 of the sum spectrum
    h = r["hdisp"].view(np.complex64)
    ss = r["sumspec"]
    assert np.array_equal(h[:DISPDIM // 2], ss[:DISPDIM // 2, :DISPDIM // 2 + 1])
    assert np.array_equal(h[DISPDIM // 2:], ss[n - DISPDIM // 2:, :DISPDIM // 2 + 1])
