"""MI355X-native drift correction of dose-fractionated movies.

Python binding of the C ABI in include/mc_abi.h (libmc.so, HIP kernels for
gfx950).  The host driver and CLI are C++ (motioncorr1exp_amd/host); this
module exists for tests and benchmarks.  Importing it does not touch the
GPU; every call that needs the GPU raises if libmc.so is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmc.so")

MC_OK = 0
MC_DTYPE = {np.dtype(np.int8): 0, np.dtype(np.int16): 1, np.dtype(np.float32): 2,
            np.dtype(np.uint16): 6, np.dtype(np.uint8): 100}
KT = dict(spectra=0, pair_cols=1, pair_rows=2, peaks=3, sum=4, image=5, rounds=6)


class MCError(RuntimeError):
    pass


class Corrections(ctypes.Structure):
    _fields_ = [("dark", ctypes.c_void_p), ("gain", ctypes.c_void_p),
                ("dark2", ctypes.c_void_p), ("gain2", ctypes.c_void_p),
                ("xscale", ctypes.c_float), ("yscale", ctypes.c_float),
                ("second", ctypes.c_int)]


class MovieDesc(ctypes.Structure):
    _fields_ = [("frames", ctypes.c_void_p), ("frames2", ctypes.c_void_p),
                ("dtype", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("nz", ctypes.c_int), ("row_pitch", ctypes.c_size_t),
                ("frame_pitch", ctypes.c_size_t), ("nz_file", ctypes.c_int),
                ("first_frame", ctypes.c_int), ("corr", ctypes.POINTER(Corrections))]


def corrections(dark=None, gain=None, dark2=None, gain2=None, xscale=1.0, yscale=1.0,
                second=False):
    """mc_corrections from DeviceArrays of nx*ny float32 references (None:
    absent); xscale, yscale are 1/-scx and 1/-scy."""
    c = Corrections()
    c.dark = dark.ptr if dark is not None else None
    c.gain = gain.ptr if gain is not None else None
    c.dark2 = dark2.ptr if dark2 is not None else None
    c.gain2 = gain2.ptr if gain2 is not None else None
    c.xscale, c.yscale = xscale, yscale
    c.second = 1 if second else 0
    return c


class Params(ctypes.Structure):
    _fields_ = [("crop_offsetx", ctypes.c_int), ("crop_offsety", ctypes.c_int),
                ("crop_nsam", ctypes.c_int), ("bin", ctypes.c_float),
                ("nStart", ctypes.c_int), ("nEnd", ctypes.c_int),
                ("nStartSum", ctypes.c_int), ("nEndSum", ctypes.c_int),
                ("nEndSum2", ctypes.c_int), ("bfactor", ctypes.c_float),
                ("CCPeakSearchDim", ctypes.c_int), ("FrameDistOffset", ctypes.c_int),
                ("NoisePeakSize", ctypes.c_int), ("kiThresh", ctypes.c_float),
                ("bAlignToMid", ctypes.c_int), ("iter", ctypes.c_int),
                ("iterthres", ctypes.c_float), ("pad", ctypes.c_int),
                ("precision", ctypes.c_int), ("grad_steps", ctypes.c_int),
                ("grad_step_px", ctypes.c_float), ("fftscale", ctypes.c_float)]


_P = ctypes.POINTER


class Outputs(ctypes.Structure):
    _fields_ = [("pairs", _P(ctypes.c_int)), ("pair_b", _P(ctypes.c_double)),
                ("pair_cc", _P(ctypes.c_double)), ("boxmaps", _P(ctypes.c_float)),
                ("shift1", _P(ctypes.c_double)), ("ki1", _P(ctypes.c_double)),
                ("good", _P(ctypes.c_int)), ("shift2", _P(ctypes.c_double)),
                ("ki2", _P(ctypes.c_double)), ("iter_meas", _P(ctypes.c_double)),
                ("iter_cc", _P(ctypes.c_double)), ("iter_boxmaps", _P(ctypes.c_float)),
                ("rmsd", _P(ctypes.c_double)), ("final_shift", _P(ctypes.c_double)),
                ("frame_shift", _P(ctypes.c_double)), ("frame_mean", _P(ctypes.c_double)),
                ("sum_image", _P(ctypes.c_float)), ("sum2_image", _P(ctypes.c_float)),
                ("rawsum_image", _P(ctypes.c_float)), ("sum_image_dev", ctypes.c_void_p),
                ("fsc", _P(ctypes.c_double)), ("disp_raw_fft", _P(ctypes.c_short)),
                ("disp_corr_sum", _P(ctypes.c_short)), ("disp_corr_fft", _P(ctypes.c_short))]

DISPDIM = 400  # MC_DISPDIM


class Report(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nsamUnbin", ctypes.c_int), ("nframe", ctypes.c_int),
                ("npairs", ctypes.c_int), ("ngood", ctypes.c_int), ("niter", ctypes.c_int),
                ("converged", ctypes.c_int), ("sum2_written", ctypes.c_int),
                ("rawsum_written", ctypes.c_int), ("rmsd", ctypes.c_double),
                ("ms_total", ctypes.c_float), ("ms_spectra", ctypes.c_float),
                ("ms_pairs", ctypes.c_float), ("ms_peaks", ctypes.c_float),
                ("ms_solve", ctypes.c_float), ("ms_iter", ctypes.c_float),
                ("ms_sum", ctypes.c_float), ("sharp_j0", ctypes.c_double),
                ("sharp_j1", ctypes.c_double), ("grad_accepted", ctypes.c_int)]


class BatchItem(ctypes.Structure):
    _fields_ = [("movie", _P(MovieDesc)), ("params", _P(Params)), ("out", _P(Outputs)),
                ("rep", _P(Report))]


class BatchReport(ctypes.Structure):
    _fields_ = [("groups", ctypes.c_int), ("sub_batches", ctypes.c_int),
                ("graph_captures", ctypes.c_int), ("graph_replays", ctypes.c_int),
                ("ms_total", ctypes.c_double)]


class TuneReport(ctypes.Structure):
    _fields_ = [("r2c_variant", ctypes.c_int), ("chunk_frames", ctypes.c_int),
                ("pair_cols_variant", ctypes.c_int), ("pair_rows_variant", ctypes.c_int),
                ("ms_stage1", ((ctypes.c_float * 4) * 4) * 2), ("ms_pair_cols", ctypes.c_float * 4),
                ("ms_pair_rows", ctypes.c_float * 4), ("cols_plan", ctypes.c_int),
                ("ms_cols_plan", ctypes.c_float * 6)]


class MaskElem(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int), ("z", ctypes.c_float)]


# Every symbol declared in include/mc_abi.h.
ABI_SYMBOLS = [
    "mc_last_error", "mc_version", "mc_init_gpu", "mc_reset_gpu", "mc_device_count",
    "mc_mem_check", "mc_mem_alloc", "mc_mem_zero", "mc_mem_free", "mc_mem_h2d", "mc_mem_d2h",
    "mc_mem_d2d", "mc_mem_bin_d2d", "mc_mem_bin_d2h", "mc_fft_plan", "mc_ifft_plan",
    "mc_fft2d", "mc_ifft2d", "mc_fft_destroy", "mc_sync", "mc_add", "mc_multiply_num",
    "mc_fft_log_modulus", "mc_make_pos_list", "mc_shift_cc", "mc_shift", "mc_find_shift",
    "mc_find_shift_noalloc",
    "mc_context_create", "mc_context_destroy", "mc_align_movie", "mc_load_frames",
    "mc_set_spectra", "mc_get_spectra", "mc_pair_shifts", "mc_find_peaks", "mc_sum_image",
    "mc_pair_list", "mc_lsq", "mc_threshold", "mc_kernel_timing", "mc_kernel_timing_groups",
    "mc_kernel_timing_get",
    "mc_sharpness", "mc_align_batch", "mc_autotune", "mc_tune_load", "mc_fsc_bins",
    "mc_fsc_table", "mc_disp_short",
]

_lib = None


def lib():
    """Load libmc.so (no GPU call happens at load time)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MCError("libmc.so not built: run __graft_entry__.build() (%s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, i, f, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_size_t
        L.mc_last_error.restype = ctypes.c_char_p
        L.mc_version.restype = ctypes.c_char_p
        L.mc_mem_alloc.argtypes = [_P(vp), sz]
        L.mc_mem_zero.argtypes = [vp, sz]
        L.mc_mem_free.argtypes = [vp]
        L.mc_mem_h2d.argtypes = [vp, vp, sz]
        L.mc_mem_d2h.argtypes = [vp, vp, sz]
        L.mc_mem_d2d.argtypes = [vp, vp, sz]
        L.mc_mem_check.argtypes = [_P(sz), _P(sz)]
        L.mc_mem_bin_d2d.argtypes = [vp, vp, i, i]
        L.mc_mem_bin_d2h.argtypes = [vp, vp, i, i]
        L.mc_fft_plan.argtypes = [_P(vp), i]
        L.mc_ifft_plan.argtypes = [_P(vp), i]
        L.mc_fft2d.argtypes = [vp, vp]
        L.mc_ifft2d.argtypes = [vp, vp]
        L.mc_fft_destroy.argtypes = [vp]
        L.mc_add.argtypes = [vp, vp, sz]
        L.mc_multiply_num.argtypes = [vp, f, sz]
        L.mc_fft_log_modulus.argtypes = [vp, vp, i, f]
        L.mc_make_pos_list.argtypes = [vp, i, f]
        L.mc_shift_cc.argtypes = [vp, vp, vp, f, f, i]
        L.mc_shift.argtypes = [vp, vp, f, f, i]
        L.mc_find_shift.argtypes = [vp, i, vp, i, _P(f), _P(f), _P(f), i]
        L.mc_find_shift_noalloc.argtypes = [vp, i, vp, i, _P(f), _P(f), _P(f), i]
        L.mc_context_create.argtypes = [i, _P(vp)]
        L.mc_context_destroy.argtypes = [vp]
        L.mc_align_movie.argtypes = [vp, _P(MovieDesc), _P(Params), _P(Outputs), _P(Report)]
        L.mc_load_frames.argtypes = [vp, _P(MovieDesc), _P(Params), _P(Report)]
        L.mc_set_spectra.argtypes = [vp, vp, i, i]
        L.mc_get_spectra.argtypes = [vp, vp]
        L.mc_pair_shifts.argtypes = [vp, vp, i, f, i, i, vp, vp, vp]
        L.mc_find_peaks.argtypes = [vp, vp, i, i, i, i, vp, vp]
        L.mc_sum_image.argtypes = [vp, vp, vp]
        L.mc_pair_list.argtypes = [i, i, vp, i]
        L.mc_lsq.argtypes = [vp, i, i, vp, vp]
        L.mc_threshold.argtypes = [vp, i, i, vp, d, vp]
        L.mc_kernel_timing.argtypes = [vp, i]
        L.mc_kernel_timing_groups.argtypes = [vp, ctypes.c_uint]
        L.mc_kernel_timing_get.argtypes = [vp, i, _P(i), _P(d)]
        L.mc_sharpness.argtypes = [vp, vp, f, _P(d), vp]
        L.mc_align_batch.argtypes = [vp, _P(BatchItem), i, _P(BatchReport)]
        L.mc_autotune.argtypes = [vp, _P(MovieDesc), _P(Params), ctypes.c_char_p, _P(TuneReport)]
        L.mc_tune_load.argtypes = [vp, ctypes.c_char_p, _P(i)]
        L.mc_fsc_bins.argtypes = [i]
        L.mc_fsc_table.argtypes = [vp, vp, vp, i, d, d, vp]
        L.mc_disp_short.argtypes = [vp, vp, i, vp]
        _lib = L
    return _lib


def check(rc):
    if rc != MC_OK:
        raise MCError("libmc error %d: %s" % (rc, lib().mc_last_error().decode(errors="replace")))
    return rc


def check_count(rc):
    """A count returned by libmc, or MCError for a negative error code."""
    if rc < 0:
        check(rc)
    return rc


def _ptr(a, ct):
    return a.ctypes.data_as(_P(ct)) if a is not None else None


class DeviceArray:
    """A device allocation holding a copy of a host array (libmc memory API)."""

    def __init__(self, host=None, nbytes=None):
        L = lib()
        self.nbytes = host.nbytes if host is not None else int(nbytes)
        p = ctypes.c_void_p()
        check(L.mc_mem_alloc(ctypes.byref(p), self.nbytes))
        self.ptr = p.value
        if host is not None:
            h = np.ascontiguousarray(host)
            check(L.mc_mem_h2d(self.ptr, h.ctypes.data, self.nbytes))

    def to_host(self, dtype, shape):
        out = np.empty(shape, dtype)
        check(lib().mc_mem_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().mc_mem_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def default_params(nframe, **kw):
    """APARA defaults (reference dosefgpu_driftcorr.cpp:8-58) clamped for a
    stack of nframe frames, as VerifyPara does (:60-118)."""
    p = dict(crop_offsetx=0, crop_offsety=0, crop_nsam=0, bin=1.0, nStart=0, nEnd=0,
             nStartSum=0, nEndSum=0, nEndSum2=0, bfactor=150.0, CCPeakSearchDim=96,
             FrameDistOffset=2, NoisePeakSize=0, kiThresh=1.0, bAlignToMid=1, iter=0,
             iterthres=0.0, pad=0, precision=0, grad_steps=0, grad_step_px=0.0,
             fftscale=0.0001)
    p.update(kw)
    if p["nStart"] < 0:
        p["nStart"] = 0
    if p["nEnd"] >= nframe or p["nEnd"] <= p["nStart"]:
        p["nEnd"] = nframe - 1
    if p["nStartSum"] < p["nStart"]:
        p["nStartSum"] = p["nStart"]
    if p["nEndSum"] > p["nEnd"] or p["nEndSum"] <= p["nStartSum"]:
        p["nEndSum"] = p["nEnd"]
    if p["nEndSum2"] > p["nEnd"]:
        p["nEndSum2"] = p["nEnd"]
    if p["nEndSum2"] <= p["nStartSum"]:
        p["nEndSum2"] = 0
    return p


def to_params(d):
    P = Params()
    for k, v in d.items():
        setattr(P, k, v)
    return P


class Context:
    """A libmc context (one HIP stream + resident buffers) on one device."""

    def __init__(self, device=0):
        self.L = lib()
        h = ctypes.c_void_p()
        check(self.L.mc_context_create(device, ctypes.byref(h)))
        self.h = h.value

    def close(self):
        if self.h:
            self.L.mc_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def movie_desc(dev_frames, shape, dtype, dev_frames2=None, nz_file=0, corr=None):
        nz, ny, nx = shape
        m = MovieDesc()
        if corr is not None:
            m.corr = ctypes.pointer(corr)
        m.frames = dev_frames.ptr
        m.frames2 = dev_frames2.ptr if dev_frames2 is not None else None
        m.dtype = MC_DTYPE[np.dtype(dtype)]
        m.nx, m.ny, m.nz = nx, ny, nz
        m.row_pitch = 0
        m.frame_pitch = 0
        m.nz_file = nz_file
        m.first_frame = 0
        return m

    def _outputs(self, shape, p, want, sum_dev):
        F = p.nEnd - p.nStart + 1
        K = F - 1
        P = self.L.mc_pair_list(F, p.FrameDistOffset, None, 0)
        B = p.CCPeakSearchDim
        n = _binned_n(shape, p)
        res = {}
        sizes = dict(pairs=((P, 2), np.int32), pair_b=((P, 2), np.float64),
                     pair_cc=((P,), np.float64), boxmaps=((P, B, B), np.float32),
                     shift1=((K, 2), np.float64), ki1=((P,), np.float64),
                     good=((P,), np.int32), shift2=((K, 2), np.float64), ki2=((P,), np.float64),
                     iter_meas=((max(p.iter, 1), F, 2), np.float64),
                     iter_cc=((max(p.iter, 1), F), np.float64),
                     iter_boxmaps=((max(p.iter, 1), F, B, B), np.float32),
                     rmsd=((max(p.iter, 1),), np.float64), final_shift=((K, 2), np.float64),
                     frame_shift=((F, 2), np.float64), frame_mean=((F,), np.float64),
                     sum_image=((n, n), np.float32), sum2_image=((n, n), np.float32),
                     rawsum_image=((n, n), np.float32),
                     fsc=((max(self.L.mc_fsc_bins(n), 0), 4), np.float64),
                     disp_raw_fft=((DISPDIM, DISPDIM), np.int16),
                     disp_corr_sum=((DISPDIM, DISPDIM), np.int16),
                     disp_corr_fft=((DISPDIM, DISPDIM), np.int16))
        o = Outputs()
        for k in want:
            shp, dt = sizes[k]
            res[k] = np.zeros(shp, dt)
            ct = {np.int32: ctypes.c_int, np.int16: ctypes.c_short,
                  np.float32: ctypes.c_float}.get(dt, ctypes.c_double)
            setattr(o, k, _ptr(res[k], ct))
        if sum_dev is not None:
            o.sum_image_dev = sum_dev.ptr
        return o, res

    @staticmethod
    def _finish(res, r):
        res["report"] = {k: getattr(r, k) for k, _ in Report._fields_}
        if "good" in res:
            res["good"] = res["good"][: r.ngood]
        if "ki2" in res:
            res["ki2"] = res["ki2"][: r.ngood]
        for k in ("iter_meas", "iter_cc", "iter_boxmaps", "rmsd"):
            if k in res:
                res[k] = res[k][: r.niter]
        return res

    def align_movie(self, dev_frames, shape, dtype, params, dev_frames2=None, want=(),
                    nz_file=0, sum_dev=None, corr=None):
        """Runs mc_align_movie.  `want` names optional outputs to return;
        corr: an mc_corrections (see corrections())."""
        p = to_params(params)
        m = self.movie_desc(dev_frames, shape, dtype, dev_frames2, nz_file, corr)
        o, res = self._outputs(shape, p, want, sum_dev)
        r = Report()
        check(self.L.mc_align_movie(self.h, ctypes.byref(m), ctypes.byref(p), ctypes.byref(o),
                                    ctypes.byref(r)))
        return self._finish(res, r)

    def align_batch(self, movies):
        """Runs mc_align_batch over a list of dicts with keys dev, shape,
        dtype, params and optionally dev2, want, nz_file, sum_dev, corr.  Returns
        (per-movie results as align_movie returns them, batch report)."""
        n = len(movies)
        items = (BatchItem * n)()
        keep, results = [], []
        for i, mv in enumerate(movies):
            p = to_params(mv["params"])
            m = self.movie_desc(mv["dev"], mv["shape"], mv["dtype"], mv.get("dev2"),
                                mv.get("nz_file", 0), mv.get("corr"))
            o, res = self._outputs(mv["shape"], p, mv.get("want", ()), mv.get("sum_dev"))
            r = Report()
            keep.append((p, m, o, r))
            results.append(res)
            items[i].movie = ctypes.pointer(m)
            items[i].params = ctypes.pointer(p)
            items[i].out = ctypes.pointer(o)
            items[i].rep = ctypes.pointer(r)
        br = BatchReport()
        check(self.L.mc_align_batch(self.h, items, n, ctypes.byref(br)))
        out = [self._finish(res, k[3]) for res, k in zip(results, keep)]
        return out, {k: getattr(br, k) for k, _ in BatchReport._fields_}

    def autotune(self, dev_frames, shape, dtype, params, cache_path=None):
        """Runs mc_autotune on this movie; returns the report as a dict."""
        p = to_params(params)
        m = self.movie_desc(dev_frames, shape, dtype)
        r = TuneReport()
        check(self.L.mc_autotune(self.h, ctypes.byref(m), ctypes.byref(p),
                                 cache_path.encode() if cache_path else None, ctypes.byref(r)))
        return {"r2c_variant": r.r2c_variant, "chunk_frames": r.chunk_frames,
                "pair_cols_variant": r.pair_cols_variant,
                "pair_rows_variant": r.pair_rows_variant,
                "ms_stage1": [[list(row) for row in blk] for blk in r.ms_stage1],
                "ms_pair_cols": list(r.ms_pair_cols), "ms_pair_rows": list(r.ms_pair_rows),
                "cols_plan": r.cols_plan, "ms_cols_plan": list(r.ms_cols_plan)}

    def tune_load(self, cache_path):
        n = ctypes.c_int()
        check(self.L.mc_tune_load(self.h, cache_path.encode(), ctypes.byref(n)))
        return n.value

    def load_frames(self, dev_frames, shape, dtype, params, dev_frames2=None, corr=None):
        p = to_params(params)
        m = self.movie_desc(dev_frames, shape, dtype, dev_frames2, corr=corr)
        r = Report()
        check(self.L.mc_load_frames(self.h, ctypes.byref(m), ctypes.byref(p), ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in Report._fields_}

    def set_spectra(self, padded):
        padded = np.ascontiguousarray(padded, np.float32)
        nf, n, _ = padded.shape
        check(self.L.mc_set_spectra(self.h, padded.ctypes.data, n, nf))

    def get_spectra(self, nframe, n):
        out = np.zeros((nframe, n, n + 2), np.float32)
        check(self.L.mc_get_spectra(self.h, out.ctypes.data))
        return out

    def pair_shifts(self, pairs, bfactor, box, wnoise=-1, want_boxes=False):
        pairs = np.ascontiguousarray(pairs, np.int32)
        P = len(pairs)
        b = np.zeros((P, 2))
        cc = np.zeros(P)
        boxes = np.zeros((P, box, box), np.float32) if want_boxes else None
        check(self.L.mc_pair_shifts(self.h, pairs.ctypes.data, P, bfactor, box, wnoise,
                                    b.ctypes.data, cc.ctypes.data,
                                    boxes.ctypes.data if want_boxes else None))
        return b, cc, boxes

    def find_peaks(self, boxmaps, n, wnoise=-1):
        boxmaps = np.ascontiguousarray(boxmaps, np.float32)
        P, box, _ = boxmaps.shape
        b = np.zeros((P, 2))
        cc = np.zeros(P)
        check(self.L.mc_find_peaks(self.h, boxmaps.ctypes.data, P, box, n, wnoise,
                                   b.ctypes.data, cc.ctypes.data))
        return b, cc

    def sum_image(self, frame_shift, n):
        fs = np.ascontiguousarray(frame_shift, np.float64)
        img = np.zeros((n, n), np.float32)
        check(self.L.mc_sum_image(self.h, fs.ctypes.data, img.ctypes.data))
        return img

    def sharpness(self, frame_shift, bfactor):
        fs = np.ascontiguousarray(frame_shift, np.float64)
        J = ctypes.c_double()
        g = np.zeros_like(fs)
        check(self.L.mc_sharpness(self.h, fs.ctypes.data, bfactor, ctypes.byref(J), g.ctypes.data))
        return J.value, g

    def fsc_table(self, dev_a, dev_b, n, direction):
        """mc_fsc_table on two device half spectra in the row-major layout
        (n rows of n/2+1 complex).  Returns (nbox, 2): parallel (D) and
        perpendicular (U) columns."""
        nb = check_count(self.L.mc_fsc_bins(n))
        out = np.zeros((nb, 2))
        check(self.L.mc_fsc_table(self.h, dev_a.ptr, dev_b.ptr, n, float(direction[0]),
                                  float(direction[1]), out.ctypes.data))
        return out

    def disp_short(self, dev_buf, size):
        """mc_disp_short: 8-bit display (int16) of `size` device floats."""
        out = np.zeros(size, np.int16)
        check(self.L.mc_disp_short(self.h, dev_buf.ptr, size, out.ctypes.data))
        return out

    def kernel_timing(self, enable, groups=None):
        """Time every kernel group (enable) or none; with `groups` (names of
        KT), only those groups."""
        if groups is None:
            check(self.L.mc_kernel_timing(self.h, 1 if enable else 0))
        else:
            mask = sum(1 << KT[g] for g in groups) if enable else 0
            check(self.L.mc_kernel_timing_groups(self.h, mask))

    def kernel_time(self, group):
        c = ctypes.c_int()
        t = ctypes.c_double()
        check(self.L.mc_kernel_timing_get(self.h, KT[group], ctypes.byref(c), ctypes.byref(t)))
        return c.value, t.value


def _verify_crop_size(nx, ny, offx, offy, nsamout, bin_):
    """func.cpp:80-103 (host geometry, used only to size output arrays)."""
    if offx < 0 or offx >= nx or offy < 0 or offy >= ny or bin_ <= 0:
        return 0
    if nsamout <= 0:
        nsamout = min(nx - offx, ny - offy)
    nsamout = min(nsamout, nx - offx, ny - offy)
    if abs(bin_ - 1.0) < 0.0001:
        return nsamout // 2 * 2
    if abs(bin_ - 2.0) < 0.0001:
        return nsamout // 4 * 4
    return nsamout // 2 * 2


def _binned_n(shape, p):
    nz, ny, nx = shape
    nsu = p.pad if p.pad > 0 else _verify_crop_size(nx, ny, p.crop_offsetx, p.crop_offsety,
                                                    p.crop_nsam, p.bin)
    return int(np.float32(nsu) / np.float32(p.bin)) // 2 * 2


def pair_list(nframe, fod):
    L = lib()
    c = L.mc_pair_list(nframe, fod, None, 0)
    out = np.zeros((c, 2), np.int32)
    L.mc_pair_list(nframe, fod, out.ctypes.data, c)
    return out


def lsq(pairs, nframe, b):
    pairs = np.ascontiguousarray(pairs, np.int32)
    b = np.ascontiguousarray(b, np.float64)
    s = np.zeros((nframe - 1, 2))
    check(lib().mc_lsq(pairs.ctypes.data, len(pairs), nframe, b.ctypes.data, s.ctypes.data))
    return s


def threshold(pairs, nframe, ki, kit):
    pairs = np.ascontiguousarray(pairs, np.int32)
    ki = np.ascontiguousarray(ki, np.float64)
    good = np.zeros(len(pairs), np.int32)
    ng = lib().mc_threshold(pairs.ctypes.data, len(pairs), nframe, ki.ctypes.data, kit,
                            good.ctypes.data)
    return good[:ng].copy()
