"""ctypes binding of the CPU oracle (oracle/liboracle_f64.so, liboracle_f32.so).

Test infrastructure: the oracle is the checker.  Only tests, the smoke check
and the cpu_baseline leg of bench.py load it.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# MC_SANITIZED=1: the AddressSanitizer/UBSan build (oracle/_san, `make -C
# oracle sanitize`; tests/test_sanitize.py sets it)
ORACLE_DIR = os.path.join(ROOT, "oracle", "_san" if os.environ.get("MC_SANITIZED") else "")


class OrParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("crop_offsetx", ctypes.c_int), ("crop_offsety", ctypes.c_int),
        ("crop_nsam", ctypes.c_int), ("bin", ctypes.c_float), ("pad", ctypes.c_int),
        ("nStart", ctypes.c_int), ("nEnd", ctypes.c_int),
        ("nStartSum", ctypes.c_int), ("nEndSum", ctypes.c_int), ("nEndSum2", ctypes.c_int),
        ("bfactor", ctypes.c_float), ("box", ctypes.c_int), ("fod", ctypes.c_int),
        ("nps", ctypes.c_int), ("kit", ctypes.c_float), ("atm", ctypes.c_int),
        ("iter", ctypes.c_int), ("iterthres", ctypes.c_float),
        ("full_c2r", ctypes.c_int), ("nthreads", ctypes.c_int), ("fftscale", ctypes.c_float),
    ]


def _result_struct(real_t):
    P = ctypes.POINTER
    dp = P(ctypes.c_double)
    ip = P(ctypes.c_int)
    rp = P(real_t)

    class OrResult(ctypes.Structure):
        _fields_ = [
            ("n", ctypes.c_int), ("nsamUnbin", ctypes.c_int), ("nframe", ctypes.c_int),
            ("npairs", ctypes.c_int), ("ncols", ctypes.c_int), ("ngood", ctypes.c_int),
            ("niter", ctypes.c_int),
            ("pairs", ip), ("pair_b", dp), ("pair_cc", dp), ("boxmaps", rp),
            ("shift1", dp), ("ki1", dp), ("good", ip), ("shift2", dp), ("ki2", dp),
            ("iter_meas", dp), ("rmsd", dp), ("final_shift", dp), ("frame_shift", dp),
            ("sum_image", rp), ("sum2_image", rp), ("rawsum_image", rp), ("spectra", rp),
            ("sum2_written", ctypes.c_int), ("rawsum_written", ctypes.c_int),
            ("fsc", dp), ("disp_raw_fft", P(ctypes.c_short)),
            ("disp_corr_sum", P(ctypes.c_short)), ("disp_corr_fft", P(ctypes.c_short)),
            ("disp_written", ctypes.c_int),
        ]
    return OrResult


class Oracle:
    """One precision build of the oracle."""

    def __init__(self, precision="f64"):
        path = os.path.join(ORACLE_DIR, "liboracle_%s.so" % precision)
        if not os.path.exists(path):
            raise FileNotFoundError(path + " (run oracle/Makefile or __graft_entry__.build())")
        self.lib = ctypes.CDLL(path)
        self.real = np.float64 if precision == "f64" else np.float32
        self.c_real = ctypes.c_double if precision == "f64" else ctypes.c_float
        self.Result = _result_struct(self.c_real)
        L = self.lib
        vp = ctypes.c_void_p
        L.or_fft_c2c.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.or_r2c2d.argtypes = [vp, ctypes.c_int]
        L.or_c2r2d.argtypes = [vp, ctypes.c_int]
        L.or_verify_crop_size.argtypes = [ctypes.c_int] * 5 + [ctypes.c_float]
        L.or_binned_size.argtypes = [ctypes.c_int, ctypes.c_float]
        L.or_weight.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int]
        L.or_weight.restype = self.c_real
        L.or_cross_power.argtypes = [vp, vp, ctypes.c_int, ctypes.c_float]
        L.or_phase_shift.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.or_find_shift.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, vp, vp, ctypes.c_int]
        L.or_find_shift.restype = ctypes.c_double
        L.or_find_shift_box.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_int]
        L.or_find_shift_box.restype = ctypes.c_double
        L.or_pruned_box.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
        L.or_bin_crop.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
        L.or_pad2fft.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int]
        L.or_pair_list.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int]
        L.or_lsq.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp]
        L.or_residual.argtypes = [vp, ctypes.c_int, vp, vp, vp]
        L.or_threshold.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_double, vp]
        L.or_align_movie.argtypes = [vp, vp, ctypes.POINTER(OrParams), ctypes.POINTER(self.Result)]
        L.or_sharpness.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, vp, vp]
        L.or_sharpness.restype = ctypes.c_double
        L.or_fsc_bins.argtypes = [ctypes.c_int, vp]
        L.or_fsc.argtypes = [vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, vp]
        L.or_logmod_image.argtypes = [vp, ctypes.c_int, ctypes.c_float, vp]
        L.or_fft_display.argtypes = [vp, ctypes.c_int, ctypes.c_float, ctypes.c_int, vp]
        L.or_sum_display.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
        L.or_disp_short.argtypes = [vp, ctypes.c_int, vp]

    # -- FFT -------------------------------------------------------------
    def fft_c2c(self, x, sign=-1):
        a = np.ascontiguousarray(np.stack([x.real, x.imag], -1).astype(self.real))
        self.lib.or_fft_c2c(a.ctypes.data, len(x), sign)
        return a[..., 0] + 1j * a[..., 1]

    def r2c2d(self, img):
        n = img.shape[0]
        buf = np.zeros((n, n + 2), self.real)
        buf[:, :n] = img
        self.lib.or_r2c2d(buf.ctypes.data, n)
        return buf.view(np.complex128 if self.real == np.float64 else np.complex64)

    def c2r2d(self, spec):
        n = spec.shape[0]
        buf = np.ascontiguousarray(spec.astype(np.complex128 if self.real == np.float64 else np.complex64))
        b = buf.view(self.real).reshape(n, n + 2).copy()
        self.lib.or_c2r2d(b.ctypes.data, n)
        return b[:, :n].copy()

    def padded(self, spec):
        """complex (n, n/2+1) -> real padded layout (n, n+2) in this precision"""
        ct = np.complex128 if self.real == np.float64 else np.complex64
        return np.ascontiguousarray(spec.astype(ct)).view(self.real).reshape(spec.shape[0], -1).copy()

    def weight(self, n, bfactor, kx, ky):
        return self.lib.or_weight(n, bfactor, kx, ky)

    def cross_power(self, a, b, n, bfactor):
        """dsum=a (first), dfft=b (conjugated); complex (n, n/2+1) in/out"""
        A = self.padded(a)
        B = self.padded(b)
        self.lib.or_cross_power(B.ctypes.data, A.ctypes.data, n, bfactor)
        return B.view(np.complex128 if self.real == np.float64 else np.complex64)

    def phase_shift(self, spec, sx, sy):
        n = spec.shape[0]
        B = self.padded(spec)
        self.lib.or_phase_shift(B.ctypes.data, n, sx, sy)
        return B.view(np.complex128 if self.real == np.float64 else np.complex64)

    def find_shift_box(self, boxmap, n, wnoise=-1):
        box = boxmap.shape[0]
        bm = np.ascontiguousarray(boxmap, self.real)
        sx = ctypes.c_double()
        sy = ctypes.c_double()
        cc = self.lib.or_find_shift_box(bm.ctypes.data, n, box, ctypes.addressof(sx), ctypes.addressof(sy), wnoise)
        return sx.value, sy.value, cc

    def find_shift(self, cmap, box, wnoise=-1):
        n = cmap.shape[0]
        buf = np.zeros((n, n + 2), self.real)
        buf[:, :n] = cmap
        bm = np.zeros((box, box), self.real)
        sx = ctypes.c_double()
        sy = ctypes.c_double()
        cc = self.lib.or_find_shift(buf.ctypes.data, n, bm.ctypes.data, box,
                                    ctypes.addressof(sx), ctypes.addressof(sy), wnoise)
        return sx.value, sy.value, cc, bm

    def pruned_box(self, X, box):
        n = X.shape[0]
        B = self.padded(X)
        out = np.zeros((box, box), self.real)
        self.lib.or_pruned_box(B.ctypes.data, n, box, out.ctypes.data)
        return out

    def pair_list(self, nframe, fod):
        c = self.lib.or_pair_list(nframe, fod, None, 0)
        p = np.zeros((c, 2), np.int32)
        self.lib.or_pair_list(nframe, fod, p.ctypes.data, c)
        return p

    def lsq(self, pairs, nframe, b):
        pairs = np.ascontiguousarray(pairs, np.int32)
        b = np.ascontiguousarray(b, np.float64)
        s = np.zeros((nframe - 1, 2), np.float64)
        rc = self.lib.or_lsq(pairs.ctypes.data, len(pairs), nframe, b.ctypes.data, s.ctypes.data)
        if rc:
            raise RuntimeError("or_lsq rc=%d" % rc)
        return s

    def residual(self, pairs, s, b):
        pairs = np.ascontiguousarray(pairs, np.int32)
        ki = np.zeros(len(pairs))
        self.lib.or_residual(pairs.ctypes.data, len(pairs), np.ascontiguousarray(s, np.float64).ctypes.data,
                             np.ascontiguousarray(b, np.float64).ctypes.data, ki.ctypes.data)
        return ki

    def threshold(self, pairs, nframe, ki, kit):
        pairs = np.ascontiguousarray(pairs, np.int32)
        good = np.zeros(len(pairs), np.int32)
        ng = self.lib.or_threshold(pairs.ctypes.data, len(pairs), nframe,
                                   np.ascontiguousarray(ki, np.float64).ctypes.data, kit, good.ctypes.data)
        return good[:ng].copy()

    def sharpness(self, spectra_padded, n, bfactor, s, want_grad=True):
        sp = np.ascontiguousarray(spectra_padded, self.real)
        nf = sp.shape[0]
        s = np.ascontiguousarray(s, np.float64)
        g = np.zeros(2 * nf)
        J = self.lib.or_sharpness(sp.ctypes.data, nf, n, bfactor, s.ctypes.data,
                                  g.ctypes.data if want_grad else None)
        return J, g.reshape(nf, 2)

    def fsc_bins(self, n):
        st = ctypes.c_int()
        nb = self.lib.or_fsc_bins(n, ctypes.addressof(st))
        return nb, st.value

    def fsc(self, a, b, direction):
        """a, b: complex (n, n/2+1) half spectra.  Returns (nbox, 2): columns
        parallel (D) and perpendicular (U) to `direction`."""
        n = a.shape[0]
        A = self.padded(a)
        B = self.padded(b)
        nb, _ = self.fsc_bins(n)
        out = np.zeros((nb, 2))
        rc = self.lib.or_fsc(A.ctypes.data, B.ctypes.data, n, float(direction[0]),
                             float(direction[1]), out.ctypes.data)
        if rc:
            raise RuntimeError("or_fsc rc=%d" % rc)
        return out

    # -- quick-look images --------------------------------------------------
    def logmod_image(self, spec, scale=1e-4):
        n = spec.shape[0]
        S = self.padded(spec)
        img = np.zeros((n, n), self.real)
        if self.lib.or_logmod_image(S.ctypes.data, n, scale, img.ctypes.data):
            raise RuntimeError("or_logmod_image")
        return img

    def fft_display(self, spec, scale=1e-4, d=400):
        n = spec.shape[0]
        S = self.padded(spec)
        out = np.zeros((d, d), self.real)
        if self.lib.or_fft_display(S.ctypes.data, n, scale, d, out.ctypes.data):
            raise RuntimeError("or_fft_display")
        return out

    def sum_display(self, spec, d=400):
        n = spec.shape[0]
        S = self.padded(spec)
        out = np.zeros((d, d), self.real)
        if self.lib.or_sum_display(S.ctypes.data, n, d, out.ctypes.data):
            raise RuntimeError("or_sum_display")
        return out

    def disp_short(self, buf):
        """Returns the 8-bit display (int16) or None when the clipped range is
        empty (the reference leaves the image unwritten)."""
        b = np.ascontiguousarray(buf, np.float32)
        out = np.zeros(b.shape, np.int16)
        rc = self.lib.or_disp_short(b.ctypes.data, b.size, out.ctypes.data)
        return None if rc else out

    # -- whole movie -------------------------------------------------------
    def align_movie(self, frames, params, frames2=None, want_spectra=False, want_boxes=True,
                    want_disp=False):
        """frames: (nz, ny, nx) float32.  params: dict of OrParams fields."""
        frames = np.ascontiguousarray(frames, np.float32)
        nz, ny, nx = frames.shape
        p = OrParams()
        defaults = dict(nx=nx, ny=ny, nz=nz, crop_offsetx=0, crop_offsety=0, crop_nsam=0,
                        bin=1.0, pad=0, nStart=0, nEnd=nz - 1, nStartSum=0, nEndSum=nz - 1,
                        nEndSum2=0, bfactor=150.0, box=96, fod=2, nps=0, kit=1.0, atm=1,
                        iter=0, iterthres=0.0, full_c2r=1, nthreads=0, fftscale=0.0)
        defaults.update(params)
        for k, v in defaults.items():
            setattr(p, k, v)
        # sizes
        if p.pad > 0:
            nsu = p.pad
        else:
            nsu = self.lib.or_verify_crop_size(nx, ny, p.crop_offsetx, p.crop_offsety, p.crop_nsam, p.bin)
        n = self.lib.or_binned_size(nsu, p.bin)
        nframe = p.nEnd - p.nStart + 1
        npairs = self.lib.or_pair_list(nframe, p.fod, None, 0)
        K = nframe - 1
        box = p.box
        out = dict(
            pairs=np.zeros((npairs, 2), np.int32), pair_b=np.zeros((npairs, 2)),
            pair_cc=np.zeros(npairs), shift1=np.zeros((K, 2)), ki1=np.zeros(npairs),
            good=np.zeros(npairs, np.int32), shift2=np.zeros((K, 2)), ki2=np.zeros(npairs),
            iter_meas=np.zeros((max(p.iter, 1), nframe, 2)), rmsd=np.zeros(max(p.iter, 1)),
            final_shift=np.zeros((K, 2)), frame_shift=np.zeros((nframe, 2)),
            sum_image=np.zeros((n, n), self.real), sum2_image=np.zeros((n, n), self.real),
            rawsum_image=np.zeros((n, n), self.real),
            fsc=np.zeros((self.lib.or_fsc_bins(n, None), 4)),
        )
        if want_boxes:
            out["boxmaps"] = np.zeros((npairs, box, box), self.real)
        if want_spectra:
            out["spectra"] = np.zeros((nframe, n, n + 2), self.real)
        if want_disp:
            for k in ("disp_raw_fft", "disp_corr_sum", "disp_corr_fft"):
                out[k] = np.zeros((400, 400), np.int16)
        r = self.Result()
        for k, v in out.items():
            setattr(r, k, v.ctypes.data_as(type(getattr(r, k))))
        f2 = None
        if frames2 is not None:
            frames2 = np.ascontiguousarray(frames2, np.float32)
            f2 = frames2.ctypes.data
        rc = self.lib.or_align_movie(frames.ctypes.data, f2, ctypes.byref(p), ctypes.byref(r))
        if rc:
            raise RuntimeError("or_align_movie rc=%d" % rc)
        out["n"] = r.n
        out["nsamUnbin"] = r.nsamUnbin
        out["nframe"] = r.nframe
        out["npairs"] = r.npairs
        out["ngood"] = r.ngood
        out["niter"] = r.niter
        out["good"] = out["good"][: r.ngood]
        out["ki2"] = out["ki2"][: r.ngood]
        out["iter_meas"] = out["iter_meas"][: r.niter]
        out["rmsd"] = out["rmsd"][: r.niter]
        out["sum2_written"] = r.sum2_written
        out["rawsum_written"] = r.rawsum_written
        out["disp_written"] = r.disp_written
        return out
