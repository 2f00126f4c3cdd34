"""Deterministic synthetic dose-fractionated movies with known drift.

Phantom: a Gaussian random field with power spectrum 1/(1+(k/0.05)^2), k in
cycles/px, standardised to g; intensity I = dose * exp(0.5 g - 0.125) (mean
= dose).  Frame f = Poisson(I(x - t_f)) where t_f is applied by an exact
Fourier phase shift on a periodic canvas 1.25x the frame, then cropped.
Drift t_f is a cumulative random walk (t_0 = 0) with N(0, step^2) steps per
axis, optionally plus a quadratic term a*f^2 per axis.

The convention matches the alignment path: the content of frame f is
phi(x - t_f), so the measured shift of pair (i, j) is t_i - t_j and the
aligned sum registers every frame onto the reference frame.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import ctypes

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_synth_lib = None


def _poisson_lib():
    global _synth_lib
    if _synth_lib is None:
        path = os.path.join(_HERE, "libmc_synth.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path + " (run __graft_entry__.build())")
        lib = ctypes.CDLL(path)
        lib.mc_synth_poisson.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                         ctypes.c_uint64]
        _synth_lib = lib
    return _synth_lib

try:
    import scipy.fft as _fft

    def _rfft2(a, workers):
        return _fft.rfft2(a, workers=workers)

    def _irfft2(a, s, workers):
        return _fft.irfft2(a, s=s, workers=workers)
except ImportError:  # pragma: no cover
    def _rfft2(a, workers):
        return np.fft.rfft2(a)

    def _irfft2(a, s, workers):
        return np.fft.irfft2(a, s=s)


def _workers():
    n = os.cpu_count() or 1
    return max(1, min(n, 16))


def _canvas_size(nx, ny):
    m = int(np.ceil(1.25 * max(nx, ny)))
    m += m & 1
    while True:  # next 5-smooth even size
        k = m
        for f in (2, 3, 5):
            while k % f == 0:
                k //= f
        if k == 1:
            return m
        m += 2


def make_drift(nframe, step, seed, quad_sigma=0.0):
    """Cumulative random walk with t_0 = 0 (+ optional quadratic a f^2)."""
    rng = np.random.default_rng(seed + 7919)
    steps = rng.normal(0.0, step, size=(nframe, 2))
    steps[0] = 0.0
    t = np.cumsum(steps, axis=0)
    if quad_sigma > 0:
        a = rng.normal(0.0, quad_sigma, size=2)
        f = np.arange(nframe, dtype=np.float64)
        t = t + f[:, None] ** 2 * a[None, :]
    return t


def make_movie(nx, ny, nframe, dose=1.0, step=0.7, seed=1001, quad_sigma=0.0,
               dtype=np.float32, drift=None, poisson=True):
    """Returns (frames[nframe, ny, nx] of dtype, drift t[nframe, 2] (x, y))."""
    m = _canvas_size(nx, ny)
    rng = np.random.default_rng(seed)
    w = _workers()
    white = rng.standard_normal((m, m)).astype(np.float32)
    W = _rfft2(white, w)
    ky = np.fft.fftfreq(m)[:, None].astype(np.float32)
    kx = np.fft.rfftfreq(m)[None, :].astype(np.float32)
    k2 = kx * kx + ky * ky
    W *= np.sqrt(1.0 / (1.0 + k2 / np.float32(0.05 ** 2))).astype(np.float32)
    W[0, 0] = 0
    g = _irfft2(W, (m, m), w)
    g = (g - g.mean()) / g.std()
    phantom = np.exp(0.5 * g - 0.125).astype(np.float32) * np.float32(dose)
    P = _rfft2(phantom, w)
    if drift is None:
        drift = make_drift(nframe, step, seed, quad_sigma)
    frames = np.empty((nframe, ny, nx), dtype=dtype)
    kxs = np.fft.rfftfreq(m)[None, :]
    kys = np.fft.fftfreq(m)[:, None]
    lib = _poisson_lib() if poisson else None

    def one(f):
        tx, ty = drift[f]
        ph = (np.exp(-2j * np.pi * kys * ty).astype(np.complex64)
              * np.exp(-2j * np.pi * kxs * tx).astype(np.complex64))
        img = _irfft2(P * ph, (m, m), 1)[:ny, :nx]
        np.maximum(img, 0, out=img)
        if poisson:
            lam = np.ascontiguousarray(img, np.float32)
            cnt = np.empty_like(lam)
            lib.mc_synth_poisson(lam.ctypes.data, cnt.ctypes.data, lam.size,
                                 (seed * 1000003 + f * 7919 + 17) & 0xFFFFFFFFFFFFFFFF)
            frames[f] = cnt.astype(dtype)
        else:
            frames[f] = img.astype(dtype)
        return f

    with ThreadPoolExecutor(max_workers=w) as ex:
        list(ex.map(one, range(nframe)))
    return frames, drift


# ------------------------------------------------------ on-device movies
_gpu_lib = None


def _synth_gpu_lib():
    global _gpu_lib
    if _gpu_lib is None:
        path = os.path.join(_HERE, "libmc_synth_gpu.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path + " (run __graft_entry__.build())")
        lib = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        i = ctypes.c_int
        lib.mcs_gpu_movie.argtypes = [vp, i, i, i, i, vp, vp, i, ctypes.c_float,
                                      ctypes.c_ulonglong]
        _gpu_lib = lib
    return _gpu_lib


def phantom_waves(nw, seed, kc=0.05, kmax=0.5):
    """nw plane waves (k_x, k_y in cycles/px, phase in cycles) whose radial
    frequency density is proportional to k P(k), P(k) = 1/(1+(k/kc)^2), on
    (0, kmax]: inverse CDF k = kc sqrt(exp(u ln(1 + kmax^2/kc^2)) - 1)."""
    rng = np.random.default_rng(seed + 104729)
    u = rng.random(nw)
    k = kc * np.sqrt(np.expm1(u * np.log1p((kmax / kc) ** 2)))
    th = rng.random(nw) * 2 * np.pi
    ph = rng.random(nw)
    return np.ascontiguousarray(np.stack([k * np.cos(th), k * np.sin(th), ph], 1), np.float64)


def make_movie_gpu(nx, ny, nframe, dose=1.0, step=0.7, seed=1004, dtype=np.uint8, nwaves=256,
                   drift=None):
    """Movie generated in device memory (libmc_synth_gpu.so): a random-wave
    phantom moved exactly by the drift and Poisson-sampled.  Returns
    (motioncorr1exp_amd.DeviceArray of [nframe, ny, nx] dtype, drift)."""
    import motioncorr1exp_amd as mc

    if drift is None:
        drift = make_drift(nframe, step, seed)
    drift = np.ascontiguousarray(drift, np.float64)
    waves = phantom_waves(nwaves, seed)
    code = {np.dtype(np.uint8): 0, np.dtype(np.float32): 2}[np.dtype(dtype)]
    d = mc.DeviceArray(nbytes=nframe * ny * nx * np.dtype(dtype).itemsize)
    rc = _synth_gpu_lib().mcs_gpu_movie(d.ptr, code, nx, ny, nframe, drift.ctypes.data,
                                        waves.ctypes.data, nwaves, dose,
                                        (seed * 1000003 + 29) & 0xFFFFFFFFFFFFFFFF)
    if rc != 0:
        raise RuntimeError("mcs_gpu_movie failed (HIP error %d)" % rc)
    return d, drift

