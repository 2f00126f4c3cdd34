"""Writes tests/golden/fft_golden.npz: FP64 numpy (pocketfft) 2-D real FFT
vectors used to pin the oracle's transforms.  Inputs are seeded random
images of sizes 16, 26, 32 and 64 plus the 26 x 26 ramp h[y][x] = x + 1 of
the reference's own FFT self-check (src-1125/cufunc.cu:523-582).

    python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    out = {}
    rng = np.random.default_rng(20131)
    for n in (16, 26, 32, 64):
        img = rng.standard_normal((n, n))
        out["img_%d" % n] = img
        out["rfft2_%d" % n] = np.fft.rfft2(img)
    ramp = np.tile(np.arange(1, 27, dtype=np.float64), (26, 1))
    out["ramp_26"] = ramp
    out["rfft2_ramp_26"] = np.fft.rfft2(ramp)
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


if __name__ == "__main__":
    main()
