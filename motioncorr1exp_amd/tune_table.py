"""Builds the autotuner table of this device (one JSON line per geometry,
read back by mc_tune_load / Context.tune_load).

    python -m motioncorr1exp_amd.tune_table --out motioncorr1exp_amd/tune_mi355x.jsonl

Each geometry is tuned by mc_autotune on a synthetic movie of its shape
generated in device memory (random-wave phantom, Poisson counts).  The
default list covers the benchmark workloads: 3838x3710x40 float32 padded to
4096 and at its default square crop (n = 3710, chirp-z stage 1); 4096x4096x24
uint8 with 2x binning; uint8 movies padded to every
canvas size of the batched form (1024 ... 4096); 512x512x10 float32.
"""
import argparse
import os

import numpy as np

import motioncorr1exp_amd as mc
from motioncorr1exp_amd.synth import make_movie_gpu

# (nx, ny, frames, dtype, pad, bin, fod, box)
GEOMETRIES = [
    (3838, 3710, 40, np.float32, 4096, 1.0, 2, 96),
    (3838, 3710, 40, np.float32, 0, 1.0, 2, 96),
    (4096, 4096, 24, np.uint8, 0, 2.0, 11, 96),
    (4000, 4000, 30, np.uint8, 4096, 1.0, 2, 96),
    (3800, 3800, 30, np.uint8, 3840, 1.0, 2, 96),
    (3000, 3000, 30, np.uint8, 3072, 1.0, 2, 96),
    (2500, 2500, 30, np.uint8, 2560, 1.0, 2, 96),
    (2000, 2000, 30, np.uint8, 2048, 1.0, 2, 96),
    (1900, 1900, 30, np.uint8, 1920, 1.0, 2, 96),
    (1500, 1500, 30, np.uint8, 1536, 1.0, 2, 96),
    (1250, 1250, 30, np.uint8, 1280, 1.0, 2, 96),
    (1024, 1024, 30, np.uint8, 1024, 1.0, 2, 96),
    (512, 512, 10, np.float32, 0, 1.0, 4, 96),
]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args()
    if os.path.exists(args.out):
        os.remove(args.out)
    ctx = mc.Context(args.device)
    for i, (nx, ny, F, dt, pad, bin_, fod, box) in enumerate(GEOMETRIES):
        dev, _ = make_movie_gpu(nx, ny, F, dose=1.0, step=0.7, seed=500 + i, dtype=dt)
        prm = mc.default_params(F, pad=pad, bin=bin_, FrameDistOffset=fod, CCPeakSearchDim=box)
        rep = ctx.autotune(dev, (F, ny, nx), dt, prm, cache_path=args.out)
        print("%dx%dx%d %s pad %d bin %g: %s pair cols ms %s rows ms %s column plans ms %s" % (
            nx, ny, F, np.dtype(dt).name, pad, bin_,
            {k: v for k, v in rep.items() if not k.startswith("ms_")},
            ["%.3f" % v for v in rep["ms_pair_cols"]], ["%.3f" % v for v in rep["ms_pair_rows"]],
            ["%.3f" % v for v in rep["ms_cols_plan"]]), flush=True)
        dev.free()
    ctx.close()


if __name__ == "__main__":
    main()
