#!/usr/bin/env python3
"""Benchmark of the MI355X drift-correction path on its headline workload.

Workload (one step = one movie aligned end to end, frames resident in HBM):
3838 x 3710 x 40 float32 counts, mean 0.8 e/px/frame, seeded Gaussian random
field phantom, cumulative random-walk drift N(0, 0.7^2) px per axis, Poisson
noise; padded to 4096 x 4096 (region mean fill), bin 1, all pairs with frame
distance >= 18 (-fod 2, 253 pairs), B-factor 150, peak box 96, -itr 0.
A second line item times the same movie with -itr 3.

Prints one JSON line.  Errors are stated against the FP64 CPU oracle run on
the same movie: per-frame shift RMS / max (px) and 1/32-grid flips, aligned
sum relative L2 error end to end and stage-isolated (oracle shifts fed to
the GPU sum).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "movies aligned/sec at 3838x3710x40; shift RMS err vs FP64 oracle (px)"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


FP32_PEAK_TFLOPS = 157.3


def fft_flops(group, F, n, h, P, box):
    """Algorithmic FP32 flops per movie of an FFT-bound kernel group.
    spectra: per frame, h row R2Cs (an n/2-point complex FFT, 5 (n/2)
    log2(n/2), plus the unpack, 10 per bin) and n/2+1 column FFTs
    (5 n log2 n).  pair_cols: per (pair, kx) item, the weighted product (9
    per element), n/128 inverse 128-point FFTs as radix 16 (168 flops per
    butterfly) then radix 8 with twiddles (98), and the combine of box rows
    from n/128 terms (8 each)."""
    import math
    if group == "spectra":
        rows = h * (5 * (n // 2) * math.log2(n // 2) + 10 * (n // 2))
        cols = (n // 2 + 1) * 5 * n * math.log2(n)
        return F * (rows + cols)
    if group == "pair_cols":
        item = 9 * n + (n // 16) * 168 + (n // 8) * 98 + box * (n // 128) * 8
        return P * (n // 2 + 1) * item
    return None


def compulsory_bytes(F, n, nx, ny, in_elem, P, box, itr):
    """Compulsory HBM bytes of one movie: every tensor read once per consuming
    stage, every output written once (stage 1: frames in, spectra out; pairs:
    spectra in, box maps out; per align-to-sum round: spectra twice + sum;
    final sum: spectra in, sum out, image out)."""
    S = n * (n // 2 + 1) * 8
    b = F * nx * ny * in_elem + F * S          # stage 1
    b += F * S + P * box * box * 4             # pairs
    b += itr * (2 * F * S + 2 * S + F * box * box * 4)  # align-to-sum rounds
    b += F * S + S + S + 4 * n * n             # final sum + C2R
    return b


def _host_cpu():
    """nproc, the CPUs this process may run on, and the CPU model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "affinity": aff, "model": model,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def _sample_seconds(L, n, nx, ny, reps):
    """Per-unit times of the FP32 port on one padded 4096^2 frame: crop/pad +
    R2C, pair cross-power + full C2R + FindShift, phase shift + add, C2R."""
    import ctypes

    import numpy as np
    rng = np.random.default_rng(5)
    frame = rng.poisson(0.8, size=(ny, nx)).astype(np.float32)
    buf = np.zeros((n, n + 2), np.float32)
    buf2 = np.zeros((n, n + 2), np.float32)
    box = np.zeros((96, 96), np.float32)
    sx, sy = ctypes.c_double(), ctypes.c_double()
    t0 = time.perf_counter()
    for _ in range(reps):
        L.or_pad2fft(frame.ctypes.data, nx, ny, buf.ctypes.data, 0, 0, n)
        L.or_r2c2d(buf.ctypes.data, n)
    t_frame = (time.perf_counter() - t0) / reps
    spec = buf.copy()
    t0 = time.perf_counter()
    for _ in range(reps):
        buf2[:] = spec
        L.or_cross_power(buf2.ctypes.data, spec.ctypes.data, n, ctypes.c_float(150.0))
        L.or_c2r2d(buf2.ctypes.data, n)
        L.or_find_shift(buf2.ctypes.data, n, box.ctypes.data, 96, ctypes.byref(sx),
                        ctypes.byref(sy), -1)
    t_pair = (time.perf_counter() - t0) / reps
    acc = np.zeros_like(spec)
    t0 = time.perf_counter()
    for _ in range(reps):
        buf2[:] = spec
        L.or_phase_shift(buf2.ctypes.data, n, 0.37, -1.21)
        acc += buf2
    t_shift = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    L.or_c2r2d(acc.ctypes.data, n)
    t_c2r = time.perf_counter() - t0
    return t_frame, t_pair, t_shift, t_c2r


def cpu_baseline(frames, pad, P, F):
    """The FP32 port of the reference algorithm (oracle/liboracle_f32.so: the
    reference's dataflow with a full C2R per pair, own FFT, OpenMP) timed on
    this host.  `value`: one whole movie of the workload with every CPU this
    process may use (the box's share), measured.  `one_thread`: the same
    port on one thread, on a bounded sample (one frame R2C, one pair, one
    phase shift, one C2R) scaled to the movie; the all-thread sample scaled
    the same way is kept as a cross-check of the scaling."""
    import ctypes

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import Oracle

    host = _host_cpu()
    threads = host["affinity"]
    if host["OMP_NUM_THREADS"]:
        threads = min(threads, max(1, int(host["OMP_NUM_THREADS"])))
    o = Oracle("f32")
    L = o.lib
    L.or_pad2fft.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                             ctypes.c_int, ctypes.c_int, ctypes.c_int]
    gomp = ctypes.CDLL("libgomp.so.1")
    nz, ny, nx = frames.shape
    t0 = time.perf_counter()
    o.align_movie(frames, dict(pad=pad, bin=1.0, fod=2, bfactor=150.0, box=96, iter=0,
                               full_c2r=1, nthreads=threads), want_boxes=False)
    t_movie = time.perf_counter() - t0
    gomp.omp_set_num_threads(threads)
    ta = _sample_seconds(L, pad, nx, ny, 1)
    gomp.omp_set_num_threads(1)
    t1 = _sample_seconds(L, pad, nx, ny, 1)
    gomp.omp_set_num_threads(threads)

    def scaled(t):
        return F * t[0] + P * t[1] + F * t[2] + t[3]
    T1, TA = scaled(t1), scaled(ta)
    return {
        "value": 1.0 / t_movie, "unit": "movies/s", "cores": threads, "kind": "port",
        "sample": ("one whole %dx%dx%d movie (pad %d, %d pairs, itr 0) through the FP32 port of "
                   "the reference path (oracle/liboracle_f32.so: full C2R per pair, own FFT, "
                   "OpenMP on %d threads), measured: %.1f s" % (nx, ny, F, pad, P, threads, t_movie)),
        "host": host,
        "one_thread": {"value": 1.0 / T1, "unit": "movies/s", "cores": 1,
                       "seconds_per_movie": T1,
                       "sample": "1x frame pad+R2C %.3fs, 1x pair cross-power+full C2R+FindShift "
                                 "%.3fs, 1x phase shift+add %.3fs, 1x C2R %.3fs, scaled to %d frames "
                                 "+ %d pairs + %d shifts + 1 C2R" % (t1 + (F, P, F))},
        "all_threads_sample_cross_check": {"seconds_per_movie": TA,
                                           "measured_seconds_per_movie": t_movie},
    }


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_3838x3710x40_pad4096.txt")
# canvas sizes of the batching form: the supported transform sizes from 1024
PAD_SIZES = (1024, 1280, 1536, 1920, 2048, 2560, 3072, 3840, 4096)


def committed_traffic(group, launches_per_movie=1.0):
    """HBM bytes per launch of a kernel group from the committed rocprofv3
    counter summary (FETCH_SIZE and WRITE_SIZE runs of this workload, scaled
    by the copy calibration): the sum over the group's kernels of their bytes
    per dispatch times their dispatches per movie (`per_movie`; a launch of
    the column pass is several kx-segment dispatches), divided by the group's
    launches per movie."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    tot, seen = 0.0, False
    with open(TRAFFIC_FILE) as f:
        for line in f:
            if not line.startswith("traffic "):
                continue
            kv = dict(tok.split("=", 1) for tok in line.split()[1:])
            if kv.get("group") == group:
                tot += float(kv["bytes"]) * float(kv.get("per_movie", "1"))
                seen = True
    return tot / launches_per_movie if seen else None


def batch_leg(n_movies, steps, seed=1004, tune=True, tune_cache=None, oracle=True):
    """Batching form: n_movies movies of seeded uneven size (square sides
    uniform over the even values 1024..4096, padded with the region mean to
    the next supported transform size), 8..50 frames, dose uniform in [0.5, 1.5],
    random-walk drift N(0, 0.7^2), uint8 Poisson counts of the random-field
    phantom (synth.make_movie); bin 1, fod 2,
    bft 150, pbx 96, 3 align-to-sum rounds.  Times mc_align_batch (one call =
    one step over all movies) and the same movies through mc_align_movie one
    by one; reports batch-vs-single shift differences (max and RMS) and, for
    the smallest and the largest canvas, the batch results against the FP64
    oracle."""
    import numpy as np

    import motioncorr1exp_amd as mc
    from motioncorr1exp_amd.synth import make_movie

    rng = np.random.default_rng(seed)
    t0 = time.time()
    specs = []
    for i in range(n_movies):
        side = int(rng.integers(512, 2049)) * 2
        F = int(rng.integers(8, 51))
        dose = float(rng.uniform(0.5, 1.5))
        specs.append((side, F, dose, min(z for z in PAD_SIZES if z >= side)))
    pick = sorted(range(n_movies), key=lambda i: (specs[i][3], specs[i][1]))
    pick = (pick[0], pick[-1])
    movies, host = [], {}
    for i, (side, F, dose, pad) in enumerate(specs):
        # generated on the host one movie at a time, uploaded, host copy dropped
        frames, drift = make_movie(side, side, F, dose=dose, step=0.7, seed=seed * 100 + i,
                                   dtype=np.uint8)
        if i in pick:
            host[i] = frames
        prm = mc.default_params(F, pad=pad, bin=1.0, FrameDistOffset=2, bfactor=150.0,
                                CCPeakSearchDim=96, iter=3)
        movies.append(dict(dev=mc.DeviceArray(frames), shape=(F, side, side), dtype=np.uint8,
                           params=prm, want=("frame_shift",), drift=drift))
    log("batch: %d random-field movies generated in %.1fs" % (n_movies, time.time() - t0))
    ctx = mc.Context(0)
    if tune:
        # shipped table first; geometries it lacks are tuned on their first movie
        have = set()
        if tune_cache and os.path.exists(tune_cache):
            ctx.tune_load(tune_cache)
            with open(tune_cache) as f:
                for line in f:
                    e = json.loads(line)
                    have.add((e["nu"], e["dtype"], e["box"], e["precision"]))
        for m in movies:
            key = (m["params"]["pad"], mc.MC_DTYPE[np.dtype(m["dtype"])], 96, 0)
            if key not in have:
                have.add(key)
                ctx.autotune(m["dev"], m["shape"], m["dtype"], m["params"])
    res, br = ctx.align_batch(movies)  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        res, br = ctx.align_batch(movies)
    mc.check(mc.lib().mc_sync())
    dt = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    seq = [ctx.align_movie(m["dev"], m["shape"], m["dtype"], m["params"], want=("frame_shift",))
           for m in movies]
    mc.check(mc.lib().mc_sync())
    dts = time.perf_counter() - t0
    diff = np.concatenate([(a["frame_shift"] - b["frame_shift"]).ravel() for a, b in zip(res, seq)])
    err = []
    for m, r in zip(movies, res):
        F = m["shape"][0]
        truth = m["drift"] - m["drift"][F // 2 + 1]  # RefFrame = nz/2+1
        err.append(r["frame_shift"] - truth)
    err = np.concatenate(err)
    sizes = {}
    for m in movies:
        n = m["params"]["pad"]
        sizes[n] = sizes.get(n, 0) + 1
    if oracle:  # one more (untimed) batch call that also returns these movies' sums
        res2, _ = ctx.align_batch([dict(m, want=("frame_shift", "sum_image")) if i in pick else m
                                   for i, m in enumerate(movies)])
    ctx.close()
    out = {"movies": n_movies, "steps": steps, "movies_per_s": n_movies / dt,
           "ms_per_batch": dt * 1e3, "sequential_movies_per_s": n_movies / dts,
           "speedup_vs_sequential": dts / dt,
           "frames_total": int(sum(m["shape"][0] for m in movies)),
           "movies_per_transform_size": {str(k): v for k, v in sorted(sizes.items())},
           "batch_report": br,
           "batch_vs_single_shift_max_px": float(np.abs(diff).max()),
           "batch_vs_single_shift_rms_px": float(np.sqrt(np.mean(diff ** 2))),
           "vs_truth_rms_px": float(np.sqrt(np.mean(err ** 2))),
           "workload": "%d movies, sides U{1024..4096} even padded to the next 2^k*{1,3,5,15}, frames U{8..50}, "
                       "dose U[0.5,1.5], uint8 Poisson counts of the random-field phantom, bin 1, "
                       "fod 2, bft 150, pbx 96, itr 3" % n_movies}
    if oracle:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from _oracle import Oracle

        t0 = time.time()
        checks = []
        for i in pick:
            m, r = movies[i], res2[i]
            assert np.array_equal(r["frame_shift"], res[i]["frame_shift"])
            frames = host[i].astype(np.float32)
            o = Oracle("f64").align_movie(frames, dict(pad=m["params"]["pad"], fod=2, box=96,
                                                       iter=3, full_c2r=0, nthreads=16),
                                          want_boxes=False)
            d = r["frame_shift"] - o["frame_shift"]
            checks.append({"canvas": m["params"]["pad"], "frames": m["shape"][0],
                           "shift_rms_err_px": float(np.sqrt(np.mean(d ** 2))),
                           "shift_max_err_px": float(np.abs(d).max()),
                           "frames_beyond_0.01px": int((np.abs(d) > 0.01).any(1).sum()),
                           "sum_rel_err": float(np.linalg.norm(r["sum_image"] - o["sum_image"]) /
                                                np.linalg.norm(o["sum_image"]))})
        out["batch_vs_fp64_oracle"] = checks
        log("batch: oracle checks in %.1fs" % (time.time() - t0))
    for m in movies:
        m["dev"].free()
    log("batch: %s" % {k: v for k, v in out.items() if k != "workload"})
    return out


def grad_leg(ctx, dev, sum_dev, params, nx, ny, F, steps, oracle=True):
    """Differentiation form: 3838x3710x40 float32 frames with sub-pixel drift
    (random-walk steps N(0, 0.4 px)) plus a quadratic slow drift a f^2, a ~
    N(0, 0.02^2) px per axis; the same alignment followed by 20
    adjoint-gradient refinement steps of the per-frame shifts.  Reports the
    time per movie with the refinement, the sharpness before and after, and
    the shifts before and after against the known drift and against the FP64
    oracle's peak-search shifts of the same movie."""
    import numpy as np

    import motioncorr1exp_amd as mc
    from motioncorr1exp_amd.synth import make_movie
    frames, drift = make_movie(nx, ny, F, dose=0.8, step=0.4, seed=1005, quad_sigma=0.02)
    L = mc.lib()
    mc.check(L.mc_mem_h2d(dev.ptr, frames.ctypes.data, frames.nbytes))
    mc.check(L.mc_sync())
    pg = dict(params, grad_steps=20, grad_step_px=0.05)

    def run(prm):
        return ctx.align_movie(dev, frames.shape, frames.dtype, prm, want=("frame_shift",),
                               sum_dev=sum_dev)
    r0 = run(params)
    run(pg)
    mc.check(L.mc_sync())
    k = max(2, steps // 2)
    t0 = time.perf_counter()
    for _ in range(k):
        r1 = run(pg)
    mc.check(L.mc_sync())
    dt = (time.perf_counter() - t0) / k
    truth = drift - drift[F // 2 + 1]
    rep = r1["report"]
    res = {"grad_steps": 20, "ms_per_movie": dt * 1e3, "movies_per_s": 1.0 / dt,
           "sharpness_before": rep["sharp_j0"], "sharpness_after": rep["sharp_j1"],
           "accepted_steps": rep["grad_accepted"],
           "vs_truth_rms_px_before": float(np.sqrt(np.mean((r0["frame_shift"] - truth) ** 2))),
           "vs_truth_rms_px_after": float(np.sqrt(np.mean((r1["frame_shift"] - truth) ** 2)))}
    if oracle:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from _oracle import Oracle
        o = Oracle("f64").align_movie(frames, dict(pad=params["pad"], fod=2, box=96, iter=0,
                                                   full_c2r=0, nthreads=16), want_boxes=False)
        res["vs_fp64_oracle_rms_px_before"] = float(np.sqrt(np.mean((r0["frame_shift"] - o["frame_shift"]) ** 2)))
        res["vs_fp64_oracle_rms_px_after"] = float(np.sqrt(np.mean((r1["frame_shift"] - o["frame_shift"]) ** 2)))
        res["oracle_vs_truth_rms_px"] = float(np.sqrt(np.mean((o["frame_shift"] - truth) ** 2)))
    log("differentiation form: %s" % res)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-oracle", action="store_true", help="skip the FP64 error check")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-itr3", action="store_true")
    ap.add_argument("--no-precision", action="store_true", help="skip the MFMA hi/lo form")
    ap.add_argument("--no-refgeom", action="store_true",
                    help="skip the reference's default geometry (n = 3710, no padding)")
    ap.add_argument("--frames", type=int, default=40)
    ap.add_argument("--no-batch", action="store_true", help="skip the batching form")
    ap.add_argument("--no-grad", action="store_true", help="skip the differentiation form")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes, no autotuner")
    ap.add_argument("--retune", action="store_true", help="autotune even if the table has an entry")
    ap.add_argument("--tune-cache", default=os.path.join(ROOT, "motioncorr1exp_amd", "tune_mi355x.jsonl"),
                    help="autotuner table loaded before timing (mc_tune_load)")
    ap.add_argument("--batch-movies", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--calib-copy", action="store_true",
                    help="only copy 1 GiB device-to-device (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU per box; --gpus %d is not supported "
              "(movies are independent: several GPUs would each take whole movies from "
              "a shared queue)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import motioncorr1exp_amd as mc
    from motioncorr1exp_amd.synth import make_movie

    if args.calib_copy:
        nb = 1 << 30
        a = mc.DeviceArray(nbytes=nb)
        b = mc.DeviceArray(nbytes=nb)
        mc.check(mc.lib().mc_mem_zero(a.ptr, nb))
        mc.check(mc.lib().mc_sync())
        mc.check(mc.lib().mc_mem_d2d(b.ptr, a.ptr, nb))
        mc.check(mc.lib().mc_sync())
        print(json.dumps({"calib_copy_bytes": nb}))
        return

    nx, ny, F, pad = 3838, 3710, args.frames, 4096
    t0 = time.time()
    frames, drift = make_movie(nx, ny, F, dose=0.8, step=0.7, seed=1003)
    log("generated %dx%dx%d movie in %.1fs" % (nx, ny, F, time.time() - t0))
    params = mc.default_params(F, pad=pad, bin=1.0, FrameDistOffset=2, bfactor=150.0,
                               CCPeakSearchDim=96, iter=0)
    ctx = mc.Context(0)
    dev = mc.DeviceArray(nbytes=frames.nbytes)
    sum_dev = mc.DeviceArray(nbytes=pad * pad * 4)
    L = mc.lib()
    # host -> device upload of the movie (pageable host memory), outside `value`
    mc.check(L.mc_mem_h2d(dev.ptr, frames.ctypes.data, frames.nbytes))  # first touch
    mc.check(L.mc_sync())
    t0 = time.perf_counter()
    mc.check(L.mc_mem_h2d(dev.ptr, frames.ctypes.data, frames.nbytes))
    mc.check(L.mc_sync())
    t_h2d = time.perf_counter() - t0

    def step(prm, want=("frame_shift",)):
        return ctx.align_movie(dev, frames.shape, frames.dtype, prm, want=want, sum_dev=sum_dev)

    tune = None
    if not args.no_tune:
        # launch shapes from the project's autotuner: the shipped table for
        # this device when it has this geometry, else tuned now on this movie
        # (untimed)
        loaded = 0
        if args.tune_cache and os.path.exists(args.tune_cache) and not args.retune:
            loaded = ctx.tune_load(args.tune_cache)
        if loaded:
            tune = {"source": os.path.relpath(args.tune_cache, ROOT), "entries": loaded}
        else:
            t0 = time.time()
            tune = ctx.autotune(dev, frames.shape, frames.dtype, params)
            tune["source"] = "tuned in this run"
            log("autotune in %.1fs" % (time.time() - t0))
        log("launch shapes: %s" % {k: v for k, v in tune.items() if not k.startswith("ms_")})

    for _ in range(args.warmup):
        step(params)
    mc.check(L.mc_sync())
    # per-group kernel breakdown from an instrumented pass outside the timed
    # region (events around every group cost about 10 us each)
    nb = max(2, min(args.steps, 5))
    ctx.kernel_timing(True)
    for _ in range(nb):
        step(params)
    mc.check(L.mc_sync())
    kern = {}
    for g in mc.KT:
        c, t = ctx.kernel_time(g)
        kern[g] = {"launches": c, "ms_per_launch": (t / c) if c else 0.0, "ms_per_movie": t / nb}
    ctx.kernel_timing(False)
    dom = max(kern, key=lambda k: kern[k]["ms_per_movie"])
    # timed region: HIP events around the dominant group only
    ctx.kernel_timing(True, groups=(dom,))
    mc.check(L.mc_sync())
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step(params)
    mc.check(L.mc_sync())
    dt = time.perf_counter() - t0
    c, t = ctx.kernel_time(dom)
    live = {"group": dom, "launches": c, "ms_per_launch": (t / c) if c else 0.0,
            "ms_per_movie": t / args.steps}
    ctx.kernel_timing(False)
    ms = dt * 1e3 / args.steps
    rep = res["report"]
    P = rep["npairs"]
    n = rep["n"]
    log("itr0: %.3f ms/movie; %s %.3f ms/launch in the timed region; breakdown pass %s" % (
        ms, dom, live["ms_per_launch"], {k: round(v["ms_per_movie"], 3) for k, v in kern.items()}))

    # roofline of the dominant kernel group, from its launches in the timed region
    S = n * (n // 2 + 1) * 8
    unit_bytes = {
        # frames in + spectra out per frame (padded 4096^2 canvas)
        "spectra": F * (nx * ny * 4 + S),
        # every spectrum read once + the pair maps of one launch
        "pair_cols": F * S + P * (n // 2 + 1) * 96 * 8,
        "pair_rows": P * (n // 2 + 1) * 96 * 8 + P * 96 * 96 * 4,
        "peaks": P * 96 * 96 * 4,
        "sum": F * S + S,
        "image": S + 4 * n * n,
    }
    alg = unit_bytes.get(dom, 0)
    per_launch_ms = live["ms_per_launch"]
    # launches of the dominant group per movie
    lpm = live["launches"] / args.steps
    achieved = alg / lpm / (per_launch_ms * 1e-3) / 1e9 if per_launch_ms > 0 else 0.0
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": committed_traffic(dom, lpm),
                "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT),
                "algorithmic_bytes_per_launch": alg / lpm}
    # the FFT groups are bound by vector-ALU issue: their arithmetic against
    # the FP32 peak as well
    fl = fft_flops(dom, F, n, ny, P, 96)
    if fl and per_launch_ms > 0:
        tf = fl / lpm / (per_launch_ms * 1e-3) / 1e12
        roofline["compute"] = {"flops_per_launch": fl / lpm, "achieved": tf,
                               "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                               "frac": tf / FP32_PEAK_TFLOPS}
    comp = compulsory_bytes(F, n, nx, ny, 4, P, 96, 0)
    movies_s = args.steps / dt
    whole = {"compulsory_bytes_per_movie": comp,
             "achieved_GBs": comp * movies_s / 1e9,
             "frac_of_8TBs": comp * movies_s / 1e9 / HBM_PEAK_GBS}

    out = {
        "metric": METRIC, "value": movies_s, "unit": "movies/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded random-field phantom, Poisson counts, known drift)",
        "config": {"workload": "3838x3710x40 f32 counts dose 0.8, pad 4096, bin 1, fod 2 "
                               "(253 pairs), bft 150, pbx 96, itr 0",
                   "movies": 1, "frames": F, "nx": nx, "ny": ny, "pad": pad, "pairs": P,
                   "form": "plain FP32"},
        "roofline": roofline, "whole_movie_roofline": whole, "kernels": kern, "kernels_source": "instrumented pass of %d movies outside the timed region" % nb,
        "dominant_kernel_live": live, "autotune": tune,
        "stage_ms": {k: rep[k] for k in ("ms_spectra", "ms_pairs", "ms_solve", "ms_iter",
                                         "ms_sum", "ms_total")},
        "host_upload": {
            "bytes": int(frames.nbytes), "ms": t_h2d * 1e3,
            "GBs": frames.nbytes / t_h2d / 1e9,
            # one movie uploaded then aligned, no overlap
            "movies_per_s_serial": 1.0 / (t_h2d + ms * 1e-3),
            # upload of the next movie overlapped with the alignment of this one
            "movies_per_s_overlapped": 1.0 / max(t_h2d, ms * 1e-3),
        },
    }

    if not args.no_itr3:
        p3 = dict(params, iter=3)
        step(p3)
        mc.check(L.mc_sync())
        t0 = time.perf_counter()
        k3 = max(2, args.steps // 2)
        for _ in range(k3):
            r3 = step(p3)
        mc.check(L.mc_sync())
        dt3 = time.perf_counter() - t0
        out["itr3"] = {"movies_per_s": k3 / dt3, "ms_per_movie": dt3 * 1e3 / k3,
                       "niter": r3["report"]["niter"]}
        log("itr3: %.3f ms/movie" % (dt3 * 1e3 / k3))

    if not args.no_refgeom:
        # the reference's default geometry: square crop min(nx, ny) = 3710 =
        # 2*5*7*53, no padding (stage 1 in two levels, 35 x 53 rows and 70 x 53
        # columns; chirp-z for the image)
        pr = dict(params, pad=0)
        step(pr)
        mc.check(L.mc_sync())
        ctx.kernel_timing(True)
        t0 = time.perf_counter()
        kr = max(2, args.steps // 2)
        for _ in range(kr):
            rr = step(pr)
        mc.check(L.mc_sync())
        dtr = time.perf_counter() - t0
        kernr = {g: ctx.kernel_time(g)[1] / kr for g in mc.KT}
        ctx.kernel_timing(False)
        out["reference_geometry"] = {
            "workload": "the same movie at the reference's default square crop, n = %d "
                        "(no -pad; two-level stage-1 transforms), itr 0" % rr["report"]["n"],
            "n": rr["report"]["n"], "movies_per_s": kr / dtr, "ms_per_movie": dtr * 1e3 / kr,
            "kernels_ms_per_movie": kernr}
        log("reference geometry n=%d: %.3f ms/movie %s" % (rr["report"]["n"], dtr * 1e3 / kr,
                                                         {k: round(v, 3) for k, v in kernr.items()}))

    if not args.no_precision:
        # precision form: pair-stage DFTs as FP16 hi/lo MFMA products
        p1 = dict(params, precision=1)
        step(p1)
        mc.check(L.mc_sync())
        ctx.kernel_timing(True)
        t0 = time.perf_counter()
        k1 = max(2, args.steps // 2)
        for _ in range(k1):
            r1 = step(p1)
        mc.check(L.mc_sync())
        dt1 = time.perf_counter() - t0
        kern1 = {g: ctx.kernel_time(g)[1] / k1 for g in mc.KT}
        cnt1 = {g: ctx.kernel_time(g)[0] / k1 for g in mc.KT}
        ctx.kernel_timing(False)
        d1 = r1["frame_shift"] - res["frame_shift"]
        dom1 = max(kern1, key=lambda k: kern1[k])
        # the form's dominant group against the HBM roofline: its algorithmic
        # bytes (the FP32 form's, plus the column maxima the hi/lo row pass
        # reads) over its measured time per launch
        alg1 = unit_bytes.get(dom1, 0) + (P * (n // 2 + 1) * 4 * 2 if dom1 in ("pair_cols", "pair_rows") else 0)
        ms1 = kern1[dom1] / max(cnt1[dom1], 1e-9)
        ach1 = alg1 / max(cnt1[dom1], 1e-9) / (ms1 * 1e-3) / 1e9 if ms1 > 0 else 0.0
        out["precision_form"] = {
            "form": "MFMA FP16 hi/lo pair stage (FP32 accumulation)",
            "movies_per_s": k1 / dt1, "ms_per_movie": dt1 * 1e3 / k1,
            "kernels_ms_per_movie": kern1,
            "roofline": {"bound": "hbm", "kernel": dom1, "achieved": ach1, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": ach1 / HBM_PEAK_GBS,
                         "algorithmic_bytes_per_launch": alg1 / max(cnt1[dom1], 1e-9)},
            "shift_rms_diff_vs_fp32_px": float(np.sqrt(np.nanmean(d1 ** 2))),
        }
        log("precision form: %.3f ms/movie, shift RMS vs FP32 %.2e px"
            % (dt1 * 1e3 / k1, out["precision_form"]["shift_rms_diff_vs_fp32_px"]))

    if not args.no_oracle:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from _oracle import Oracle

        t0 = time.time()
        full = ctx.align_movie(dev, frames.shape, frames.dtype, params,
                               want=("frame_shift", "final_shift", "pair_b", "sum_image"))
        o = Oracle("f64").align_movie(
            frames, dict(pad=pad, bin=1.0, fod=2, bfactor=150.0, box=96, iter=0, full_c2r=0,
                         nthreads=min(os.cpu_count() or 1, 16)), want_boxes=False)
        log("oracle FP64 in %.1fs" % (time.time() - t0))
        d = full["frame_shift"] - o["frame_shift"]
        d = d[np.isfinite(d).all(1)]
        flips = int(np.sum(np.abs(full["pair_b"] - o["pair_b"]) > 0.01))
        rel = float(np.linalg.norm(full["sum_image"] - o["sum_image"]) /
                    np.linalg.norm(o["sum_image"]))
        iso = ctx.sum_image(o["frame_shift"], n)
        rel_iso = float(np.linalg.norm(iso - o["sum_image"]) / np.linalg.norm(o["sum_image"]))
        truth = drift - drift[F // 2 + 1]  # RefFrame = nz/2+1 (DFAlign.cu:1133)
        out["accuracy"] = {
            "shift_rms_err_px": float(np.sqrt(np.mean(d ** 2))),
            "shift_max_err_px": float(np.abs(d).max()),
            "pair_components_beyond_0.01px": flips,
            "frames_beyond_0.01px": int(np.sum((np.abs(d) > 0.01).any(1))),
            "sum_rel_err": rel, "sum_rel_err_stage_isolated": rel_iso,
            "vs_truth_rms_px": float(np.sqrt(np.mean((full["frame_shift"] - truth) ** 2))),
        }
        out["shift_rms_err_px"] = out["accuracy"]["shift_rms_err_px"]
        log("accuracy %s" % out["accuracy"])
        if "precision_form" in out:
            p1 = dict(params, precision=1)
            f1 = ctx.align_movie(dev, frames.shape, frames.dtype, p1,
                                 want=("frame_shift", "sum_image"))
            d1 = f1["frame_shift"] - o["frame_shift"]
            d1 = d1[np.isfinite(d1).all(1)]
            out["precision_form"]["shift_rms_err_vs_fp64_oracle_px"] = float(np.sqrt(np.mean(d1 ** 2)))
            out["precision_form"]["sum_rel_err_vs_fp64_oracle"] = float(
                np.linalg.norm(f1["sum_image"] - o["sum_image"]) / np.linalg.norm(o["sum_image"]))

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(frames, pad, P, F)
        out["gpu_over_cpu"] = movies_s / out["cpu_baseline"]["value"]
        log("cpu baseline %s" % out["cpu_baseline"])

    if not args.no_grad:
        out["differentiation_form"] = grad_leg(ctx, dev, sum_dev, params, nx, ny, F, args.steps,
                                               oracle=not args.no_oracle)

    ctx.close()
    dev.free()
    sum_dev.free()
    if not args.no_batch:
        out["batch"] = batch_leg(args.batch_movies, args.batch_steps, tune=not args.no_tune,
                                 tune_cache=None if args.retune else args.tune_cache,
                                 oracle=not args.no_oracle)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
