"""Benchmark: MotionCor3 frame alignment + dose-weighted sum on one MI355X.

Workload (one step = one movie): 5760 x 4092 px, 50 frames of float32
Poisson counts at 1.0 e-/px/frame, 20k seeded blobs, global drift plus a
3 px quadratic-in-(x,y) x cubic-in-t local deformation; 5x5 patches,
-Bft 500 100, -PixSize 0.83, -kV 300, -FmDose 1.0 (sum, _DW, _DWS), all other
flags at the reference defaults (-Iter 15 -Tol 0.1 -Group 1 4).  The movie is
rendered on the GPU before timing and stays resident in HBM; a step is one
mc_align_movie call (measurement, correction and the three output sums).

Prints one JSON line.  The roofline figure is for the dominant kernel class,
timed with HIP events on the library's stream during the timed steps.  The
CPU baseline is the FP64 oracle (oracle/) on the timed movie itself (all 50
frames, on every host core): its wall time, and the GPU's shifts for that
movie against its shifts ("shift RMS err vs FP64").
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "movies/s @5760x4092x50 frames, 5x5 patches; shift RMS err vs FP64 (px)"
W, H, N = 5760, 4092, 50
PEAK_HBM_GBS = 8000.0


def algorithmic_bytes(kernel, g, launches):
    """Compulsory HBM bytes of one launch of `kernel` in this workload:
    each stage reads its inputs once and writes its outputs once; `launches`
    is the kernel class's launch count per movie (the forward classes run in
    the forward and local-correction passes, the inverse ones in the second
    only)."""
    cx, xcx = W // 2 + 1, g["xcf"][0] // 2 + 1
    S_f = 8.0 * cx * H                         # full half spectrum
    R_f = 4.0 * W * H                          # full real frame
    S_x = 8.0 * xcx * g["xcf"][1]              # xcf half spectrum
    if kernel == "fft_cols_fwd":               # T -> S (+ xcf crop in the first pass)
        return (N * 2 * S_f + N * S_x + N * 2 * S_f) / launches
    if kernel == "fft_rows_r2c":               # real -> T
        return (2 * N * (R_f + S_f)) / launches
    if kernel == "fft_cols_inv":               # S -> T
        return N * 2 * S_f / launches
    if kernel == "fft_rows_c2r":               # T -> real
        return N * (S_f + R_f) / launches
    if kernel == "local_warp":                 # all frames, read + write
        return N * 2 * R_f
    if kernel == "dw_sums":                    # N spectra in, 3 sums out
        return N * S_f + 3 * S_f
    raise ValueError(kernel)


def path_bytes(g, iters_global, iters_patch, nsum, P=25):
    """Compulsory HBM bytes of one movie through the whole path: every stage
    reads its inputs once and writes its outputs once (FFT = one read + one
    write), adjacent elementwise stages fused."""
    cx, H_ = W // 2 + 1, H
    S_f, R_f = 8.0 * cx * H_, 4.0 * W * H_
    xw, xh = g["xcf"]
    S_x, R_x = 8.0 * (xw // 2 + 1) * xh, 4.0 * xw * xh
    pw, ph = g["patch"]
    S_p, R_p = 8.0 * (pw // 2 + 1) * ph, 4.0 * pw * ph
    b = N * (R_f + S_f) + 2 * N * S_x + iters_global * (2 * N + 2) * S_x
    b += ((N + 1) * S_x + N * (S_x + R_x) + 2 * N * S_f + P * N * (R_p + S_p)
          + P * iters_patch * (2 * N + 2) * S_p + N * (2 * S_f + 2 * R_f))
    b += N * S_f + nsum * S_f
    return b


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-calibrated), or None when none was recorded."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_hbm_bytes.md"))):
        for line in open(path):
            f = line.split()
            if len(f) == 4 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
                return int(f[3])
    return None


def host_info():
    """CPU model, logical CPUs and the oracle's OpenMP thread count."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"cpu_model": model, "nproc": n}


def cpu_baseline(ctx, frames, gres, params_kw, frames_needed):
    """The FP64 oracle on the timed movie itself (all its frames, or the first
    `frames_needed`): its wall time is the CPU baseline, and the GPU's shifts
    for that movie are compared with its shifts."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    # the oracle prints each stage's wall time on stderr as it finishes (a
    # long run shows its progress)
    os.environ.setdefault("ORC_TIMING", "1")
    import oracle as O

    from motioncor3_amd import lib as L

    po = O.default_params(**{k: v for k, v in params_kw.items()
                             if k not in ("spectra_fmt", "frame_fmt", "movie_pipeline")})
    if frames_needed >= N:
        host = frames.numpy()
        g = gres
        what = "the whole timed movie (%d frames)" % N
    else:
        # the first frames of the timed movie, and the GPU on the same frames
        sub = ctx.empty((frames_needed, H, W), np.float32)
        ctx.copy(sub, frames, sub.nbytes)
        host = sub.numpy()
        g = ctx.align_movie(sub, L.default_params(**params_kw))
        what = "the first %d of the timed movie's %d frames" % (frames_needed, N)
    t0 = time.time()
    ores = O.align_movie(host, po)
    dt = time.time() - t0
    dg = g["global_shift"] - ores["global_shift"]
    rms_g = float(np.sqrt(np.mean(np.sum(dg ** 2, axis=1))))
    same = bool(np.array_equal(g["patch_starts"], ores["patch_starts"]))
    dp = g["patch_shift"] - ores["patch_shift"]
    rms_p = float(np.sqrt(np.mean(np.sum(dp ** 2, axis=2))))
    # each output sum against the oracle's own (its own shifts: a shift
    # difference moves the skip rule and the warp's pixel truncation too)
    gsums = [g["sums"][i].numpy() for i in range(len(ores["sums"]))]
    sum_l2 = [float(np.linalg.norm(gsums[i] - ores["sums"][i]) / np.linalg.norm(ores["sums"][i]))
              for i in range(len(ores["sums"]))]
    # and against the oracle's correction with the GPU's shifts (the -InAln
    # replay): the sums' own arithmetic, without the discrete effects
    t1 = time.time()
    orep = O.align_movie(host, po, replay_global=g["global_shift"], replay_patch=g["patch_shift"],
                         replay_centers=g["patch_center"], replay_bad=g["patch_bad"])
    dt_rep = time.time() - t1
    sum_l2_rep = [float(np.linalg.norm(gsums[i] - orep["sums"][i]) / np.linalg.norm(orep["sums"][i]))
                  for i in range(len(orep["sums"]))]
    del orep, host
    threads = O.lib().orc_num_threads()
    hi = host_info()
    movie_s = dt * N / frames_needed if frames_needed < N else dt
    cb = dict(value=1.0 / movie_s, unit="movies/s", cores=threads, kind="port",
              sample="FP64 oracle, full path (alignment, correction, three sums) on %s, same "
                     "flags: %.1f s wall%s" % (what, dt, "" if frames_needed >= N else
                                               ", scaled x%.2f to a movie" % (N / frames_needed)),
              host=dict(hi, omp_threads=threads))
    acc = dict(global_rms_vs_fp64_px=rms_g, patch_rms_vs_fp64_px=rms_p,
               global_max_vs_fp64_px=float(np.max(np.abs(dg))),
               patch_max_vs_fp64_px=float(np.max(np.abs(dp))),
               fp64_frames=min(frames_needed, N), same_patch_starts=same,
               same_bad_patches=bool(np.array_equal(g["patch_bad"], ores["patch_bad"])),
               sum_rel_l2_vs_fp64=sum_l2, sum_rel_l2_vs_fp64_replay=sum_l2_rep,
               fp64_replay_s=dt_rep,
               fp64_global_iters=int(ores["global_iters"]),
               fp64_patch_iters_max=int(np.max(ores["patch_iters"])))
    return cb, acc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--kernel", default="fft_cols_fwd", help="kernel class for the roofline")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-frames", type=int, default=N, metavar="K",
                    help="frames of the timed movie the FP64 oracle runs on (default: all 50; "
                         "fewer scale its time to a movie)")
    ap.add_argument("--spectra", default="f32", choices=["f32", "f16", "bf16"],
                    help="storage of the xcf and patch spectrum stacks (arithmetic stays FP32)")
    ap.add_argument("--frames", default="f32", choices=["f32", "f16", "bf16"],
                    help="storage of the full-frame spectra (forward transforms, local-correction "
                         "output, what the sums read); arithmetic stays FP32")
    ap.add_argument("--batch", type=int, default=8,
                    help="movies per step: K > 1 aligns K resident movies in one mc_align_movies "
                         "call (each movie's forward transform overlaps the previous movie's "
                         "alignment; --no-pipeline: their alignment iterations share launches)")
    ap.add_argument("--no-pipeline", action="store_true",
                    help="with --batch K > 1, batch the alignment iterations of the K movies "
                         "instead of pipelining them")
    ap.add_argument("--no-h2d", action="store_true",
                    help="skip timing the host-to-device upload of one movie")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="one plain device-to-device copy of MIB MiB before the steps, as the "
                         "known-size reference for hardware byte counters")
    ap.add_argument("--no-radix", action="store_true",
                    help="let the autotuner time tile variants of the default radix orders only "
                         "(by default it also times the other compiled radix orders)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default FFT tiles instead of timing the compiled variants "
                         "(mc_autotune) before the steps")
    ap.add_argument("--ingest", type=int, default=0, metavar="K",
                    help="also time K movies delivered as 8-bit raw frames in pinned host memory "
                         "(gain-corrected on the GPU by mc_load_frames), each load overlapping "
                         "the previous movie's alignment")
    ap.add_argument("--eer", type=int, default=0, metavar="K",
                    help="also time rendering a synthetic EER movie of K fractions (4096^2 camera, "
                         "about 100k electrons per fraction, 10 fractions per frame) with "
                         "mc_load_eer from pinned host memory")
    ap.add_argument("--mixed", type=int, default=0, metavar="K",
                    help="also time K movies of uneven size in one mc_align_movies call "
                         "(alternately 5760x4092x50 and 4096x4096x40, same flags): with their "
                         "alignment iterations batched over one job table per stage, and "
                         "pipelined size by size")
    ap.add_argument("--prior", type=int, default=0, metavar="K",
                    help="train the learned shift prior on K synthetic movies of the workload "
                         "(labels from the plain path), then time the guided path")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the path is data-parallel over movies "
              "(one movie per GPU); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from motioncor3_amd import lib as L
    from motioncor3_amd.synth import MovieRecipe

    recipe = MovieRecipe(W, H, N, 20000, seed=20250101, dose=1.0, local_max=3.0)
    ctx = L.Context(0)
    frames = ctx.empty((N, H, W), np.float32)
    ctx.synth_render(frames, recipe)
    kw = dict(patch=(5, 5), bft=(500, 100), pix_size=0.83, kv=300, fm_dose=1.0,
              spectra_fmt={"f32": 0, "f16": 1, "bf16": 2}[args.spectra],
              frame_fmt={"f32": 0, "f16": 1, "bf16": 2}[args.frames],
              movie_pipeline=0 if args.no_pipeline else 1)
    prm = L.default_params(**kw)
    ns, ow, oh = ctx.output_geometry(prm, W, H, N)
    sums = [ctx.empty((oh, ow), np.float32) for _ in range(ns)]
    tiles = None
    if not args.no_autotune:
        rep = ctx.autotune(W, H, N, patch=(5, 5), radix=not args.no_radix)
        tiles = [ln.strip() for ln in rep.splitlines() if "->" in ln]
    if args.copy_calib > 0:
        nb = args.copy_calib << 20
        a = ctx.empty((nb // 4,), np.float32)
        b = ctx.empty((nb // 4,), np.float32)
        ctx.copy(b, a)
        ctx.sync()
        del a, b
    K = max(1, args.batch)
    movies, msums = [frames], [sums]
    for k in range(1, K):
        rk = MovieRecipe(W, H, N, 20000, seed=20250101 + k, dose=1.0, local_max=3.0)
        fk = ctx.empty((N, H, W), np.float32)
        ctx.synth_render(fk, rk)
        movies.append(fk)
        msums.append([ctx.empty((oh, ow), np.float32) for _ in range(ns)])

    prior_info = None
    if args.prior > 0:
        # training data: the plain path's own first-iteration features and
        # converged patch shifts on movies of the same workload (other seeds)
        t_p = time.perf_counter()
        fx = ctx.empty((N, H, W), np.float32)
        X, T, pit = [], [], []
        for k in range(args.prior):
            ctx.synth_render(fx, MovieRecipe(W, H, N, 20000, seed=7000 + k, dose=1.0, local_max=3.0))
            rr = ctx.align_movie(fx, prm, sums=sums, prior_data=True)
            X.append(rr["prior_features"])
            T.append(rr["prior_targets"].reshape(rr["prior_targets"].shape[0], -1))
            pit.append(rr["patch_iters"])
            if (k + 1) % 250 == 0:
                print("bench: prior training data %d/%d movies" % (k + 1, args.prior),
                      file=sys.stderr, flush=True)
        del fx
        X, T = np.concatenate(X), np.concatenate(T)
        t_d = time.perf_counter()
        prior = L.Prior(ctx, N, 128, seed=1)
        nv = max(1, len(X) // 8)
        xtr, ttr = ctx.upload(X[nv:]), ctx.upload(T[nv:])
        epochs = 150
        hist = prior.train(xtr, ttr, epochs=epochs, batch=64, lr=5e-4, seed=0)
        yv = prior.forward(ctx.upload(X[:nv]))
        t_t = time.perf_counter()
        prior.attach()
        prm = L.default_params(prior=1, **kw)
        prior_info = {"train_movies": args.prior, "samples": int(len(X)), "epochs": epochs,
                      "params": int(prior.n_params), "data_s": t_d - t_p, "train_s": t_t - t_d,
                      "train_mse": float(hist[-1]),
                      "heldout_mse": float(((yv - T[:nv]) ** 2).mean()),
                      "heldout_mse_zero_correction": float((T[:nv] ** 2).mean()),
                      "plain_patch_iters_mean_on_training_movies": float(np.mean(pit))}

    def step():
        if K == 1:
            return [ctx.align_movie(frames, prm, sums=sums)]
        return ctx.align_movies(movies, prm, sums=msums)

    for _ in range(args.warmup):
        res = step()[0]

    ctx.set_kernel_timing(args.kernel)
    ctx.sync()
    t0 = time.perf_counter()
    iters = []
    for _ in range(args.steps):
        res_all = step()
        res = res_all[0]
        iters.append((res["global_iters"], int(np.max(res["patch_iters"]))))
    # movie 0 (the first movie of a step) is `frames`
    res_first = res_all[0]
    ctx.sync()
    t1 = time.perf_counter()
    kms, klaunch = ctx.kernel_timing()
    ctx.set_kernel_timing(None)
    # the same kernel class with the frame chunks on one stream (not the
    # timed configuration): its launches then run without the second
    # stream's kernels beside them
    iso = L.default_params(chunk_streams=1, **kw)
    ctx.set_kernel_timing(args.kernel)
    ctx.align_movie(frames, iso, sums=sums)
    ctx.sync()
    iso_ms, iso_n = ctx.kernel_timing()
    ctx.set_kernel_timing(None)

    ms = (t1 - t0) * 1000.0 / (args.steps * K)
    g = L.geometry(W, H, 5, 5)
    # frame chunks per pass from the launches counted in the timed calls
    per_launch = algorithmic_bytes(args.kernel, g, max(1, round(klaunch / (args.steps * K))))
    avg_ms = kms / max(klaunch, 1)
    achieved = per_launch / (avg_ms * 1e-3) / 1e9 if klaunch else None

    gi, pi = iters[-1]
    pb = path_bytes(g, gi, pi, ns)
    # the FP64 oracle on the timed movie, before the optional legs below
    # reuse its device buffer
    cpu_leg = None
    if not args.no_cpu:
        print("bench: GPU leg done (%.2f ms/movie); timing the FP64 oracle on %d frames of the "
              "timed movie" % (ms, min(args.cpu_frames, N)), file=sys.stderr, flush=True)
        cpu_leg = cpu_baseline(ctx, frames, res_first, kw, max(2, min(args.cpu_frames, N)))
        print("bench: FP64 oracle done (%s)" % cpu_leg[0]["sample"], file=sys.stderr, flush=True)
    h2d = None
    if not args.no_h2d:
        host = frames.numpy()
        ctx.sync()
        t2 = time.perf_counter()
        frames.upload(host)
        ctx.sync()
        h2d = (time.perf_counter() - t2) * 1000.0
        del host

    ingest = None
    if args.ingest > 0:
        # raw 8-bit counts of two movies in pinned host memory, a gain reference
        raws = []
        for k in range(2):
            rk = MovieRecipe(W, H, N, 20000, seed=20250201 + k, dose=1.0, local_max=3.0)
            ctx.synth_render(frames, rk)
            hk = ctx.host_array((N, H, W), np.uint8)
            hk.a[...] = np.clip(frames.numpy(), 0, 255).astype(np.uint8)
            raws.append(hk)
        ctx.set_refs(np.random.default_rng(11).uniform(0.95, 1.05, (H, W)).astype(np.float32), None)
        bufs = [frames, ctx.empty((N, H, W), np.float32)]
        ctx.sync()
        tl = time.perf_counter()
        ctx.load_frames(raws[0], 0, W, H, N, bufs[0])
        load_ms = (time.perf_counter() - tl) * 1000.0
        K = args.ingest
        ctx.sync()
        t2 = time.perf_counter()
        if len(movies) >= K and K > 1:
            # every load queued at once on the ingest stream, then one
            # pipelined mc_align_movies call: each movie's forward transform
            # waits for its own load only
            flow = "mc_align_movies over %d movies loaded asynchronously" % K
            for m in range(K):
                ctx.load_frames(raws[m % 2], 0, W, H, N, movies[m], async_=True)
            ctx.align_movies(movies[:K], prm, sums=msums[:K])
        else:
            flow = "load of movie m+1 overlapping mc_align_movie of movie m"
            ctx.load_frames(raws[0], 0, W, H, N, bufs[0], async_=True)
            for m in range(K):
                if m + 1 < K:
                    ctx.load_frames(raws[(m + 1) % 2], 0, W, H, N, bufs[(m + 1) % 2], async_=True)
                ctx.align_movie(bufs[m % 2], prm, sums=sums)
        ctx.sync()
        t3 = time.perf_counter()
        raw_b = N * H * W
        # bad-pixel detection and correction on the loaded movie
        ctx.sync()
        tb0 = time.perf_counter()
        bad, nbad, _ = ctx.detect_bad(bufs[0])
        tb1 = time.perf_counter()
        ctx.correct_bad(bufs[0], bad)
        tb2 = time.perf_counter()
        ingest_bad = {"detect_ms": (tb1 - tb0) * 1000.0, "correct_ms": (tb2 - tb1) * 1000.0,
                      "bad_pixels": nbad}
        del bad
        ingest = {"movies": K, "raw_mode": 0, "raw_bytes_per_movie": raw_b,
                  "bad_pixels": ingest_bad,
                  "load_ms": load_ms, "load_gbs": raw_b / (load_ms * 1e-3) / 1e9,
                  "end_to_end_movies_per_s": K / (t3 - t2),
                  "flow": flow,
                  "note": "first load not overlapped; loads go host (pinned) -> device by DMA "
                          "on the ingest stream, then one gain/crop kernel"}
        ctx.set_refs(None, None)
        for hk in raws:
            hk.free()
        del bufs

    mixed = None
    if args.mixed > 0:
        # movies of two sizes; the 4096x4096 ones get 40 frames
        shapes = [(W, H, N) if k % 2 == 0 else (4096, 4096, 40) for k in range(args.mixed)]
        mv = []
        for k, (w, h, n) in enumerate(shapes):
            fk = ctx.empty((n, h, w), np.float32)
            ctx.synth_render(fk, MovieRecipe(w, h, n, 20000, seed=20250301 + k, dose=1.0, local_max=3.0))
            mv.append(fk)
        mixed = {"movies": [list(sh) for sh in shapes]}
        for name, pipe in (("batched", 0), ("pipelined", 1)):
            pm = L.default_params(**dict(kw, movie_pipeline=pipe))
            ctx.align_movies(mv, pm)  # warm-up (plans, buffers, graphs)
            ctx.sync()
            tm = time.perf_counter()
            reps = 3
            for _ in range(reps):
                ctx.align_movies(mv, pm)
            ctx.sync()
            dt = (time.perf_counter() - tm) / reps
            mixed[name] = {"movies_per_s": len(mv) / dt, "ms_per_call": dt * 1000.0}
        mixed["note"] = ("batched: each alignment iteration stage is one launch over the jobs of "
                         "both sizes (one host thread per size); pipelined: the sizes one after "
                         "another, each movie's forward transform beside the previous alignment")
        del mv

    eer = None
    if args.eer > 0:
        from motioncor3_amd.synth import eer_stream_7bit
        cw = ch = 4096
        nfr, per, rng = args.eer, 100000, np.random.default_rng(5)
        streams, nev = [], 0
        for k in range(nfr):
            q = np.unique(rng.integers(0, cw * ch, per))
            streams.append(eer_stream_7bit(q, rng.integers(0, 16, len(q)), cw * ch))
            nev += len(q)
        off = np.cumsum([0] + [len(x) for x in streams]).astype(np.int64)
        host = ctx.host_array((int(off[-1]),), np.uint8)
        host.a[...] = np.concatenate(streams)
        del streams
        nout = nfr // 10
        ist, isz = [10 * k for k in range(nout)], [10] * nout
        efr = ctx.empty((nout, ch, cw), np.float32)
        ctx.load_eer(host, off, cw, ch, 7, 1, efr, int_start=ist, int_size=isz)  # warm-up
        reps = 3
        ctx.sync()
        te = time.perf_counter()
        for _ in range(reps):
            ctx.load_eer(host, off, cw, ch, 7, 1, efr, int_start=ist, int_size=isz)
        ctx.sync()
        e_ms = (time.perf_counter() - te) * 1000.0 / reps
        eer = {"fractions": nfr, "camera": [cw, ch], "electrons": nev, "code_bytes": int(off[-1]),
               "rendered_frames": nout, "ms_per_movie": e_ms,
               "electrons_per_s": nev / (e_ms * 1e-3),
               "note": "7-bit codes in pinned host memory; the time includes their DMA, "
                       "the rendering (one thread per fraction) and the byte-count frames"}
        host.free()
        del efr

    # ground truth of the global shifts: the drift plus the frame mean of
    # the local field (reported shift = -displacement + a constant)
    res0 = res_first
    gs = res0["global_shift"]
    d = recipe.global_truth()
    e = (gs - gs.mean(0)) + (d - d.mean(0))
    truth_rms = float(np.sqrt(np.mean(np.sum(e ** 2, axis=1))))
    d0 = recipe.displacement
    e0 = (gs - gs.mean(0)) + (d0 - d0.mean(0))
    truth_rms_drift_only = float(np.sqrt(np.mean(np.sum(e0 ** 2, axis=1))))
    # patch shifts against the local field at each patch centre, less the
    # frame mean the global alignment already took out (reported shift =
    # -residual displacement + a constant per patch)
    pc, ps = res0["patch_center"], res0["patch_shift"]
    res_x = np.stack([recipe.local_field(pc[:, 0], pc[:, 1], t)[0] - recipe.mean_local(t)[0]
                      for t in range(N)])
    res_y = np.stack([recipe.local_field(pc[:, 0], pc[:, 1], t)[1] - recipe.mean_local(t)[1]
                      for t in range(N)])
    ep = ps + np.stack([res_x, res_y], axis=2)
    ep = ep - ep.mean(0, keepdims=True)
    patch_truth_rms = float(np.sqrt(np.mean(np.sum(ep ** 2, axis=2))))

    out = {
        "metric": METRIC,
        "value": 1000.0 / ms,
        "unit": "movies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms * K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "5760x4092 px x 50 frames, 5x5 patches, Bft 500/100, PixSize 0.83, "
                        "kV 300, FmDose 1.0 (sum + DW + DWS), Iter 15, Tol 0.1, Group 1 4",
            "movie": "20k Gaussian blobs, 1 e-/px/frame Poisson, drift + 3 px local field",
            "iterations_global_patch": iters[-1],
            "spectra_storage": args.spectra,
            "movies_per_call": K,
            "movie_pipeline": K > 1 and not args.no_pipeline,
            "spectra_storage_used": [["f32", "f16", "bf16"][v] for v in res["spectra_fmt_used"]],
            "frame_storage_used": ["f32", "f16", "bf16"][res["frame_fmt_used"]],
            "learned_prior": args.prior > 0,
            "fft_tiles": tiles if tiles is not None else "defaults",
        },
        "roofline": {
            "kernel": args.kernel,
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": (achieved / PEAK_HBM_GBS) if achieved else None,
            "traffic": measured_traffic(args.kernel),
            "algorithmic_bytes_per_launch": per_launch,
            "avg_launch_ms": avg_ms,
            "launches": klaunch,
            "concurrency": "frame chunks alternate over 2 streams; a launch's time includes "
                           "sharing the GPU with the other stream's kernels",
            "isolated": {"avg_launch_ms": iso_ms / max(iso_n, 1),
                         "achieved": per_launch / (iso_ms / max(iso_n, 1) * 1e-3) / 1e9,
                         "frac": per_launch / (iso_ms / max(iso_n, 1) * 1e-3) / 1e9 / PEAK_HBM_GBS,
                         "note": "one extra call with chunk_streams=1, outside the timed steps"},
            "path_bytes_per_movie": pb,
            "path_achieved": pb / (ms * 1e-3) / 1e9,
            "path_frac": pb / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
        },
        # per movie, averaged over the movies of the last timed call (the
        # first one transforms its frames itself; the others wait on the
        # forward transform that ran beside the previous movie)
        "stages_ms": {k: 1000.0 * float(np.mean([r["seconds_stage"][k] for r in res_all]))
                      for k in res["seconds_stage"]},
        "stages_ms_first_movie": {k: v * 1000.0 for k, v in res["seconds_stage"].items()},
        "accuracy": {"global_shift_rms_vs_truth_px": truth_rms,
                     "patch_shift_rms_vs_truth_px": patch_truth_rms,
                     "global_shift_rms_vs_drift_only_px": truth_rms_drift_only,
                     "truth": "global drift plus the frame mean of the local field"},
    }
    if ingest is not None:
        out["ingest"] = ingest
    if eer is not None:
        out["eer"] = eer
    if mixed is not None:
        out["mixed"] = mixed
    if prior_info is not None:
        prior_info["fallbacks_last_step"] = res["prior_fallbacks"]
        prior_info["patch_iters_mean_last_step"] = float(np.mean(res["patch_iters"]))
        out["prior"] = prior_info
    if h2d is not None:
        out["h2d"] = {"ms_per_movie": h2d, "pageable_host_memory": True,
                      "pcie_inclusive_movies_per_s": 1000.0 / (ms + h2d)}
    if cpu_leg is not None:
        out["cpu_baseline"], acc = cpu_leg
        out["accuracy"].update(acc)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
