"""Benchmark of the per-frame measure path on one MI355X.

A frame step is the source's frame and every active encode's frame through CAMBI (all scales, top-k), the window
energies, the lasting texture, PSNR and the mean colours (frames with chroma planes, read in place) and the collect.
A step (--steps, --warmup) is one batch of frame steps, one round of the batched form.  Inputs are synthetic and
generated into device memory before the timed region.  `value` is the batched form's frame steps per second (graph
replay, integer energies); the plain per-frame-launch FP64 form is timed on the same frames in the same process.  The
CPU baseline is the oracle run on the host cores over a bounded sample of the same frames.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1..5]
"""
import argparse
import importlib
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "scored frames/s per encode, 1080p & 4K 10-bit; % HBM roofline; batched vs plain"
HBM_PEAK_GBS = 8000.0


def config(n, es, frames):
    """(source desc, [(encode desc, length)], source bytes, shift, workload text) of configuration n."""
    mk = lambda **kw: es.SynthDesc(**kw)
    seed = 0x5EED0000 + n
    if n == 1:
        w, h = 1920, 1080
        src = mk(kind=0, seed=seed, full_w=w, full_h=h, w=w, h=h, depth=10, grain_q8=256)
        enc = [(mk(kind=0, seed=seed, salt=1, full_w=w, full_h=h, w=w, h=h, depth=10, box3=1), frames)]
        return src, enc, 2, 0, "C1 plain: 1 source + 1 encode, 1920x1080 10-bit ramp + grain vs 4-code banded box3"
    if n == 2:
        w, h = 3840, 2160
        src = mk(kind=1, seed=seed, full_w=w, full_h=h, w=w, h=h, depth=10, grain_q8=256)
        enc = [(mk(kind=1, seed=seed, salt=1, full_w=w, full_h=h, w=w, h=h, depth=10, grain_q8=0, quant_step=2,
                   quant_where=2, block=8, block_pct=50), frames)]
        return src, enc, 2, 0, "C2: 1 source + 1 encode, 3840x2160 10-bit tile mosaic, encode deblocked/degrained/dark-quantised"
    if n == 3:
        w, h = 1920, 1080
        src = mk(kind=1, seed=seed, full_w=w, full_h=h, w=w, h=h, depth=10, grain_q8=256)
        crops = [(1920, 1080), (1920, 800), (1920, 1040), (1800, 1080)]
        dist = [dict(blur_q8=154, grain_q8=256), dict(quant_step=2, grain_q8=256), dict(grain_q8=0),
                dict(block=16, block_pct=50, grain_q8=256)]
        lens = [1.0, 1100 / 1440, 720 / 1440, 333 / 1440]
        enc = []
        for i, ((cw, ch), d, l) in enumerate(zip(crops, dist, lens)):
            cx, cy = ((w - cw) // 2) & ~1, ((h - ch) // 2) & ~1
            enc.append((mk(kind=1, seed=seed, salt=i, full_w=w, full_h=h, crop_x=cx, crop_y=cy, w=cw, h=ch, depth=10, **d),
                        max(13, int(frames * l))))
        return src, enc, 2, 0, "C3: 1920x1080 source + 4 encodes of uneven crops and lengths"
    if n == 4:
        w, h = 1920, 1080
        src = mk(kind=1, seed=seed, full_w=w, full_h=h, w=w, h=h, depth=8, grain_q8=256)
        dist = [dict(grain_q8=128), dict(quant_step=4, quant_where=1, grain_q8=256), dict(block=8, block_pct=25, grain_q8=256)]
        enc = [(mk(kind=1, seed=seed, salt=i, full_w=w, full_h=h, w=w, h=h, depth=10, **d), frames) for i, d in enumerate(dist)]
        return src, enc, 1, 2, "C4: 8-bit 1920x1080 source (1-byte samples, shift 2) + 3 10-bit encodes, graph replay"
    if n == 5:
        w, h = 3840, 1600
        src = mk(kind=1, seed=seed, full_w=w, full_h=h, w=w, h=h, depth=10, grain_q8=256)
        rng = np.random.default_rng(seed)
        enc = []
        for i in range(8):
            d = dict(blur_q8=int(256 * rng.uniform(0.3, 1.2)), quant_step=int(rng.choice([1, 2, 4])),
                     grain_q8=int(256 * rng.uniform(0, 1)))
            length = max(13, int(frames * (1440 - i * (940 / 7)) / 1440))
            enc.append((mk(kind=1, seed=seed, salt=i, full_w=w, full_h=h, w=w, h=h, depth=10, **d), length))
        return src, enc, 2, 0, "C5: 3840x1600 source + 8 encodes, uneven lengths"
    raise SystemExit(f"no configuration {n}")


STAGE_KERNEL = {"cambi_cvalues": "void cambi_cvalues_kernel<{layout}>", "cambi_base": "cambi_base_kernel",
                "window_stats": "void window_stats_int_kernel<16>", "persist": "void lasting_kernel<16>",
                "frame_sums": "frame_sums_kernel"}
# rocprofv3 --pmc passes of each workload (FETCH_SIZE, WRITE_SIZE and the SQ counters of the dominant kernel),
# summarised per dispatch; see each file's "command"
PMC_FILES = {(1, 1920, 1080): "pmc_1920x1080_ramp.json", (2, 3840, 2160): "pmc_3840x2160_mosaic.json",
             (3, 1920, 1080): "pmc_1920x1080_crops_4enc.json", (4, 1920, 1080): "pmc_1920x1080_8bit_3enc.json"}
PMC_MISSING = {5: "no PMC summary: rocprofv3's FETCH_SIZE and WRITE_SIZE passes over this workload stopped printing "
                  "while the run generated its frames and were killed; the kernel trace alone is in "
                  "profiles/measure_path_3840x1600_mosaic_8enc.md"}


def kernel_name(stage, bmax, mfma):
    layout = f"CvSharedM<{1 if bmax == 32 else 2}> " if mfma else f"CvShared<{bmax}> "
    return STAGE_KERNEL[stage].format(layout=layout)


def pmc_profile(cfg, w, h, kernel):
    """(per-dispatch record of `kernel`, the profiled run's input bytes per launch, file name) from the committed PMC
    summary of this workload, or (None, None, reason)."""
    name = PMC_FILES.get((cfg, w, h))
    if name is None:
        return None, None, PMC_MISSING.get(cfg, "no rocprofv3 PMC summary of this workload")
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            prof = json.load(f)
        return prof["kernels"][kernel], prof["b_in_per_launch"], f"profiles/{name}"
    except (OSError, KeyError, ValueError):
        return None, None, f"profiles/{name} holds no entry for {kernel}"


def host_cpu():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return f"{line.split(':', 1)[1].strip()}, {os.cpu_count()} logical CPUs visible"
    except OSError:
        pass
    return None


def load_oracle():
    spec = importlib.util.spec_from_file_location("es_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


TUNE_BATCH_MAX = 64               # the largest batch the autotuner tries
TUNE_FRAMES = 3 * TUNE_BATCH_MAX  # it times three batches of it
PLAIN_WARM = 16                   # untimed frame steps of the plain form (lasting texture starts at frame 12)


def frame_pool(steps, warmup, tune, batch, cap):
    """Distinct frames generated per file: up to `cap` the run's frame steps (the largest batch the run may take times
    its steps), and with the autotuner its frames at least; later frame steps reuse the pool's frames in turn."""
    most_batch = TUNE_BATCH_MAX if tune else batch
    run = min(cap, (warmup + steps) * most_batch)
    return max(TUNE_FRAMES, run) if tune else run


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8,
                    help="timed steps; a step is one batch of frame steps (the autotuner's batch, or --batch)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed steps before the timed ones")
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--batch", type=int, default=32, help="frame steps per step (with --tune 0)")
    ap.add_argument("--tile-h", type=int, default=0, help="c-value tile height (with --tune 0)")
    ap.add_argument("--tile-h0", type=int, default=0, help="c-value tile height at full resolution, 0: --tile-h (with --tune 0)")
    ap.add_argument("--cv-bmax", type=int, default=0, help="c-value LDS layout, 32 or 64 (with --tune 0)")
    ap.add_argument("--cv-waves", type=int, default=0, help="persistent c-value waves per CU (with --tune 0)")
    ap.add_argument("--cv-tile-blocks", type=int, default=0, help="c-value kernel as one block per tile (with --tune 0)")
    ap.add_argument("--cv-mfma", type=int, default=0,
                    help="c-value window counts on the matrix cores (with --tune 0)")
    ap.add_argument("--cv-queues", type=int, default=0,
                    help="c-value tile queues: one shared queue (0) or one per XCD (1) (with --tune 0)")
    ap.add_argument("--early-energies", type=int, default=0,
                    help="start the energy chain with the CAMBI pyramid rather than after it (with --tune 0)")
    ap.add_argument("--plain-steps", type=int, default=48, help="frame steps of the plain form timed")
    ap.add_argument("--cpu-frames", type=int, default=-1, help="frames of the CPU baseline sample (-1: per config)")
    ap.add_argument("--pool", type=int, default=256,
                    help=f"distinct frames generated per file at most (at least {TUNE_FRAMES}); frame steps beyond "
                         "the pool reuse its frames in turn")
    ap.add_argument("--solo-rounds", type=int, default=2,
                    help="rounds of the serial run that times each kernel alone (0: none; the roofline kernel is then "
                         "the CAMBI chain's longest)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", type=int, default=1, help="run the autotuner before the timed region (1) or not (0)")
    ap.add_argument("--tune-cache", default=None, help="JSON file to keep autotuner results in")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run the calibration copy kernel over this many MiB (for HBM byte counters)")
    ap.add_argument("--serial", type=int, default=0,
                    help="run each round's CAMBI and energy chains one after the other (1) instead of overlapped (0)")
    ap.add_argument("--chroma", type=int, default=1,
                    help="frames carry chroma planes: PSNR and mean colours in every step (1) or luma only (0)")
    ap.add_argument("--cache-run", type=int, default=1,
                    help="also time a second run whose source CAMBI comes from the source-CAMBI cache")
    ap.add_argument("--repeats", type=int, default=3,
                    help="timed runs of the batched form, each a fresh scorer with the same warm-up and frame steps; "
                         "value is the first, the line also gives all of them and their median")
    ap.add_argument("--ingest", type=int, default=50,
                    help="time this many launches of the frame-ingest kernel on a P010 frame of the source's size")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py runs on one GPU only: this path has no multi-GPU form here (asked for {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        raise SystemExit("--steps must be at least 1 and --warmup at least 0")

    es = importlib.import_module("encode-score_amd")
    es.init(0)
    # frames per file: enough for the autotuner's trials and, up to --pool, for the run; later steps reuse them
    pool = frame_pool(args.steps, args.warmup, args.tune, args.batch, args.pool)
    src_d, encs, sbytes, sshift, workload = config(args.config, es, pool)
    W, H = src_d.w, src_d.h
    print(f"[bench] config {args.config}: {workload}; generating {pool} frames per file", file=sys.stderr, flush=True)

    def gen(desc, nbytes):
        frames = []
        for f in range(pool):
            d = es.DeviceArray((desc.h, desc.w), np.uint8 if nbytes == 1 else np.uint16)
            es.synth_frame(desc, f, d, bytes=nbytes)
            frames.append(d)
        return frames

    if args.copy_mib:
        nbytes = args.copy_mib << 20
        a, b = es.DeviceArray(nbytes, np.uint8), es.DeviceArray(nbytes, np.uint8)
        for _ in range(3):
            es.copy_kernel(b, a, nbytes)
        es.synchronize()
        del a, b

    def chroma_desc(desc, salt):
        """A half-size picture of the same generator for a chroma plane."""
        d = es.SynthDesc()
        for name, _ in type(d)._fields_:
            setattr(d, name, getattr(desc, name))
        d.w, d.h, d.full_w, d.full_h = desc.w // 2, desc.h // 2, desc.full_w // 2, desc.full_h // 2
        d.crop_x, d.crop_y, d.salt = desc.crop_x // 2, desc.crop_y // 2, desc.salt + salt
        return d

    src_frames = gen(src_d, sbytes)
    enc_frames = [gen(d, 2) for d, _ in encs]
    src_chroma = [gen(chroma_desc(src_d, 100 * k), sbytes) for k in (1, 2)] if args.chroma else None
    enc_chroma = [[gen(chroma_desc(d, 100 * k), 2) for k in (1, 2)] for d, _ in encs] if args.chroma else None
    es.synchronize()
    src_fmt = "yuv420p" if sbytes == 1 else "yuv420p10"

    def ptrs(f0, n, encodes=None):
        idx = range(len(encs)) if encodes is None else encodes
        s = [src_frames[f % pool].ptr for f in range(f0, f0 + n)]
        e = [[enc_frames[i][f % pool].ptr for i in idx] for f in range(f0, f0 + n)]
        return s, e

    def yuv(f0, n, encodes=None):
        """Frames with chroma of steps f0 .. f0 + n - 1 (source, and rows of the chosen encodes)."""
        idx = range(len(encs)) if encodes is None else encodes
        s = [es.frame_view(src_fmt, src_frames[f % pool], src_chroma[0][f % pool], src_chroma[1][f % pool],
                           W * sbytes, W // 2 * sbytes) for f in range(f0, f0 + n)]
        e = [[es.frame_view("yuv420p10", enc_frames[i][f % pool], enc_chroma[i][0][f % pool],
                            enc_chroma[i][1][f % pool], encs[i][0].w * 2, encs[i][0].w) for i in idx]
             for f in range(f0, f0 + n)]
        return s, e

    def submit(scorer, f0, n, stream=None, pts=None, encodes=None):
        if n <= 0:
            return
        if args.chroma:
            scorer.submit_frames(*yuv(f0, n, encodes), stream, pts=None if pts is None else pts[f0:f0 + n])
        elif pts is not None:
            idx = range(len(encs)) if encodes is None else encodes
            s_ = [es.frame_view(src_fmt, src_frames[f % pool], None, None, W * sbytes, W) for f in range(f0, f0 + n)]
            e_ = [[es.frame_view("yuv420p10", enc_frames[i][f % pool], None, None, encs[i][0].w * 2, encs[i][0].w)
                   for i in idx] for f in range(f0, f0 + n)]
            scorer.submit_frames(s_, e_, stream, pts=pts[f0:f0 + n])
        else:
            scorer.submit(*ptrs(f0, n, encodes), stream)

    def timed(scorer, f0, n):
        es.synchronize()
        start, stop = es.Event(), es.Event()
        stream = es.Stream()
        t0 = time.perf_counter()
        start.record(stream)
        submit(scorer, f0, n, stream)
        scorer.sync()
        stop.record(stream)
        stream.synchronize()
        es.synchronize()
        wall = time.perf_counter() - t0
        return wall, start.elapsed_ms(stop) / 1e3

    # the autotuner, on the pool's first frames, outside the timed region; if it cannot run, the defaults
    tuned = None
    batch, tile_h, cv_bmax, cv_waves, cv_tb = args.batch, args.tile_h, args.cv_bmax, args.cv_waves, bool(args.cv_tile_blocks)
    tile_h0 = args.tile_h0
    early, mfma, queues = bool(args.early_energies), bool(args.cv_mfma), args.cv_queues
    if args.tune:
        tune_descs = [es.EncodeDesc(d.w, d.h, pool, d.w, 2, 0) for d, _ in encs]
        s_t, e_t = ptrs(0, TUNE_FRAMES)
        t0 = time.perf_counter()
        try:
            tuned = es.autotune(W, H, tune_descs, s_t, e_t, src_bytes=sbytes, src_shift=sshift, cache=args.tune_cache,
                                serial_stages=bool(args.serial))
            batch, tile_h, cv_bmax, cv_waves = tuned["batch"], tuned["tile_h"], tuned["cv_bmax"], tuned["cv_waves"]
            cv_tb, early, mfma = tuned["cv_tile_blocks"], tuned["early_energies"], tuned["cv_mfma"]
            queues, tile_h0 = tuned.get("cv_queues", 0), tuned.get("tile_h0", 0)
        except es.EsError as exc:
            tuned = {"fallback": f"the defaults and the command line's settings: the autotuner failed ({exc})"}
        tuned["seconds"] = round(time.perf_counter() - t0, 2)
        print(f"[bench] autotuner: {tuned}", file=sys.stderr, flush=True)

    # the run: warmup + steps batches of `batch` frame steps; encode lengths follow the configuration's proportions
    warm, timed_frames = args.warmup * batch, args.steps * batch
    total = warm + timed_frames
    src_d, encs, sbytes, sshift, workload = config(args.config, es, total)
    descs = [es.EncodeDesc(d.w, d.h, length, d.w, 2, 0) for d, length in encs]
    lengths = [length for _, length in encs]

    # batched form (graph replay), with live per-stage events
    def batched(**kw):
        return es.Scorer(es.ES_BATCHED, W, H, kw.pop("descs", descs), batch=batch, tile_h=tile_h, tile_h0=tile_h0,
                         use_graph=True, src_bytes=sbytes, src_shift=sshift, cv_bmax=cv_bmax, cv_waves=cv_waves,
                         cv_tile_blocks=cv_tb, early_energies=early, cv_mfma=mfma, cv_queues=queues, **kw)

    bat = batched(profile=True, serial_stages=bool(args.serial))
    submit(bat, 0, warm)
    bat.sync()
    warm_stage, warm_rounds = bat.stage_ms()
    print(f"[bench] warmup done ({args.warmup} steps of {batch} frame steps)", file=sys.stderr, flush=True)
    wall, dev = timed(bat, warm, timed_frames)
    stage, rounds = bat.stage_ms()
    stage = {k: stage[k] - warm_stage[k] for k in stage}
    rounds -= warm_rounds
    frame_steps_per_s = timed_frames / wall
    # the repeats: the same settings, warm-up and timed frame steps, each on a scorer of its own
    repeat_values = [round(frame_steps_per_s, 3)]
    for _ in range(max(args.repeats, 1) - 1):
        again = batched(profile=True, serial_stages=bool(args.serial))
        submit(again, 0, warm)
        again.sync()
        repeat_values.append(round(timed_frames / timed(again, warm, timed_frames)[0], 3))
        del again
    encode_frames = sum(min(length, total) - min(length, warm) for length in lengths)
    print(f"[bench] batched: {frame_steps_per_s:.1f} frame steps/s ({wall * 1e3 / args.steps:.3f} ms per step of "
          f"{batch})", file=sys.stderr, flush=True)

    # The input bytes of the timed frame steps (the algorithmic bytes of the path): each luma plane read once, the
    # source once per step and every active encode's plane (B_in); with chroma, 1.5 times that.
    def active(f):
        return [d for d, length in encs if f < length]

    def b_in(f0, f1):
        return sum((sbytes * W * H + sum(2 * d.w * d.h for d in active(f))) if active(f) else 0 for f in range(f0, f1))

    b_in_timed = b_in(warm, total)
    launches = max(rounds, 1)
    # each kernel's own bytes per launch (each input plane read once, each output written once), for comparison
    scale_px = lambda n: n * (1 + 1 / 4 + 1 / 16 + 1 / 64 + 1 / 256)
    kernel_bytes = dict.fromkeys(STAGE_KERNEL, 0.0)
    crops_of = lambda f: {(d.w, d.h) for d in active(f)}
    for f in range(warm, total):
        src_px = sum(cw * ch for cw, ch in crops_of(f))
        enc_px = sum(d.w * d.h for d in active(f))
        if not enc_px:
            continue
        kernel_bytes["cambi_cvalues"] += 2 * scale_px(src_px + enc_px)
        kernel_bytes["cambi_base"] += sbytes * src_px + 2 * enc_px + 2 * (src_px + enc_px)
        kernel_bytes["window_stats"] += sbytes * src_px + 2 * enc_px + 4 * (src_px + enc_px)
        kernel_bytes["persist"] += 4 * (src_px + enc_px) + 4 * 17 * src_px / batch
        kernel_bytes["frame_sums"] += 1.5 * (sbytes * src_px + 2 * enc_px) if args.chroma else 0
    kernel_bytes = {k: v / launches for k, v in kernel_bytes.items()}

    # Which kernel dominates, from each kernel's time alone on the GPU.  With the chains overlapped, a stage of the
    # energy chain (on the side stream) waits there for the CUs the c-value kernel's persistent waves hold, so its
    # event span is not its kernel time: a short serial run of the same scorer configuration (one untimed round,
    # then --solo-rounds, stage events around each kernel) gives the solo times.  The roofline then takes the
    # dominant kernel's mean launch time from the timed region when it runs on the main stream there (the CAMBI
    # chain, or every stage when serial), and from the serial run otherwise.
    if args.serial:
        solo, solo_launches = stage, launches
    elif args.solo_rounds <= 0:  # no serial run: the main stream's stages only
        solo = {k: (stage[k] if k.startswith("cambi_") else 0.0) for k in stage}
        solo_launches = launches
    else:
        cal = batched(profile=True, serial_stages=True)
        submit(cal, 0, batch)
        cal.sync()
        s0, r0 = cal.stage_ms()
        submit(cal, batch, args.solo_rounds * batch)
        cal.sync()
        s1, r1 = cal.stage_ms()
        solo = {k: s1[k] - s0[k] for k in s1}
        solo_launches = max(r1 - r0, 1)
        del cal
    dominant = max(STAGE_KERNEL, key=lambda k: solo[k])
    on_main = args.serial or dominant.startswith("cambi_")
    mean_ms = stage[dominant] / launches if on_main else solo[dominant] / solo_launches
    b_in_launch = b_in_timed / launches
    achieved = b_in_launch / (mean_ms / 1e3) / 1e9
    kname = kernel_name(dominant, cv_bmax or 32, mfma)
    rec, prof_b_in, prof_src = pmc_profile(args.config, W, H, kname)
    roofline = {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None,
                "algorithmic_bytes_per_launch": int(b_in_launch),
                "algorithmic_bytes": "the luma planes of the launch's frame steps, each read once (source once per "
                                     "step, every active encode)",
                "frame_steps_per_launch": round(timed_frames / launches, 3),
                "mean_launch_ms": round(mean_ms, 5),
                "timing": "HIP events on the scorer's work stream over the timed region" if on_main else
                          "HIP events on the scorer's work stream, serial run (the kernel alone)",
                "kernel_bytes_per_launch": int(kernel_bytes[dominant]),
                "kernel_bytes_frac": round(kernel_bytes[dominant] / (mean_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 6),
                "kernel_bytes": "the kernel's own inputs read once and outputs written once (for the c-value kernel "
                                "the packed planes of all five scales)",
                "solo_launch_ms": {k: round(solo[k] / solo_launches, 5) for k in STAGE_KERNEL},
                "counters_source": prof_src}
    if rec is not None:
        # HBM bytes per dispatch from the PMC passes (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel
        # calibration), scaled by this run's input bytes per launch over the profiled run's (a kernel's traffic goes
        # with the pictures of its launch)
        scale = b_in_launch / prof_b_in
        traffic = rec["bytes"] * scale
        roofline["traffic"] = int(traffic)
        roofline["traffic_over_algorithmic"] = round(traffic / b_in_launch, 3)
        roofline["traffic_over_kernel_bytes"] = round(traffic / kernel_bytes[dominant], 3)
        hbm_frac = traffic / (mean_ms / 1e3) / (HBM_PEAK_GBS * 1e9)
        roofline["traffic_frac"] = round(hbm_frac, 4)
        derived = rec.get("derived", {})
        if "mfma_i8_ops" in rec:
            mfma_frac = rec["mfma_i8_ops"] * scale / (mean_ms / 1e3) / 5.0e15
            roofline["mfma_i8_frac"] = round(mfma_frac, 5)
            roofline["bound"] = "mfma" if mfma_frac > hbm_frac else "hbm"
        roofline["counters"] = derived
    step_gbs = b_in_timed / wall / 1e9

    result = {
        "metric": METRIC,
        "value": round(frame_steps_per_s, 3),
        "unit": "frame steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic",
        "config": {"workload": f"{workload}; encode lengths {'/'.join(map(str, lengths))} frame steps (the run's "
                               f"length, {total} frame steps, in the configuration's proportions); {pool} distinct "
                               f"frames per file, reused in turn",
                   "config": args.config, "width": W, "height": H, "encodes": len(encs),
                   "step": f"one batch of {batch} frame steps (each the source's frame and every active encode's)",
                   "frame_steps_per_step": batch, "tile_h": tile_h or 32, "tile_h0": tile_h0 or tile_h or 32,
                   "cv_bmax": cv_bmax or 32,
                   "cv_waves": "one block per tile" if cv_tb else (cv_waves or 16),
                   "cv_method": "matrix cores" if mfma else "prefix scans",
                   "cv_queues": "one block per tile" if cv_tb else ("one per XCD, an eighth of every scale each"
                                                                     if queues > 0 else "one shared queue"),
                   "chroma": bool(args.chroma), "stages": "serial" if args.serial else
                   ("energies overlapped from the pyramid" if early else "energies overlapped from the c-values"),
                   "form": "batched, integer energies, graph replay"},
        "autotuner": tuned,
        "repeats": {"frame_steps_per_s": repeat_values, "median": float(np.median(repeat_values)),
                    "runs": "each a scorer of its own with the same warm-up and timed frame steps; value is the first"},
        "encode_frames_per_s": round(encode_frames / wall, 3),
        "device_ms_per_step": round(dev * 1e3 / args.steps, 4),
        "hbm_frac_b_in": round(step_gbs / HBM_PEAK_GBS, 6),
        "hbm_frac_b_in_with_chroma": round(1.5 * step_gbs / HBM_PEAK_GBS, 6) if args.chroma else None,
        "stage_ms_per_frame_step": {k: round(v / timed_frames, 5) for k, v in stage.items()},
        "roofline": roofline,
    }

    # the source-CAMBI cache (src/cache.cpp): a first run against the source fills it, a second run reads every
    # source CAMBI from it and skips computing it; the second run is timed like the first
    if args.cache_run:
        import tempfile
        with tempfile.TemporaryDirectory() as tmp:
            source = os.path.join(tmp, "source")
            with open(source, "wb") as fh:
                fh.write(b"synthetic source" * 4096)
            pts = np.arange(total, dtype=np.int64) * 1001
            for k in range(2):
                sc = batched(serial_stages=bool(args.serial))
                sc.use_cache(source, tmp)
                submit(sc, 0, warm, pts=pts)
                sc.sync()
                es.synchronize()
                t0 = time.perf_counter()
                submit(sc, warm, timed_frames, pts=pts)
                sc.sync()
                es.synchronize()
                cwall = time.perf_counter() - t0
                hits = sc.cache_hits()
                cam_cached = sc.cambi(0, min(lengths[0], total))
                del sc
            ref_cam = bat.cambi(0, min(lengths[0], total))
            result["source_cambi_cached"] = {
                "value": round(timed_frames / cwall, 3), "unit": "frame steps/s", "hits": hits,
                "speedup": round(wall / cwall, 3),
                "equal_figures": bool(np.array_equal(cam_cached[0], ref_cam[0]) and np.array_equal(cam_cached[1], ref_cam[1]))}
            print(f"[bench] source CAMBI cached: {timed_frames / cwall:.1f} frame steps/s", file=sys.stderr, flush=True)

    # plain per-frame-launch FP64 form on the same frames
    if not args.no_plain:
        n_plain = args.plain_steps
        n_run = PLAIN_WARM + n_plain
        pl = es.Scorer(es.ES_PLAIN, W, H, [es.EncodeDesc(d.w, d.h, min(d.length, n_run), d.pitch, 2, 0) for d in descs],
                       src_bytes=sbytes, src_shift=sshift)
        submit(pl, 0, PLAIN_WARM)
        pl.sync()
        pwall, _ = timed(pl, PLAIN_WARM, n_plain)
        plain_sps = n_plain / pwall
        result["plain_fp64"] = {"value": round(plain_sps, 3), "unit": "frame steps/s", "frame_steps": n_plain,
                                "ms_per_frame_step": round(pwall * 1e3 / n_plain, 4)}
        result["batched_over_plain"] = round(frame_steps_per_s / plain_sps, 3)
        # the two forms must agree exactly on the frames both scored
        n_both = [min(d.length, n_run, total) for d in descs]
        result["plain_equals_batched_cambi"] = all(
            all(np.array_equal(a, b) for a, b in zip(bat.cambi(i, n), pl.cambi(i, n))) for i, n in enumerate(n_both))
        if args.chroma:
            result["plain_equals_batched_psnr"] = all(
                bool(np.array_equal(bat.psnr(i, n), pl.psnr(i, n))) for i, n in enumerate(n_both))
        print(f"[bench] plain: {plain_sps:.2f} frame steps/s", file=sys.stderr, flush=True)

    # CPU baseline: the oracle on the host cores over a bounded sample of the same frames (encode 0, timed), then the
    # oracle's figures for every encode over the same sample against the GPU's
    if not args.no_cpu:
        orc = load_oracle()
        cpu_frames = args.cpu_frames if args.cpu_frames >= 0 else (16 if W * H > 4e6 else 24)
        cpu_frames = min(cpu_frames, pool, total)
        threads = min(16, os.cpu_count() or 1)
        runs, cams, psnrs = [], [], []
        cpu_t = 0.0
        for i, (d, length) in enumerate(encs):
            n = min(cpu_frames, length)
            cx, cy = ((W - d.w) // 2) & ~1, ((H - d.h) // 2) & ~1
            run = orc.Run(d.w, d.h, threads=threads)
            crop = lambda a, k: a[cy // k:cy // k + d.h // k, cx // k:cx // k + d.w // k]
            cam, psnr_cpu = [], []
            if args.chroma:
                frames = [((crop(src_frames[f].host().astype(np.uint16) << sshift, 1),
                            crop(src_chroma[0][f].host().astype(np.uint16) << sshift, 2),
                            crop(src_chroma[1][f].host().astype(np.uint16) << sshift, 2)),
                           (enc_frames[i][f].host(), enc_chroma[i][0][f].host(), enc_chroma[i][1][f].host()))
                          for f in range(n)]
                t0 = time.perf_counter()
                for s_, e_ in frames:
                    a_, b_, p_ = run.frame_yuv(s_, e_)
                    cam.append((a_, b_))
                    psnr_cpu.append(p_)
            else:
                frames = [(crop(src_frames[f].host().astype(np.uint16) << sshift, 1), enc_frames[i][f].host())
                          for f in range(n)]
                t0 = time.perf_counter()
                for s_, e_ in frames:
                    cam.append(run.frame(s_, e_))
            if i == 0:
                cpu_t = time.perf_counter() - t0
            runs.append(run)
            cams.append(np.array(cam))
            psnrs.append(psnr_cpu)
        n0 = min(cpu_frames, lengths[0])
        pair_s = n0 / cpu_t
        result["cpu_baseline"] = {"value": round(pair_s / len(encs), 4), "unit": "frame steps/s", "cores": threads,
                                  "kind": "port", "host_cpu": host_cpu(),
                                  "sample": f"the oracle's whole per-pair path (CAMBI of both pictures, energies, "
                                            f"lasting texture, collect) over the first {n0} frames of encode 0 "
                                            f"({encs[0][0].w}x{encs[0][0].h}), {cpu_t:.1f} s; the per-step rate "
                                            f"divides by the {len(encs)} encode(s)"}
        # the batched form's CAMBI of every encode over the sample equals the oracle's bit for bit
        result["cambi_bit_exact_vs_oracle"] = all(
            np.array_equal(np.stack(bat.cambi(i, len(cams[i])), axis=1), cams[i]) for i in range(len(encs)))
        # the batched form's scores of every encode over the same sample against the oracle's (FP64 formulas on exact
        # sums)
        chk = batched(descs=[es.EncodeDesc(d.w, d.h, len(cams[i]), d.w, 2, 0) for i, (d, _) in enumerate(encs)])
        submit(chk, 0, max(len(c) for c in cams))
        chk.sync()

        def err(a, b):
            if np.isnan(a) and np.isnan(b):
                return 0.0
            return abs(a - b) if not (np.isnan(a) or np.isnan(b)) else float("inf")

        scalars = ("cambi", "cambi_source", "texture", "banding", "grain", "score")
        errs = {k: 0.0 for k in scalars}
        lasting_equal, psnr_exact = True, True
        for i in range(len(encs)):
            got, ref = chk.segment(i, -1), orc.accum_summary(runs[i].merged())[0]
            for k in scalars:
                errs[k] = max(errs[k], err(got[k], ref[k]))
            lasting_equal &= bool(got["n_lasting"] == ref["n_lasting"])
            if args.chroma:
                errs["psnr"] = max(errs.get("psnr", 0.0), max(err(a, b) for a, b in zip(got["psnr"], ref["psnr"])))
                errs["color"] = max(errs.get("color", 0.0), max(
                    err(a, b) for a, b in zip(got["color_source"] + got["color_encode"],
                                              ref["color_source"] + ref["color_encode"])))
                psnr_exact &= bool(np.array_equal(chk.psnr(i, len(cams[i])), np.array(psnrs[i])))
        errs["lasting_windows_equal"] = lasting_equal
        errs["encodes_checked"] = len(encs)
        errs["frames_per_encode"] = [len(c) for c in cams]
        result["score_err_vs_oracle"] = errs
        if args.chroma:
            result["psnr_bit_exact_vs_oracle"] = psnr_exact
        print(f"[bench] cpu: {pair_s:.3f} pair-frames/s on {threads} threads", file=sys.stderr, flush=True)

    # frame ingest (the reference's frame_planes): a P010 frame of the source's size to packed planes
    if args.ingest:
        uv = es.DeviceArray((H // 2, W), np.uint16)
        es.synth_frame(es.SynthDesc(kind=1, seed=7, full_w=W, full_h=H // 2, w=W, h=H // 2, depth=10), 0, uv)
        y = src_frames[0] if sbytes == 2 else es.DeviceArray((H, W), np.uint16)
        if sbytes != 2:
            es.synth_frame(es.SynthDesc(kind=1, seed=7, full_w=W, full_h=H, w=W, h=H, depth=10), 0, y)
        fr = es.frame_view("p010", y, uv, None, W * 2, W * 2)
        oy, ou, ov = es.DeviceArray((H, W), np.uint16), es.DeviceArray((H // 2, W // 2), np.uint16), \
            es.DeviceArray((H // 2, W // 2), np.uint16)
        stream = es.Stream()
        for _ in range(3):
            es.frame_planes(fr, 0, 0, 10, oy, ou, ov, W, H, stream)
        start, stop = es.Event(), es.Event()
        start.record(stream)
        for _ in range(args.ingest):
            es.frame_planes(fr, 0, 0, 10, oy, ou, ov, W, H, stream)
        stop.record(stream)
        stream.synchronize()
        us = start.elapsed_ms(stop) * 1e3 / args.ingest
        nbytes = 1.5 * W * H * 4  # 4:2:0 16-bit samples in, packed 16-bit samples out
        result["ingest"] = {"kernel": "frame_planes_kernel<2, 1>", "frame": f"P010 {W}x{H}", "us_per_frame": round(us, 3),
                            "algorithmic_bytes": int(nbytes), "achieved_gbs": round(nbytes / us / 1e3, 1),
                            "frac": round(nbytes / us / 1e3 / HBM_PEAK_GBS, 4)}
        print(f"[bench] ingest: {us:.2f} us per frame", file=sys.stderr, flush=True)

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
