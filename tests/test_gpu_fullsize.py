"""Parity at the picture sizes of the benchmark configurations, against the oracle.

- Every c-value of every scale, bit for bit, of one 3840x2160 and one 3840x1600 picture (window 65 / pad 32 and window
  59 / pad 29: the matrix-core form's three-margin-block branch) in both c-value methods and both LDS layouts.
- About 14 frame steps of each configuration's own content through the batched scorer, every encode compared with
  the oracle's run of its (source crop, encode) pair: per-frame CAMBI bit-exact, the segments' totals, counts,
  quantiles and scores equal.

The frames are the benchmark's (bench.config), generated on the device.
"""
import importlib
import os

import numpy as np
import pytest

from test_gpu_scorer import compare_segment

pytestmark = pytest.mark.gpu
bench = importlib.import_module("bench")
THREADS = min(16, os.cpu_count() or 1)


def device_frames(es, desc, n, nbytes=2):
    frames = []
    for f in range(n):
        d = es.DeviceArray((desc.h, desc.w), np.uint8 if nbytes == 1 else np.uint16)
        es.synth_frame(desc, f, d, bytes=nbytes)
        frames.append(d)
    es.synchronize()
    return frames


@pytest.mark.parametrize("cfg", [2, 5])
@pytest.mark.parametrize("which", ["source", "encode"])
def test_every_cvalue_full_size(gpu, orc, cfg, which):
    src_d, encs, _, _, _ = bench.config(cfg, gpu, 16)
    desc = src_d if which == "source" else encs[0][0]
    w, h = desc.w, desc.h
    print(f"case C{cfg} {which} {w}x{h}", flush=True)
    dev = device_frames(gpu, desc, 1)[0]
    img = dev.host()
    p = orc.params(w, h)
    assert p.pad == {2: 32, 5: 29}[cfg]
    ref, ref_cv = orc.cambi(p, img, want_cvalues=True, threads=THREADS)
    for mfma in (True, False):
        for bmax in (32, 64):
            c = gpu.Cambi(w, h, 10, keep_cvalues=True, cv_mfma=mfma, cv_bmax=bmax)
            c.launch(dev, w)
            score = c.finish()
            for s in range(5):
                cv = c.cvalues(s)
                diff = np.count_nonzero(cv.view(np.uint32) != ref_cv[s].view(np.uint32))
                assert diff == 0, f"mfma={mfma} bmax={bmax} scale {s}: {diff} c-values differ"
            sc = c.scales()
            for s in range(5):
                assert (sc[s]["sum_fx"], sc[s]["count"], sc[s]["thr_bits"], sc[s]["above"], sc[s]["above_fx"]) == \
                       (ref.sum_fx[s], ref.count[s], ref.thr_bits[s], ref.above[s], ref.above_fx[s]), (mfma, bmax, s)
            assert score == ref.score, (mfma, bmax)
            del c
    # the picture exercises the select (more positive c-values than k at some scale) or pools all of them
    assert ref.score > 0


def scorer_vs_oracle(gpu, orc, cfg, n, encodes, per_segment=1440, settings=({},)):
    """Config `cfg`'s source and the chosen encodes for n frame steps (each encode its configuration length cut to
    n), through the batched scorer with each of `settings`, against the oracle's run of every pair."""
    src_d, encs, sbytes, shift, text = bench.config(cfg, gpu, n)
    encs = [encs[i] for i in encodes]
    W, H = src_d.w, src_d.h
    lengths = [min(length, n) for _, length in encs]
    print(f"case C{cfg} {W}x{H} encodes {[(d.w, d.h) for d, _ in encs]} lengths {lengths}", flush=True)
    src = device_frames(gpu, src_d, n, sbytes)
    enc = [device_frames(gpu, d, length) for (d, _), length in zip(encs, lengths)]
    src_host = [f.host().astype(np.uint16) << shift for f in src]
    refs = []
    for (d, _), length, frames in zip(encs, lengths, enc):
        cx, cy = ((W - d.w) // 2) & ~1, ((H - d.h) // 2) & ~1
        run = orc.Run(d.w, d.h, per_segment=per_segment, threads=THREADS)
        cam = [run.frame(s[cy:cy + d.h, cx:cx + d.w], e.host()) for s, e in zip(src_host[:length], frames)]
        refs.append((run, np.array(cam)))
    descs = [gpu.EncodeDesc(d.w, d.h, length, d.w, 2, 0) for (d, _), length in zip(encs, lengths)]
    ptrs = [[enc[e][f].ptr if f < lengths[e] else src[0].ptr for e in range(len(encs))] for f in range(n)]
    for kw in settings:
        kw = dict(dict(batch=8, use_graph=True), **kw)
        sc = gpu.Scorer(gpu.ES_BATCHED, W, H, descs, per_segment=per_segment, src_bytes=sbytes, src_shift=shift, **kw)
        sc.submit([f.ptr for f in src], ptrs)
        sc.sync()
        for e, (run, cam) in enumerate(refs):
            a, b = sc.cambi(e, lengths[e])
            assert np.array_equal(a, cam[:, 0]) and np.array_equal(b, cam[:, 1]), (kw, e)
            assert sc.segments(e) == run.segments()
            for s in range(run.segments()):
                got, hist = sc.segment(e, s, with_hist=True)
                ref = orc.accum_summary(run.segment(s))
                compare_segment(got, ref)
                assert np.array_equal(hist, ref[1]), (kw, e, s)
        del sc
    return refs


def test_c1_ramp_1080p(gpu, orc):
    """C1's own content: the gradient + radial ramp with grain against its 4-code-banded 3x3 box mean."""
    refs = scorer_vs_oracle(gpu, orc, 1, 14, [0], per_segment=8,
                            settings=({}, dict(cv_mfma=True, tile_h=64), dict(batch=16, cv_tile_blocks=True)))
    run, cam = refs[0]
    assert cam[:, 0].max() > 0  # the banded encode has CAMBI


def test_c2_mosaic_4k(gpu, orc):
    """C2 at 3840x2160: matrix cores at the autotuner's tile heights and prefix scans with the wide layout."""
    scorer_vs_oracle(gpu, orc, 2, 14, [0],
                     settings=(dict(cv_mfma=True, tile_h=64, tile_h0=128, cv_bmax=32, batch=8),
                               dict(cv_mfma=False, tile_h=32, cv_bmax=64, batch=8, cv_waves=12),
                               dict(cv_mfma=True, tile_h=128, cv_bmax=32, batch=8, cv_queues=1)))


def test_c3_uneven_crops(gpu, orc):
    """C3: the four real crops (1920x1080, 1920x800, 1920x1040, 1800x1080) and uneven lengths, every encode."""
    scorer_vs_oracle(gpu, orc, 3, 16, [0, 1, 2, 3], per_segment=9,
                     settings=(dict(cv_mfma=True, tile_h=64, batch=8), dict(cv_mfma=False, batch=16),
                               dict(cv_mfma=True, tile_h=96, batch=8, cv_queues=1, cv_waves=8)))


def test_c4_8bit_source_three_encodes(gpu, orc):
    """C4: an 8-bit source read as 1-byte samples shifted by 2 against three 10-bit encodes, graph replay."""
    scorer_vs_oracle(gpu, orc, 4, 14, [0, 1, 2], settings=(dict(cv_mfma=True, tile_h=64, tile_h0=96, batch=8),))


def test_c5_3840x1600_multi_encode(gpu, orc):
    """C5 at 3840x1600 (window 59): three of its eight encodes with their blur, quantisation and grain mix."""
    scorer_vs_oracle(gpu, orc, 5, 14, [0, 3, 7], settings=(dict(cv_mfma=True, tile_h=64, tile_h0=128, batch=8),))
