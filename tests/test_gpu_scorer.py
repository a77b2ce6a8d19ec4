"""The whole measure path (plain per-frame form and batched form) against the oracle's run of src/main.cpp's
queue/collect: per-frame CAMBI bit-exact, segment totals and histograms exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H = 256, 160


def frames_of(es, n, w=W, h=H, seed=77, enc_kw=None, crop=(0, 0), full=(W, H), src_depth=10):
    sd = es.SynthDesc(kind=1, seed=seed, full_w=full[0], full_h=full[1], w=full[0], h=full[1], depth=src_depth,
                      grain_q8=256)
    kw = dict(grain_q8=0, quant_step=2, quant_where=2, block=8, block_pct=50)
    kw.update(enc_kw or {})
    ed = es.SynthDesc(kind=1, seed=seed, salt=1, full_w=full[0], full_h=full[1], crop_x=crop[0], crop_y=crop[1], w=w,
                      h=h, depth=10, **kw)
    src = [es.synth_frame_host(sd, f, bytes=1 if src_depth == 8 else 2) for f in range(n)]
    enc = [es.synth_frame_host(ed, f) for f in range(n)]
    return src, enc


def oracle_run(orc, src, enc, per_segment, win=16, shift=0, crop=None):
    h, w = enc[0].shape
    run = orc.Run(w, h, win=win, per_segment=per_segment, threads=8)
    cam, cam_src = [], []
    for s, e in zip(src, enc):
        s = s.astype(np.uint16) << shift
        if crop is not None:
            s = s[crop[1]:crop[1] + h, crop[0]:crop[0] + w]
        a, b = run.frame(s, e)
        cam.append(a)
        cam_src.append(b)
    return run, np.array(cam), np.array(cam_src)


def compare_segment(got, ref_summary, exact=True):
    ref, _ = ref_summary
    for k, v in ref.items():
        g = got[k]
        if isinstance(v, list):
            for a, b in zip(g, v):
                assert (np.isnan(a) and np.isnan(b)) or a == b, k
        else:
            assert (isinstance(v, float) and np.isnan(v) and np.isnan(g)) or g == v, k


@pytest.mark.parametrize("mode", ["plain", "batched", "graph"])
def test_scorer_matches_oracle(gpu, orc, mode):
    n, per_segment = 30, 11
    print(f"case scorer {mode}", flush=True)
    src, enc = frames_of(gpu, n)
    run, cam, cam_src = oracle_run(orc, src, enc, per_segment)
    ds = [gpu.DeviceArray.from_host(f) for f in src]
    de = [gpu.DeviceArray.from_host(f) for f in enc]
    sc = gpu.Scorer(gpu.ES_PLAIN if mode == "plain" else gpu.ES_BATCHED, W, H,
                    [gpu.EncodeDesc(W, H, n, W, 2, 0)], per_segment=per_segment, batch=8, use_graph=mode == "graph")
    sc.submit([d.ptr for d in ds[:13]], [[d.ptr] for d in de[:13]])
    sc.submit([d.ptr for d in ds[13:]], [[d.ptr] for d in de[13:]])
    sc.sync()
    a, b = sc.cambi(0, n)
    assert np.array_equal(a, cam) and np.array_equal(b, cam_src)
    assert sc.segments(0) == run.segments()
    for s in range(run.segments()):
        got, hist = sc.segment(0, s, with_hist=True)
        ref = orc.accum_summary(run.segment(s))
        compare_segment(got, ref)
        assert np.array_equal(hist, ref[1])
    got = sc.segment(0, -1)
    compare_segment(got, orc.accum_summary(run.merged()))


@pytest.mark.parametrize("batch", [64, 128])
def test_scorer_matches_oracle_at_the_largest_batch(gpu, orc, batch):
    """Rounds of 64 and of 128 frames (the largest batch; the autotuner picks 64 to 128 on the benchmark workloads): two
    full rounds and a partial one, submitted in pieces that straddle them, in graph replay with the matrix-core
    c-values."""
    n, per_segment = 2 * batch + 22, 40
    print(f"case scorer batch {batch}", flush=True)
    src, enc = frames_of(gpu, n, seed=64)
    run, cam, cam_src = oracle_run(orc, src, enc, per_segment)
    ds = [gpu.DeviceArray.from_host(f) for f in src]
    de = [gpu.DeviceArray.from_host(f) for f in enc]
    sc = gpu.Scorer(gpu.ES_BATCHED, W, H, [gpu.EncodeDesc(W, H, n, W, 2, 0)], per_segment=per_segment, batch=batch,
                    use_graph=True, cv_mfma=True, tile_h=64)
    for a0, a1 in ((0, batch + 6), (batch + 6, 2 * batch + 3), (2 * batch + 3, n)):
        sc.submit([d.ptr for d in ds[a0:a1]], [[d.ptr] for d in de[a0:a1]])
    sc.sync()
    a, b = sc.cambi(0, n)
    assert np.array_equal(a, cam) and np.array_equal(b, cam_src)
    assert sc.segments(0) == run.segments()
    for s in range(run.segments()):
        got, hist = sc.segment(0, s, with_hist=True)
        ref = orc.accum_summary(run.segment(s))
        compare_segment(got, ref)
        assert np.array_equal(hist, ref[1])
    compare_segment(sc.segment(0, -1), orc.accum_summary(run.merged()))
    with pytest.raises(gpu.EsError):  # the scorer takes at most 128 frames per batch
        gpu.Scorer(gpu.ES_BATCHED, W, H, [gpu.EncodeDesc(W, H, n, W, 2, 0)], batch=129)


def test_batched_uneven_lengths_and_crops(gpu, orc):
    """Encodes of different crops and lengths in one run equal each (source crop, encode) pair scored alone."""
    n = 26
    fw, fh = 320, 192
    specs = [((320, 192), 26, dict(blur_q8=154, grain_q8=256)), ((320, 128), 19, dict(quant_step=2, grain_q8=256)),
             ((288, 192), 14, dict(grain_q8=0)), ((320, 128), 22, dict(block=16, block_pct=50, grain_q8=256))]
    sd = gpu.SynthDesc(kind=1, seed=99, full_w=fw, full_h=fh, w=fw, h=fh, depth=10, grain_q8=256)
    src = [gpu.synth_frame_host(sd, f) for f in range(n)]
    encs, descs = [], []
    for i, ((w, h), length, kw) in enumerate(specs):
        cx, cy = ((fw - w) // 2) & ~1, ((fh - h) // 2) & ~1
        ed = gpu.SynthDesc(kind=1, seed=99, salt=i, full_w=fw, full_h=fh, crop_x=cx, crop_y=cy, w=w, h=h, depth=10, **kw)
        encs.append([gpu.synth_frame_host(ed, f) for f in range(length)])
        descs.append(gpu.EncodeDesc(w, h, length, w, 2, 0))
    ds = [gpu.DeviceArray.from_host(f) for f in src]
    de = [[gpu.DeviceArray.from_host(f) for f in e] for e in encs]
    for graph in (False, True):
        sc = gpu.Scorer(gpu.ES_BATCHED, fw, fh, descs, per_segment=9, batch=6, use_graph=graph)
        ptrs = [[de[e][f].ptr if f < len(de[e]) else ds[0].ptr for e in range(4)] for f in range(n)]
        sc.submit([d.ptr for d in ds], ptrs)
        sc.sync()
        for e, ((w, h), length, _) in enumerate(specs):
            cx, cy = ((fw - w) // 2) & ~1, ((fh - h) // 2) & ~1
            run, cam, cam_src = oracle_run(orc, src[:length], encs[e], 9, crop=(cx, cy))
            a, b = sc.cambi(e, length)
            assert np.array_equal(a, cam) and np.array_equal(b, cam_src), (graph, e)
            for s in range(run.segments()):
                got, hist = sc.segment(e, s, with_hist=True)
                ref = orc.accum_summary(run.segment(s))
                compare_segment(got, ref)
                assert np.array_equal(hist, ref[1])


def test_plain_and_batched_agree_on_8bit_source(gpu, orc):
    """An 8-bit source read as 1-byte samples shifted by 2 against a 10-bit encode (pipeline depth 10)."""
    n = 16
    src, enc = frames_of(gpu, n, src_depth=8, enc_kw=dict(grain_q8=128))
    run, cam, cam_src = oracle_run(orc, src, enc, 1440, shift=2)
    ds = [gpu.DeviceArray.from_host(f) for f in src]
    de = [gpu.DeviceArray.from_host(f) for f in enc]
    for mode in (gpu.ES_PLAIN, gpu.ES_BATCHED):
        sc = gpu.Scorer(mode, W, H, [gpu.EncodeDesc(W, H, n, W, 2, 0)], src_bytes=1, src_shift=2, batch=16)
        sc.submit([d.ptr for d in ds], [[d.ptr] for d in de])
        sc.sync()
        a, b = sc.cambi(0, n)
        assert np.array_equal(a, cam) and np.array_equal(b, cam_src)
        compare_segment(sc.segment(0, 0), orc.accum_summary(run.segment(0)))


def test_synth_device_equals_host(gpu):
    descs = [gpu.SynthDesc(kind=0, seed=0x5EED0001, full_w=480, full_h=270, w=480, h=270, depth=10, grain_q8=256),
             gpu.SynthDesc(kind=0, seed=0x5EED0001, full_w=480, full_h=270, w=480, h=270, depth=10, box3=1),
             gpu.SynthDesc(kind=1, seed=5, full_w=400, full_h=300, crop_x=40, crop_y=20, w=320, h=240, depth=10,
                           grain_q8=100, blur_q8=300, quant_step=4, quant_where=1, block=16, block_pct=30),
             gpu.SynthDesc(kind=1, seed=6, full_w=256, full_h=144, w=256, h=144, depth=8, grain_q8=256)]
    for d in descs:
        for frame in (0, 7):
            bytes_ = 1 if d.depth == 8 else 2
            host = gpu.synth_frame_host(d, frame, bytes=bytes_)
            dev = gpu.DeviceArray(host.shape, host.dtype)
            gpu.synth_frame(d, frame, dev, bytes=bytes_)
            assert np.array_equal(dev.host(), host)


def test_profiled_graph_rounds_match_eager(gpu):
    """Stage events recorded inside the captured graph time every round and leave the results unchanged; several
    replays with eager work and synchronisations between them give the eager results."""
    n = 40
    src, enc = frames_of(gpu, n, seed=5)
    ds = [gpu.DeviceArray.from_host(f) for f in src]
    de = [gpu.DeviceArray.from_host(f) for f in enc]
    out = {}
    for graph, profile in ((False, False), (True, True)):
        sc = gpu.Scorer(gpu.ES_BATCHED, W, H, [gpu.EncodeDesc(W, H, n, W, 2, 0)], batch=4, use_graph=graph,
                        profile=profile)
        for f0 in range(0, n, 10):
            sc.submit([d.ptr for d in ds[f0:f0 + 10]], [[d.ptr] for d in de[f0:f0 + 10]])
            gpu.synchronize()
            scratch = gpu.DeviceArray.from_host(np.arange(1000, dtype=np.int64))  # eager work between replays
            assert scratch.host()[999] == 999
        sc.sync()
        out[graph] = (sc.cambi(0, n), sc.segment(0, -1, with_hist=True))
        if profile:
            ms, rounds = sc.stage_ms()
            assert rounds == (n // 10) * 3 and all(v >= 0 for v in ms.values()) and ms["cambi_cvalues"] > 0
    (a0, b0), (s0, h0) = out[False]
    (a1, b1), (s1, h1) = out[True]
    assert np.array_equal(a0, a1) and np.array_equal(b0, b1) and np.array_equal(h0, h1)
    for k, v in s0.items():
        assert (v == s1[k]) or (isinstance(v, float) and np.isnan(v) and np.isnan(s1[k])), k


@pytest.mark.parametrize("win", [32, 64])
def test_batched_larger_windows_match_oracle(gpu, orc, win):
    """32x32 and 64x64 windows (the lasting texture then adds patch sums atomically into cleared outputs)."""
    n, w, h = 24, 320, 192
    src, enc = frames_of(gpu, n, w=w, h=h, seed=31, full=(w, h))
    run, cam, cam_src = oracle_run(orc, src, enc, 1440, win=win)
    ds = [gpu.DeviceArray.from_host(f) for f in src]
    de = [gpu.DeviceArray.from_host(f) for f in enc]
    sc = gpu.Scorer(gpu.ES_BATCHED, w, h, [gpu.EncodeDesc(w, h, n, w, 2, 0)], win=win, batch=8, use_graph=True)
    sc.submit([d.ptr for d in ds], [[d.ptr] for d in de])
    sc.sync()
    got, hist = sc.segment(0, 0, with_hist=True)
    ref = orc.accum_summary(run.segment(0))
    compare_segment(got, ref)
    assert np.array_equal(hist, ref[1])


@pytest.mark.parametrize("mfma", [False, True])
@pytest.mark.parametrize("bmax,tile_h,tile_h0,waves", [(64, 32, 0, 16), (64, 16, 48, 8), (32, 48, 0, 24),
                                                       (32, 64, 128, 12), (32, 32, 96, 0), (64, 24, 0, 0),
                                                       (32, 360, 0, 16), (32, 64, 192, 16)])
def test_cvalue_layouts_and_tiles_match_oracle(gpu, orc, bmax, tile_h, tile_h0, waves, mfma):
    """Every setting the autotuner may pick (c-value passes of up to 32 or 64 numbers, tile heights, a separate tile
    height at full resolution, waves per CU or one block per tile (waves 0)) gives the oracle's results; textured
    640x360 frames have tiles with more than 32 relevant values."""
    n, w, h = 14, 640, 360
    print(f"case layout bmax {bmax} tile_h {tile_h} tile_h0 {tile_h0} waves {waves} mfma {mfma}", flush=True)
    src, enc = frames_of(gpu, n, w=w, h=h, seed=41, full=(w, h))
    run, cam, cam_src = oracle_run(orc, src, enc, 1440)
    ds = [gpu.DeviceArray.from_host(f) for f in src]
    de = [gpu.DeviceArray.from_host(f) for f in enc]
    sc = gpu.Scorer(gpu.ES_BATCHED, w, h, [gpu.EncodeDesc(w, h, n, w, 2, 0)], batch=8, use_graph=True, tile_h=tile_h,
                    tile_h0=tile_h0, cv_bmax=bmax, cv_waves=waves, cv_tile_blocks=waves == 0, cv_mfma=mfma)
    sc.submit([d.ptr for d in ds], [[d.ptr] for d in de])
    sc.sync()
    a, b = sc.cambi(0, n)
    assert np.array_equal(a, cam) and np.array_equal(b, cam_src)
    got, hist = sc.segment(0, 0, with_hist=True)
    ref = orc.accum_summary(run.segment(0))
    compare_segment(got, ref)
    assert np.array_equal(hist, ref[1])


def test_tile_height_is_bounded(gpu):
    # a lane sums a tile's c-values exactly in FP64 for tiles of at most 4096 rows
    with pytest.raises(gpu.EsError, match="tile_h"):
        gpu.Scorer(gpu.ES_BATCHED, 64, 64, [gpu.EncodeDesc(64, 64, 4, 64, 2, 0)], batch=4, tile_h=4097)
    with pytest.raises(gpu.EsError, match="tile_h"):
        gpu.Scorer(gpu.ES_BATCHED, 64, 64, [gpu.EncodeDesc(64, 64, 4, 64, 2, 0)], batch=4, tile_h0=4097)


def test_autotune_picks_valid_settings_and_caches(gpu, tmp_path):
    n = 96  # three batches of 32: batch 64 needs 192 frames and is not tried
    src, enc = frames_of(gpu, n, seed=12)
    ds = [gpu.DeviceArray.from_host(f) for f in src]
    de = [gpu.DeviceArray.from_host(f) for f in enc]
    descs = [gpu.EncodeDesc(W, H, n, W, 2, 0)]
    cache = str(tmp_path / "tune.json")
    r = gpu.autotune(W, H, descs, [d.ptr for d in ds], [[d.ptr] for d in de], cache=cache)
    assert r["batch"] in (8, 16, 32) and r["tile_h"] in (16, 24, 32, 48, 64, 96, 128)
    assert r["tile_h0"] in (0, 64, 96, 128, 192)
    assert r["cv_bmax"] in (32, 64) and r["cv_waves"] in (0, 6, 8, 12, 16, 24)
    assert r["cv_tile_blocks"] == (r["cv_waves"] == 0) and r["early_energies"] in (False, True)
    assert r["ms_per_step"] > 0 and r["trials"] >= 14 and not r["cached"]
    again = gpu.autotune(W, H, descs, [d.ptr for d in ds], [[d.ptr] for d in de], cache=cache)
    assert again["cached"] and {k: again[k] for k in ("batch", "tile_h", "tile_h0", "cv_bmax", "cv_waves")} == \
        {k: r[k] for k in ("batch", "tile_h", "tile_h0", "cv_bmax", "cv_waves")}
    with pytest.raises(gpu.EsError):  # too few frames for three batches
        gpu.autotune(W, H, descs, [d.ptr for d in ds[:10]], [[d.ptr] for d in de[:10]])


def test_overlapped_and_serial_stages_agree(gpu):
    """The CAMBI and energy chains overlapped on two streams (the default) give the figures of the serial order, in
    eager, graph and profiled-graph forms."""
    n = 24
    src, enc = frames_of(gpu, n, seed=8)
    ds = [gpu.DeviceArray.from_host(f) for f in src]
    de = [gpu.DeviceArray.from_host(f) for f in enc]
    out = []
    for serial, graph, profile, early in ((True, True, False, False), (False, False, False, False),
                                          (False, True, False, False), (False, True, True, False),
                                          (False, True, False, True), (False, False, True, True)):
        sc = gpu.Scorer(gpu.ES_BATCHED, W, H, [gpu.EncodeDesc(W, H, n, W, 2, 0)], batch=5, use_graph=graph,
                        profile=profile, serial_stages=serial, per_segment=10, early_energies=early)
        sc.submit([d.ptr for d in ds], [[d.ptr] for d in de])
        sc.sync()
        out.append((sc.cambi(0, n), [sc.segment(0, s, with_hist=True) for s in range(sc.segments(0))]))
    (a0, b0), segs0 = out[0]
    for (a, b), segs in out[1:]:
        assert np.array_equal(a, a0) and np.array_equal(b, b0)
        for (s0, h0), (s1, h1) in zip(segs0, segs):
            assert np.array_equal(h0, h1)
            for k, v in s0.items():
                assert (v == s1[k]) or (isinstance(v, float) and np.isnan(v) and np.isnan(s1[k])), k
