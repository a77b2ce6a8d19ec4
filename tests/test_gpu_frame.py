"""Frame ingest, PSNR and colour on the GPU against the oracle: es_frame_planes bit-exact for every format, crop and
depth; es_plane_sse and es_color_sums exact (every 10-bit chroma pair); the scorer reading decoded frames in place
with chroma, both forms, against the oracle's run; the JSON report in the reference's format."""
import json
import re

import numpy as np
import pytest

from framefmt import FORMATS, pack, values

pytestmark = pytest.mark.gpu


def upload(gpu, fmt, planes):
    nbytes = FORMATS[fmt][1]
    d = [None if p is None else gpu.DeviceArray.from_host(p) for p in planes]
    f = gpu.frame_view(fmt, d[0], d[1], d[2], planes[0].shape[1] * nbytes, planes[1].shape[1] * nbytes)
    return f, d


@pytest.mark.parametrize("fmt", list(FORMATS))
@pytest.mark.parametrize("crop", [(0, 0, 330, 184), (6, 4, 256, 128), (3, 5, 200, 96), (1, 0, 326, 180)])
def test_frame_planes_match_oracle(gpu, orc, fmt, crop):
    print(f"case frame_planes {fmt} {crop}", flush=True)
    rng = np.random.default_rng(sum(map(ord, fmt)) * 100 + crop[0] * 7 + crop[1])
    W, H = 330, 184
    x0, y0, w, h = crop
    assert x0 + w <= W and y0 + h <= H
    y, u, v = values(rng, fmt, W, H)
    planes = pack(fmt, y, u, v, pad=7, junk=rng)
    f, keep = upload(gpu, fmt, planes)
    for pd in ((8, 10) if FORMATS[fmt][3] == 8 else (10,)):
        oy = gpu.DeviceArray((h, w), np.uint16)
        ou = gpu.DeviceArray((h // 2, w // 2), np.uint16)
        ov = gpu.DeviceArray((h // 2, w // 2), np.uint16)
        gpu.frame_planes(f, x0, y0, pd, oy, ou, ov, w, h)
        gpu.synchronize()
        want = orc.frame_planes(fmt, *planes, x0, y0, w, h, pd)
        for g, e in zip((oy.host(), ou.host(), ov.host()), want):
            assert np.array_equal(g, e), (fmt, crop, pd)


def test_plane_sse_matches_oracle(gpu, orc):
    rng = np.random.default_rng(21)
    n = 1000003
    a = rng.integers(0, 65536, n).astype(np.uint16)
    b = rng.integers(0, 65536, n).astype(np.uint16)
    da, db = gpu.DeviceArray.from_host(a), gpu.DeviceArray.from_host(b)
    sse = gpu.DeviceArray.from_host(np.zeros(1, np.uint64))
    gpu.plane_sse(da, db, n, sse)
    gpu.synchronize()
    assert int(sse.host()[0]) == orc.plane_sse(a, b)
    gpu.plane_sse(da, db, n, sse)  # adds, as the reference's atomics do
    gpu.synchronize()
    assert int(sse.host()[0]) == 2 * orc.plane_sse(a, b)
    # not 16-byte aligned: the one-sample path
    s2 = gpu.DeviceArray.from_host(np.zeros(1, np.uint64))
    gpu.plane_sse(da.ptr + 2, db.ptr + 2, n - 5, s2)
    gpu.synchronize()
    assert int(s2.host()[0]) == orc.plane_sse(a[1:n - 4], b[1:n - 4])
    fa = rng.uniform(-500, 500, 77777).astype(np.float32)
    fb = rng.uniform(-500, 500, 77777).astype(np.float32)
    s3 = gpu.DeviceArray.from_host(np.zeros(1, np.uint64))
    gpu.plane_sse(gpu.DeviceArray.from_host(fa), gpu.DeviceArray.from_host(fb), fa.size, s3, floats=True)
    gpu.synchronize()
    assert int(s3.host()[0]) == orc.plane_sse(fa, fb)


def test_color_sums_exact_for_every_10bit_pair(gpu, orc):
    """All 2^20 (u, v) pairs at the 10-bit scale, per row of 1024 and in one call: each strength rounded to single
    precision as IEEE sqrtf does (the oracle's C library square root)."""
    u = np.repeat(np.arange(1024, dtype=np.uint16), 1024)
    v = np.tile(np.arange(1024, dtype=np.uint16), 1024)
    du, dv = gpu.DeviceArray.from_host(u), gpu.DeviceArray.from_host(v)
    sums = gpu.DeviceArray.from_host(np.zeros((1024, 3), np.uint64))
    for r in range(1024):
        gpu.color_sums(du.ptr + r * 2048, dv.ptr + r * 2048, 1024, 0, sums.ptr + r * 24)
    total = gpu.DeviceArray.from_host(np.zeros(3, np.uint64))
    gpu.color_sums(du, dv, u.size, 0, total)
    gpu.synchronize()
    got = sums.host()
    for r in range(1024):
        assert np.array_equal(got[r], orc.color_sums(u[r * 1024:(r + 1) * 1024], v[r * 1024:(r + 1) * 1024], 0)), r
    assert np.array_equal(total.host(), orc.color_sums(u, v, 0))
    # 8-bit samples shifted to the 10-bit scale, from an odd offset (the one-sample path)
    u8 = (np.arange(10001) % 256).astype(np.uint16)
    v8 = ((np.arange(10001) * 7) % 256).astype(np.uint16)
    d8u, d8v = gpu.DeviceArray.from_host(u8), gpu.DeviceArray.from_host(v8)
    s8 = gpu.DeviceArray.from_host(np.zeros(3, np.uint64))
    gpu.color_sums(d8u.ptr + 2, d8v.ptr + 2, 9999, 2, s8)
    gpu.synchronize()
    assert np.array_equal(s8.host(), orc.color_sums(u8[1:10000], v8[1:10000], 2))


# ---------------------------------------------------------------------------------------------------------------
# the scorer on decoded frames with chroma
# ---------------------------------------------------------------------------------------------------------------
W, H = 256, 160


def yuv_run(gpu, n, w, h, seed, src_fmt="yuv420p10", enc_fmt="yuv420p10"):
    """Sample values of n frame steps: the source (W x H, src_fmt's depth) and one encode cropped to w x h (centred,
    even), luma from the synthetic mosaic, chroma smooth with noise; the encode's chroma a coarser copy."""
    cx, cy = ((W - w) // 2) & ~1, ((H - h) // 2) & ~1
    sdepth, edepth = FORMATS[src_fmt][3], FORMATS[enc_fmt][3]
    sd = gpu.SynthDesc(kind=1, seed=seed, full_w=W, full_h=H, w=W, h=H, depth=sdepth, grain_q8=256)
    ed = gpu.SynthDesc(kind=1, seed=seed, salt=1, full_w=W, full_h=H, crop_x=cx, crop_y=cy, w=w, h=h, depth=edepth,
                       grain_q8=0, quant_step=2, quant_where=2, block=8, block_pct=50)
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H // 2, 0:W // 2]
    src, enc = [], []
    for f in range(n):
        top_s, top_e = (1 << sdepth) - 1, (1 << edepth) - 1
        base_u = (0.5 + 0.3 * np.sin(xx / 17.0 + f / 5.0)) * top_s
        base_v = (0.5 + 0.3 * np.cos(yy / 11.0 - f / 7.0)) * top_s
        su = np.clip(base_u + rng.normal(0, 3, base_u.shape), 0, top_s).astype(np.int64)
        sv = np.clip(base_v + rng.normal(0, 3, base_v.shape), 0, top_s).astype(np.int64)
        c = (slice(cy // 2, cy // 2 + h // 2), slice(cx // 2, cx // 2 + w // 2))
        scale = top_e / top_s
        eu = np.clip((su[c] * scale).astype(np.int64) // 4 * 4 + rng.integers(0, 2, (h // 2, w // 2)), 0, top_e)
        ev = np.clip((sv[c] * scale).astype(np.int64) // 4 * 4, 0, top_e)
        src.append((gpu.synth_frame_host(sd, f, bytes=1 if sdepth == 8 else 2).astype(np.int64), su, sv))
        enc.append((gpu.synth_frame_host(ed, f, bytes=1 if edepth == 8 else 2).astype(np.int64), eu, ev))
    return src, enc, (cx, cy)


def oracle_yuv(orc, src, enc, crop, w, h, per_segment, src_up=0, depth=10):
    run = orc.Run(w, h, depth=depth, up=10 - depth, per_segment=per_segment, threads=8)
    cx, cy = crop
    cam, cam_src, ps = [], [], []
    for s, e in zip(src, enc):
        sy = (s[0][cy:cy + h, cx:cx + w] << src_up).astype(np.uint16)
        su = (s[1][cy // 2:cy // 2 + h // 2, cx // 2:cx // 2 + w // 2] << src_up).astype(np.uint16)
        sv = (s[2][cy // 2:cy // 2 + h // 2, cx // 2:cx // 2 + w // 2] << src_up).astype(np.uint16)
        a, b, p = run.frame_yuv((sy, su, sv), tuple(x.astype(np.uint16) for x in e))
        cam.append(a)
        cam_src.append(b)
        ps.append(p)
    return run, np.array(cam), np.array(cam_src), np.array(ps)


def compare(got, ref_summary):
    ref, _ = ref_summary
    for k, v in ref.items():
        g = got[k]
        if isinstance(v, list):
            for a, b in zip(g, v):
                assert (np.isnan(a) and np.isnan(b)) or a == b, k
        else:
            assert (isinstance(v, float) and np.isnan(v) and np.isnan(g)) or g == v, k


def submit_formats(gpu, sc, src, enc, src_fmt, enc_fmt, junk=None, chunks=(9,), pts=None):
    keep, fs, fe = [], [], []
    for s, e in zip(src, enc):
        f1, d1 = upload(gpu, src_fmt, pack(src_fmt, *s, pad=3, junk=junk))
        f2, d2 = upload(gpu, enc_fmt, pack(enc_fmt, *e, pad=1, junk=junk))
        keep += d1 + d2
        fs.append(f1)
        fe.append([f2])
    i = 0
    for c in chunks + (len(fs),):
        sc.submit_frames(fs[i:i + c], fe[i:i + c], pts=None if pts is None else pts[i:i + c])
        i += c
        if i >= len(fs):
            break
    sc.sync()
    return keep


@pytest.mark.parametrize("mode", ["plain", "batched", "graph"])
@pytest.mark.parametrize("size", [(224, 128), (226, 134)])
def test_scorer_frames_with_chroma_match_oracle(gpu, orc, mode, size):
    """226 x 134: columns and rows outside whole 16 x 16 windows, whose luma squared error the batched form's frame sums
    add beside the window-stats kernel's share."""
    n, per_segment = 17, 7
    w, h = size
    print(f"case scorer frames {mode} {w}x{h}", flush=True)
    src, enc, crop = yuv_run(gpu, n, w, h, seed=41)
    run, cam, cam_src, ps = oracle_yuv(orc, src, enc, crop, w, h, per_segment)
    sc = gpu.Scorer(gpu.ES_PLAIN if mode == "plain" else gpu.ES_BATCHED, W, H, [gpu.EncodeDesc(w, h, n, w, 2, 0)],
                    per_segment=per_segment, batch=6, use_graph=mode == "graph")
    keep = submit_formats(gpu, sc, src, enc, "yuv420p10", "yuv420p10")
    a, b = sc.cambi(0, n)
    assert np.array_equal(a, cam) and np.array_equal(b, cam_src)
    assert np.array_equal(sc.psnr(0, n), ps)
    for s in range(run.segments()):
        got, hist = sc.segment(0, s, with_hist=True)
        assert got["chroma_frames"] == got["frames"]
        ref = orc.accum_summary(run.segment(s))
        compare(got, ref)
        assert np.array_equal(hist, ref[1])
    compare(sc.segment(0, -1), orc.accum_summary(run.merged()))
    del keep


def test_scorer_reads_p010_and_nv12_in_place(gpu, orc):
    """A P010 source (with nonzero unused low bits) and an NV12 8-bit source in a 10-bit pipeline, read in place, give
    the figures of their planar samples."""
    n, w, h = 14, 224, 128
    for src_fmt, enc_fmt, up in (("p010", "p010", 0), ("nv12", "yuv420p10", 2)):
        print(f"case scorer in place {src_fmt} {enc_fmt}", flush=True)
        src, enc, crop = yuv_run(gpu, n, w, h, seed=43, src_fmt=src_fmt, enc_fmt=enc_fmt)
        run, cam, cam_src, ps = oracle_yuv(orc, src, enc, crop, w, h, 1440, src_up=up)
        for mode in (gpu.ES_PLAIN, gpu.ES_BATCHED):
            sc = gpu.Scorer(mode, W, H, [gpu.EncodeDesc(w, h, n, w, 2, 0)], batch=8)
            keep = submit_formats(gpu, sc, src, enc, src_fmt, enc_fmt, junk=np.random.default_rng(1), chunks=(5,))
            a, b = sc.cambi(0, n)
            assert np.array_equal(a, cam) and np.array_equal(b, cam_src), (src_fmt, mode)
            assert np.array_equal(sc.psnr(0, n), ps)
            compare(sc.segment(0, 0), orc.accum_summary(run.segment(0)))
            del keep


def test_scorer_8bit_pipeline_matches_oracle(gpu, orc):
    """Every file 8-bit: pipeline depth 8 (src/main.cpp:382), CAMBI's 8-bit promotion, the energies and colour at the
    10-bit scale (up = 2), PSNR with peak 255 and cap 60 dB."""
    n, w, h = 14, 224, 128
    src, enc, crop = yuv_run(gpu, n, w, h, seed=47, src_fmt="yuv420p", enc_fmt="nv12")
    run, cam, cam_src, ps = oracle_yuv(orc, src, enc, crop, w, h, 6, depth=8)
    for mode in (gpu.ES_PLAIN, gpu.ES_BATCHED):
        print(f"case scorer 8-bit pipeline {mode}", flush=True)
        sc = gpu.Scorer(mode, W, H, [gpu.EncodeDesc(w, h, n, w, 1, 0)], depth=8, per_segment=6, batch=4,
                        use_graph=mode == gpu.ES_BATCHED)
        keep = submit_formats(gpu, sc, src, enc, "yuv420p", "nv12", chunks=(6,))
        a, b = sc.cambi(0, n)
        assert np.array_equal(a, cam) and np.array_equal(b, cam_src), mode
        assert np.array_equal(sc.psnr(0, n), ps)
        for s in range(run.segments()):
            compare(sc.segment(0, s), orc.accum_summary(run.segment(s)))
        del keep


def test_frames_without_chroma_and_mixing_is_an_error(gpu):
    n, w, h = 4, 224, 128
    src, enc, crop = yuv_run(gpu, n, w, h, seed=44)
    sc = gpu.Scorer(gpu.ES_BATCHED, W, H, [gpu.EncodeDesc(w, h, n, w, 2, 0)], batch=4)
    ds = gpu.DeviceArray.from_host(src[0][0].astype(np.uint16))
    de = gpu.DeviceArray.from_host(enc[0][0].astype(np.uint16))
    luma = gpu.frame_view("yuv420p10", ds, None, None, W * 2, W)
    luma_e = gpu.frame_view("yuv420p10", de, None, None, w * 2, w)
    sc.submit_frames([luma], [[luma_e]])
    f, keep = upload(gpu, "yuv420p10", pack("yuv420p10", *src[1]))
    g, keep2 = upload(gpu, "yuv420p10", pack("yuv420p10", *enc[1]))
    with pytest.raises(gpu.EsError, match="chroma"):
        sc.submit_frames([f], [[g]])
    sc.sync()
    assert np.isnan(sc.psnr(0, 1)).all()
    seg = sc.segment(0, 0)
    assert seg["chroma_frames"] == 0 and seg["frames"] == 1


PARTS_KEYS = ["frames", "score", "fidelity", "texture", "banding", "grain", "vmaf", "vmaf_min", "psnr_y", "psnr_cb",
              "psnr_cr", "cambi", "cambi_source", "lasting", "amount", "matching", "color_source", "color_encode"]


def fmt_eq(text_value, value, digits):
    """A JSON number as the reference prints it with %.{digits}f, or null for a non-finite figure."""
    if value is None or not np.isfinite(value):
        return text_value is None
    return text_value == float(f"{value:.{digits}f}")


def test_json_report_in_the_reference_format(gpu, tmp_path):
    n, per_segment, w, h = 20, 7, 224, 128
    src, enc, crop = yuv_run(gpu, n, w, h, seed=45)
    sc = gpu.Scorer(gpu.ES_BATCHED, W, H, [gpu.EncodeDesc(w, h, n, w, 2, 0), gpu.EncodeDesc(w, h, n - 3, w, 2, 0)],
                    per_segment=per_segment, batch=8)
    keep, fs, fe = [], [], []
    for i, (s, e) in enumerate(zip(src, enc)):
        f1, d1 = upload(gpu, "yuv420p10", pack("yuv420p10", *s))
        f2, d2 = upload(gpu, "yuv420p10", pack("yuv420p10", *e))
        keep += d1 + d2
        fs.append(f1)
        fe.append([f2, f2])
    sc.submit_frames(fs, fe)
    sc.sync()
    vmaf = np.linspace(80, 95, n)
    sc.set_vmaf(0, vmaf)
    info = dict(source="src.mkv", paths=["a.mkv", "b \"q\".mkv"], model="vmaf_v0.6.1", start=12.5, offset=-0.25,
                fps=24.0, segment_seconds=per_segment / 24.0, width=w, height=h)
    path = tmp_path / "run.json"
    sc.write_json(info, path)
    text = path.read_text()
    doc = json.loads(text)
    assert list(doc) == ["source", "start", "offset", "frames", "fps", "width", "height", "model", "segment_seconds",
                         "encodes"]
    assert doc["frames"] == n and doc["source"] == "src.mkv" and doc["encodes"][1]["path"] == "b \"q\".mkv"
    assert re.search(r'"start": 12\.500000, "offset": -0\.250000, "frames": 20, "fps": 24\.000000', text)
    for e, enc_doc in enumerate(doc["encodes"]):
        assert list(enc_doc) == ["path", "whole", "segments", "per_frame"]
        length = n if e == 0 else n - 3
        assert len(enc_doc["segments"]) == (length + per_segment - 1) // per_segment
        for k, part in enumerate([enc_doc["whole"]] + enc_doc["segments"]):
            assert list(part) == PARTS_KEYS
            assert list(part["lasting"]) == ["windows", "median", "worst_tenth", "total"]
            assert list(part["amount"]) == ["all", "structure", "median", "worst_tenth"]
            assert list(part["matching"]) == ["all", "structure", "median", "worst_tenth"]
            seg = sc.segment(e, k - 1)
            f = seg["frames"]
            assert part["frames"] == f
            for key, digits in (("score", 4), ("texture", 4), ("banding", 4), ("grain", 4), ("fidelity", 4)):
                assert fmt_eq(part[key], seg[key], digits), (e, k, key)
            assert fmt_eq(part["cambi"], seg["cambi"] / f, 6) and fmt_eq(part["cambi_source"], seg["cambi_source"] / f, 6)
            for i, name in enumerate(("psnr_y", "psnr_cb", "psnr_cr")):
                assert fmt_eq(part[name], seg["psnr"][i] / f, 4)
            for i in range(3):
                assert fmt_eq(part["color_source"][i], seg["color_source"][i] / f, 4)
                assert fmt_eq(part["color_encode"][i], seg["color_encode"][i] / f, 4)
            assert part["lasting"]["windows"] == seg["n_lasting"]
            if e == 0:
                assert fmt_eq(part["vmaf"], seg["vmaf"] / f, 6) and fmt_eq(part["vmaf_min"], seg["vmaf_min"], 6)
                assert part["fidelity"] == float(f"{min(max(seg['vmaf'] / f, 0), 100):.4f}")
            else:
                assert part["vmaf"] is None and part["vmaf_min"] is None and part["fidelity"] is None
        pf = enc_doc["per_frame"]
        assert list(pf) == ["vmaf", "cambi", "cambi_source", "psnr_y"]
        cam, cam_src = sc.cambi(e, length)
        assert pf["cambi"] == [float(f"{x:.6f}") for x in cam]
        assert pf["cambi_source"] == [float(f"{x:.6f}") for x in cam_src]
        assert pf["psnr_y"] == [float(f"{x:.3f}") for x in sc.psnr(e, length)[:, 0]]
        assert pf["vmaf"] == ([float(f"{x:.6f}") for x in vmaf] if e == 0 else [None] * length)
    # every number carries the reference's digits
    assert re.search(r'"score": -?\d+\.\d{4}, "fidelity": ', text)
    assert re.search(r'"cambi": \d+\.\d{6}, "cambi_source": \d+\.\d{6}', text)
    del keep


@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_source_cambi_cache_skips_cached_frames(gpu, orc, tmp_path, mode):
    """A second run against the same source reads every source CAMBI from the cache (src/main.cpp:520-521, :585-587)
    and skips computing it, with the same figures; the first run's file holds its scores as 16-byte records."""
    n, w, h = 12, 224, 128
    print(f"case cache {mode}", flush=True)
    src, enc, crop = yuv_run(gpu, n, w, h, seed=46)
    source = tmp_path / "source.mkv"
    source.write_bytes(b"a source file" * 1000)
    pts = np.arange(n, dtype=np.int64) * 1001 + 7
    m = gpu.ES_PLAIN if mode == "plain" else gpu.ES_BATCHED
    results = []
    for k in range(3):
        sc = gpu.Scorer(m, W, H, [gpu.EncodeDesc(w, h, n, w, 2, 0)], batch=5)
        if k:
            sc.use_cache(source, tmp_path / "cache")
        keep = submit_formats(gpu, sc, src, enc, "yuv420p10", "yuv420p10", chunks=(7,), pts=pts if k else None)
        results.append((sc.cambi(0, n), sc.segment(0, -1), sc.psnr(0, n), sc.cache_hits(), sc.cache_file(0)))
        del sc, keep
    (a0, b0), seg0, p0, _, _ = results[0]
    for k in (1, 2):
        (a, b), seg, p, hits, file = results[k]
        assert hits == (0 if k == 1 else n), k
        assert np.array_equal(a, a0) and np.array_equal(b, b0) and np.array_equal(p, p0)
        compare(seg, ({key: seg0[key] for key in seg0}, None))
    assert results[1][4] == results[2][4] != ""
    assert orc.cambi_cache_records(results[1][4]) == [(int(t), float(v)) for t, v in zip(pts, b0)]
