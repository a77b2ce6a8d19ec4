"""The C++ drop-in headers on the GPU: host/dropin_check.cpp (Kernels, Cambi, CambiCache) and host/dropin_main.cpp, a
caller shaped like src/main.cpp's queue/collect loop (by-name launches of frame_planes, window_stats, persist_stats,
plane_sse and color_sums with the reference's argument arrays; the reference's float figures; the collect's own
arithmetic), both against the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "encode-score_amd", "host")
THREADS = min(16, os.cpu_count() or 1)


def program(name):
    path = os.path.join(HOST, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built: run __graft_entry__.build() or make -C encode-score_amd")
    return path


def test_dropin_check_matches_the_oracle(gpu, orc, tmp_path):
    """Kernels::load, Cambi(k, w, h, depth).launch/finish and CambiCache add/find through the C++ headers."""
    source = tmp_path / "source.bin"
    source.write_bytes(b"source" * 500)
    r = subprocess.run([program("dropin_check"), "7", str(source)], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, XDG_CACHE_HOME=str(tmp_path / "cache")))
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    d = gpu.SynthDesc(kind=1, seed=7, full_w=256, full_h=144, w=256, h=144, depth=10, grain_q8=256, quant_step=4,
                      quant_where=1)
    ref = orc.cambi(orc.params(256, 144), gpu.synth_frame_host(d, 0)).score
    assert float(lines[0]) == ref
    cached, where = lines[1].split(" in ")
    assert float(cached.split()[1]) == ref and where.startswith(str(tmp_path / "cache" / "encode-score"))


def oracle_frames(es, W, H, w, h, n, seed):
    """The harness's frames (its es_synth_desc settings restated) and the oracle's run of them."""
    sx, sy = ((W - w) // 2) & ~1, ((H - h) // 2) & ~1
    sd = dict(kind=1, seed=seed, full_w=W, full_h=H, w=W, h=H, depth=10, grain_q8=256)
    ed = dict(kind=1, seed=seed, salt=1, full_w=W, full_h=H, crop_x=sx, crop_y=sy, w=w, h=h, depth=10, quant_step=2,
              quant_where=2, block=8, block_pct=50, blur_q8=128)

    def half(d, salt):
        d = dict(d, w=d["w"] // 2, h=d["h"] // 2, full_w=d["full_w"] // 2, full_h=d["full_h"] // 2,
                 crop_x=d.get("crop_x", 0) // 2, crop_y=d.get("crop_y", 0) // 2, salt=d.get("salt", 0) + salt)
        return es.SynthDesc(**d)

    planes = [es.SynthDesc(**sd), half(sd, 100), half(sd, 200), es.SynthDesc(**ed), half(ed, 100), half(ed, 200)]
    for f in range(n):
        y, u, v, ey, eu, ev = (es.synth_frame_host(d, f) for d in planes)
        yield ((y[sy:sy + h, sx:sx + w], u[sy // 2:(sy + h) // 2, sx // 2:(sx + w) // 2],
                v[sy // 2:(sy + h) // 2, sx // 2:(sx + w) // 2]), (ey, eu, ev))


def close(a, b, rel):
    if a is None or b is None or (isinstance(b, float) and np.isnan(b)):
        return (a is None or np.isnan(a)) and (b is None or np.isnan(b))
    return abs(a - b) <= rel * max(abs(a), abs(b), 1e-30)


@pytest.mark.parametrize("shape", [(320, 192, 288, 160, 20, 9), (1920, 1080, 1920, 1080, 14, 1440)])
def test_main_shaped_loop_matches_the_oracle(gpu, orc, shape):
    W, H, w, h, n, per_segment = shape
    seed = 0x5EED00AA
    print(f"case dropin_main {W}x{H} -> {w}x{h}, {n} frames", flush=True)
    r = subprocess.run([program("dropin_main"), *map(str, (W, H, w, h, n, per_segment, seed))], capture_output=True,
                       text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    run = orc.Run(w, h, per_segment=per_segment, threads=THREADS)
    cam, cam_src, psnr = [], [], []
    for s, e in oracle_frames(gpu, W, H, w, h, n, seed):
        a, b, p = run.frame_yuv(s, e)
        cam.append(a)
        cam_src.append(b)
        psnr.extend(p)
    # the CAMBI instances and the squared errors are exact: every per-frame figure equal
    assert got["cambi"] == cam and got["cambi_source"] == cam_src
    assert got["psnr"] == [float(x) for x in psnr]
    assert len(got["segments"]) == run.segments()
    for k, seg in enumerate(got["segments"]):
        ref, _ = orc.accum_summary(run.segment(k))
        assert seg["frames"] == ref["frames"]
        assert close(seg["cambi"], ref["cambi"], 1e-13) and close(seg["cambi_source"], ref["cambi_source"], 1e-13)
        for i in range(3):
            assert close(seg["psnr"][i], ref["psnr"][i], 1e-13)
            assert close(seg["color_source"][i], ref["color_source"][i], 1e-12)
            assert close(seg["color_encode"][i], ref["color_encode"][i], 1e-12)
        # the caller's totals add the reference's float figures (each the exact value rounded once): float rounding
        for key in ("all_r", "all_d", "st_r", "st_d", "sh_all_r", "sh_all_x", "sh_st_r", "sh_st_x"):
            for a, b in zip(seg[key], ref[key]):
                assert close(a, b, 1e-5), (k, key)
        # the structure masks are the same (the same float figures against the same float threshold)
        assert seg["n_amount"] == ref["n_amount"] and seg["n_matching"] == ref["n_matching"], k
        assert seg["n_lasting"] == ref["n_lasting"], k
        assert close(seg["own"], ref["own"], 1e-5) and close(seg["mine"], ref["mine"], 1e-5)
        # medians and tenths within one 0.05 bin (a percentage from float figures may land in the next bin)
        for key in ("median_amount", "tenth_amount", "median_matching", "tenth_matching"):
            for a, b in zip(seg[key], ref[key]):
                assert (a is None and np.isnan(b)) or abs(a - b) <= 0.05 + 1e-9, (k, key)
        for key in ("median_lasting", "tenth_lasting"):
            a, b = seg[key], ref[key]
            assert (a is None and np.isnan(b)) or abs(a - b) <= 0.05 + 1e-9, (k, key)
    assert sum(s["n_lasting"] for s in got["segments"]) > 0  # the run reaches the lasting texture
