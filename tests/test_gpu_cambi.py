"""CAMBI on the GPU against the CPU oracle: packed scale planes, c-values, the top-k select and the score, bit-exact."""
import numpy as np
import pytest

from conftest import banded_mosaic, staircase

pytestmark = pytest.mark.gpu


def run_gpu(gpu, img, depth=10, keep=True, bytes=2, shift=0):
    h, w = img.shape
    c = gpu.Cambi(w, h, depth, keep_cvalues=keep)
    d = gpu.DeviceArray.from_host(img)
    c.launch(d, w, bytes=bytes, shift=shift)
    score = c.finish()
    return c, score


def pictures(es):
    rng = np.random.default_rng(7)
    out = {
        "staircase": staircase(72, 200),
        "staircase_vertical": staircase(150, 90, step_px=4, base=150, vertical=True),
        "flat": np.full((64, 96), 500, np.uint16),
        "noise": rng.integers(64, 941, (80, 120)).astype(np.uint16),
        "mosaic_256x144": banded_mosaic(es, 256, 144, 11, quant_step=4, quant_where=1),
        "mosaic_odd_333x211": banded_mosaic(es, 333, 211, 12, quant_step=2, block=8, block_pct=50),
        "two_level": np.where(np.arange(96)[None, :] < 48, 400, 401).repeat(64, 0).astype(np.uint16),
    }
    return out


@pytest.mark.parametrize("name", ["staircase", "staircase_vertical", "flat", "noise", "mosaic_256x144",
                                  "mosaic_odd_333x211", "two_level"])
def test_cambi_bit_exact(gpu, orc, name):
    img = pictures(gpu)[name]
    h, w = img.shape
    p = orc.params(w, h)
    print(f"case {name} {w}x{h}", flush=True)
    ref, ref_cv = orc.cambi(p, img, want_cvalues=True)
    ims, mks = orc.pyramid(p, img)
    c, score = run_gpu(gpu, img)
    for s in range(5):
        plane = c.plane(s)
        assert np.array_equal(plane & 0x7FFF, ims[s]), f"scale {s} image"
        assert np.array_equal(plane >> 15, mks[s]), f"scale {s} mask"
        cv = c.cvalues(s)
        assert np.array_equal(cv.view(np.uint32), ref_cv[s].view(np.uint32)), f"scale {s} c-values"
    sc = c.scales()
    for s in range(5):
        assert sc[s]["sum_fx"] == ref.sum_fx[s]
        assert sc[s]["count"] == ref.count[s]
        assert sc[s]["k"] == ref.k[s]
        assert sc[s]["thr_bits"] == ref.thr_bits[s]
        assert sc[s]["above"] == ref.above[s]
        assert sc[s]["above_fx"] == ref.above_fx[s]
        assert sc[s]["pooled"] == ref.pooled[s]
    assert score == ref.score


def test_topk_select_is_exercised(gpu, orc):
    """A staircase makes more than 60% of the pixels positive, so the radix select decides the pooled value."""
    img = staircase(72, 200)
    p = orc.params(200, 72)
    ref = orc.cambi(p, img)
    assert any(ref.count[s] > ref.k[s] for s in range(5))
    c, score = run_gpu(gpu, img)
    assert score == ref.score and score > 0


def test_cambi_depth8_from8(gpu, orc):
    """cambi_from8: 8-bit pictures are promoted and anti-dithered (src/kernels.cu:522-535)."""
    rng = np.random.default_rng(3)
    img8 = (100 + np.arange(160)[None, :] // 10 + rng.integers(0, 2, (90, 160))).astype(np.uint16)
    p = orc.params(160, 90, depth=8)
    ref = orc.cambi(p, img8)
    c, score = run_gpu(gpu, img8, depth=8)
    ims, _ = orc.pyramid(p, img8)
    assert np.array_equal(c.plane(0) & 0x7FFF, ims[0])
    assert score == ref.score


def test_cambi_u8_samples_with_shift(gpu, orc):
    """An 8-bit file in a 10-bit pipeline read straight from 1-byte samples, shifted left by 2."""
    img8 = banded_mosaic(gpu, 192, 128, 21, quant_step=2).astype(np.int32)
    img8 = np.clip((img8 - 64) * 219 // 876 + 16, 16, 235).astype(np.uint8)
    p = orc.params(192, 128)
    ref = orc.cambi(p, img8.astype(np.uint16) << 2)
    h, w = img8.shape
    c = gpu.Cambi(w, h, 10)
    d = gpu.DeviceArray.from_host(img8)
    c.launch(d, w, bytes=1, shift=2)
    assert c.finish() == ref.score


def test_cambi_pitch_and_stream(gpu, orc):
    """A plane with a row pitch wider than the picture, launched on a caller stream (src/cambi.hpp:15-18)."""
    img = banded_mosaic(gpu, 256, 144, 5, quant_step=4)
    wide = np.zeros((144, 300), np.uint16)
    wide[:, :256] = img
    p = orc.params(256, 144)
    ref = orc.cambi(p, img)
    s = gpu.Stream()
    c = gpu.Cambi(256, 144)
    d = gpu.DeviceArray.from_host(wide)
    c.launch(d, 300, stream=s)
    assert c.finish() == ref.score


def test_cambi_1080p_frame(gpu, orc):
    """A full-size 1920x1080 picture of the tile mosaic with banding."""
    img = banded_mosaic(gpu, 1920, 1080, 0x5EED0002, quant_step=2, quant_where=2, block=8, block_pct=50)
    p = orc.params(1920, 1080)
    ref = orc.cambi(p, img, threads=8)
    c, score = run_gpu(gpu, img, keep=False)
    sc = c.scales()
    for s in range(5):
        assert sc[s]["sum_fx"] == ref.sum_fx[s] and sc[s]["count"] == ref.count[s]
    assert score == ref.score


def test_repeated_launches_reuse_the_instance(gpu, orc):
    c = gpu.Cambi(256, 144)
    for seed in (1, 2, 3):
        img = banded_mosaic(gpu, 256, 144, seed, quant_step=4, quant_where=1)
        d = gpu.DeviceArray.from_host(img)
        c.launch(d, 256)
        assert c.finish() == orc.cambi(orc.params(256, 144), img).score
