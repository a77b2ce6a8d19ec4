"""The oracle's frame ingest, PSNR and colour (src/kernels.cu:282-310, :453-484; src/main.cpp:273-277) pinned
against numpy slicing, numpy's IEEE float32 square root and closed forms."""
import math

import numpy as np
import pytest

from framefmt import FORMATS, expected, pack, values


@pytest.mark.parametrize("fmt", list(FORMATS))
@pytest.mark.parametrize("crop", [(0, 0, 64, 32), (6, 4, 48, 24), (3, 5, 40, 18), (0, 2, 66, 30)])
def test_frame_planes_equal_numpy_crops(orc, fmt, crop):
    rng = np.random.default_rng(sum(map(ord, fmt)) * 1000 + crop[0] * 37 + crop[1])
    W, H = 72, 40
    y, u, v = values(rng, fmt, W, H)
    x0, y0, w, h = crop
    for pd in ((8, 10) if FORMATS[fmt][3] == 8 else (10,)):
        planes = pack(fmt, y, u, v, pad=5, junk=rng)
        got = orc.frame_planes(fmt, *planes, x0, y0, w, h, pd)
        want = expected(fmt, y, u, v, x0, y0, w, h, pd)
        for g, e in zip(got, want):
            assert np.array_equal(g, e), (fmt, crop, pd)


def test_semi_planar_and_planar_layouts_agree(orc):
    """The same samples stored as NV12 and YUV420P (and P010 and YUV420P10) ingest to the same planes."""
    rng = np.random.default_rng(5)
    for a, b in (("nv12", "yuv420p"), ("p010", "yuv420p10")):
        y, u, v = values(rng, a, 64, 36)
        pa = orc.frame_planes(a, *pack(a, y, u, v, pad=3), 2, 2, 60, 32, 10)
        pb = orc.frame_planes(b, *pack(b, y, u, v, pad=1), 2, 2, 60, 32, 10)
        for x, z in zip(pa, pb):
            assert np.array_equal(x, z)


def test_plane_sse_equals_numpy(orc):
    rng = np.random.default_rng(9)
    a = rng.integers(0, 65536, 10007).astype(np.uint16)
    b = rng.integers(0, 65536, 10007).astype(np.uint16)
    want = int(((a.astype(np.int64) - b.astype(np.int64)) ** 2).sum())
    assert orc.plane_sse(a, b) == want
    fa = rng.uniform(-300, 300, 999).astype(np.float32)
    fb = rng.uniform(-300, 300, 999).astype(np.float32)
    # float form: the values truncated toward zero, as a C conversion to long long does
    want = int(((np.trunc(fa).astype(np.int64) - np.trunc(fb).astype(np.int64)) ** 2).sum())
    assert orc.plane_sse(fa, fb) == want


def test_color_sums_cover_every_10bit_pair(orc):
    """Every 10-bit (u, v) pair: the value sums, and each strength rounded as numpy's IEEE float32 square root rounds
    and then scaled by 2^23 without loss."""
    u = np.repeat(np.arange(1024, dtype=np.uint16), 1024)
    v = np.tile(np.arange(1024, dtype=np.uint16), 1024)
    got = orc.color_sums(u, v, 0)
    du = u.astype(np.int64) - 512
    dv = v.astype(np.int64) - 512
    s = np.sqrt((du * du + dv * dv).astype(np.float32))  # IEEE single precision, round to nearest
    fixed = (s.astype(np.float64) * 2.0 ** 23).astype(np.uint64)
    assert np.array_equal(s.astype(np.float64) * 2.0 ** 23, fixed.astype(np.float64))  # exact multiples of 2^-23
    assert int(got[0]) == int(u.astype(np.int64).sum()) and int(got[1]) == int(v.astype(np.int64).sum())
    assert int(got[2]) == int(fixed.sum(dtype=np.uint64))
    # 8-bit planes scaled by 4 to the 10-bit scale
    u8 = np.arange(256, dtype=np.uint16).repeat(3)
    v8 = u8[::-1].copy()
    g8 = orc.color_sums(u8, v8, 2)
    assert np.array_equal(g8, orc.color_sums(u8 << 2, v8 << 2, 0))


def test_strength_closed_forms(orc):
    assert orc.strength_fixed(0) == 0
    for a, b, c in ((3, 4, 5), (5, 12, 13), (8, 15, 17), (20, 21, 29)):
        assert orc.strength_fixed(a * a + b * b) == c << 23
    # neutral chroma has no colour strength
    n = np.full(100, 512, np.uint16)
    assert list(orc.color_sums(n, n, 0)) == [51200, 51200, 0]


def test_psnr_closed_forms(orc):
    for depth in (8, 10):
        assert orc.psnr(0, 1000, depth) == 6 * depth + 12  # the cap
        peak = (1 << depth) - 1
        assert orc.psnr(1000, 1000, depth) == pytest.approx(20 * math.log10(peak), abs=1e-12)
        assert orc.psnr(250 * 1000, 1000, depth) == pytest.approx(10 * math.log10(peak * peak / 250.0), abs=1e-12)
    # just below the cap and capped
    assert orc.psnr(1, 10**4, 10) == pytest.approx(min(10 * math.log10(1023 ** 2 * 1e4), 72), abs=1e-12)
