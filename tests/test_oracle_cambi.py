"""The CAMBI oracle pinned against things other than itself: the reference's size formulas (the size sheet it gives
for every configuration), BT.1886's closed form, independent numpy restatements of the mask, mode filter, decimation
and 8-bit preprocessing, brute-force window counts, closed-form c-values and a sort-based top-k."""
import math

import numpy as np
import pytest

from conftest import banded_mosaic, staircase

SIZES = {  # w, h -> window, mask index, k per scale (Cambi::Cambi, src/cambi.cpp:29, :39-40)
    (1920, 1080): (33, 21, [1244160, 311040, 77760, 19440, 4896]),
    (3840, 2160): (65, 24, [4976640, 1244160, 311040, 77760, 19440]),
    (1920, 800): (29, 21, [921600, 230400, 57600, 14400, 3600]),
    (1920, 1040): (33, 21, [1198080, 299520, 74880, 18720, 4680]),
    (1800, 1080): (31, 21, [1166400, 291600, 72900, 18225, 4610]),
    (3840, 1600): (59, 24, [3686400, 921600, 230400, 57600, 14400]),
}


@pytest.mark.parametrize("size", list(SIZES))
def test_params_follow_the_reference_formulas(orc, size):
    w, h = size
    p = orc.params(w, h)
    window, index, ks = SIZES[size]
    assert p.window == window and p.pad == window // 2 and p.mask_index == index
    assert list(p.k) == ks
    # the formulas themselves, restated
    assert p.window == ((65 * (w + h) // 375) >> 4) | 1
    n = (w >> 6) * (h >> 6)
    assert p.mask_index == (49 + 3 * (max(0, math.ceil(math.log2(n))) - 11) - 1) >> 1
    sw, sh = w, h
    for s in range(5):
        assert (p.sw[s], p.sh[s]) == (sw, sh)
        assert p.k[s] == min(max(int(0.6 * sw * sh), 1), sw * sh)
        sw, sh = (sw + 1) >> 1, (sh + 1) >> 1


def bt1886(sample):
    """ITU-R BT.1886 EOTF on the limited 10-bit range, white 300, black 0.01: an independent numpy restatement."""
    v = (np.clip(sample, 64, 940) - 64) / 876.0
    g = 2.4
    lw, lb = 300.0 ** (1 / g), 0.01 ** (1 / g)
    a, b = (lw - lb) ** g, lb / (lw - lb)
    return a * np.maximum(v + b, 0) ** g


def test_luminance_is_bt1886(orc):
    assert abs(orc.lib.orc_bt1886_luminance(64) - 0.01) < 1e-12
    assert abs(orc.lib.orc_bt1886_luminance(940) - 300.0) < 1e-9
    for s in (0, 64, 100, 333, 512, 800, 940, 1023):
        assert abs(orc.lib.orc_bt1886_luminance(s) - bt1886(s)) <= 1e-12 * bt1886(s)


@pytest.mark.parametrize("diff", [1, 2, 3, 4])
def test_tvi_is_the_last_visible_step(orc, diff):
    """The scan returns the code where a `diff` step stops being a 1.9% luminance rise; the condition holds on every
    code below it and fails on every code above, so libvmaf's bisection lands on the same code."""
    t = orc.lib.orc_tvi_for_diff(diff, 0.019)
    s = np.arange(64, 941)
    visible = bt1886(s + diff) - bt1886(s) > 0.019 * bt1886(s)
    assert visible[: t - 64 + 1].all() and not visible[t - 64 + 1:].any()
    p = orc.params(1920, 1080)
    assert p.tvi[diff - 1] == t + 4


def test_band_and_visibility_threshold(orc):
    p = orc.params(1920, 1080)
    assert p.vlt == 0 and p.band_base == 0 and p.band_size == p.tvi[3] + 1
    assert list(p.weight) == [1, 2, 3, 4]


def test_bt1886_defining_identities(orc):
    """BT.1886 fixes a and b by L(V = 0) = black and L(V = 1) = white; the codes below the limited range's foot and
    above its head are clipped to it; luminance rises strictly over the range."""
    L = np.array([orc.lib.orc_bt1886_luminance(s) for s in range(1024)])
    assert L[64] == pytest.approx(0.01, rel=1e-12) and L[940] == pytest.approx(300.0, rel=1e-12)
    assert (L[:64] == L[64]).all() and (L[940:] == L[940]).all()
    assert (np.diff(L[64:941]) > 0).all()
    # the exponent: between the ends, L^(1/2.4) is affine in the code
    r = L[64:941] ** (1 / 2.4)
    assert np.allclose(np.diff(r), (r[-1] - r[0]) / 876, rtol=1e-9)


def tvi_bisection(lum, diff, thr, foot=64, head=940):
    """libvmaf's get_tvi_for_diff as published: bisection over the luma range for the code whose `diff` step is a
    visible rise (delta L > thr * L) while the step from one code higher is not."""
    cond = lambda s: lum(s + diff) - lum(s) > thr * lum(s)
    lo, hi = foot, head
    while True:
        mid = lo + (hi - lo) // 2
        if not cond(mid):
            hi = mid
        elif cond(mid + 1):
            lo = mid
        else:
            return mid


@pytest.mark.parametrize("thr", [0.015, 0.019, 0.025, 0.03, 0.05])
@pytest.mark.parametrize("diff", [1, 2, 3, 4])
def test_tvi_bisection_equals_the_scan(orc, diff, thr):
    """The visibility condition holds on a prefix of the range and fails after it (the relative step of a power law
    falls with the level), so the bisection's answer is unique and equals the oracle's scan; it lies inside the range,
    away from the ends where variants of the search could differ."""
    lum = orc.lib.orc_bt1886_luminance
    t = orc.lib.orc_tvi_for_diff(diff, thr)
    assert tvi_bisection(lum, diff, thr) == t
    cond = np.array([lum(s + diff) - lum(s) > thr * lum(s) for s in range(64, 941)])
    k = t - 64 + 1
    assert cond[:k].all() and not cond[k:].any()
    assert 64 < t < 940 - diff


def test_tvi_values_at_the_reference_threshold(orc):
    """tvi + 4 for diffs 1..4 at 1.9% (src/cambi.cpp:46-49): the band every configuration uses."""
    assert list(orc.params(1920, 1080).tvi) == [182, 309, 436, 563]


@pytest.mark.parametrize("vlt", [0, 1, 11, 40, 63, 64])
def test_visibility_threshold_reading_leaves_limited_range_results_alone(es, orc, vlt):
    """get_vlt_luma(0.0) (src/cambi.cpp:50): every code's luminance is at least the black level (0.01 cd/m^2) > 0, so code 0
    already reaches the threshold; any reading of the search that gives a code up to the foot (64) derives
    band_base = max(vlt - 11, 0) <= 53 (src/cambi.cpp:51-52) and passes value + d + 1 > vlt for every sample of the
    limited range: the c-values and the score of limited-range pictures do not depend on it."""
    assert min(orc.lib.orc_bt1886_luminance(s) for s in range(1024)) > 0.0
    for img in (banded_mosaic(es, 160, 96, 8, quant_step=4, quant_where=1), staircase(48, 120, base=64)):
        h, w = img.shape
        assert img.min() >= 64
        p0 = orc.params(w, h)
        ref, ref_cv = orc.cambi(p0, img, want_cvalues=True)
        p = orc.params(w, h)
        p.vlt = vlt
        p.band_base = max(vlt - 11, 0)
        p.band_size = p.tvi[3] + 1 - p.band_base
        got, got_cv = orc.cambi(p, img, want_cvalues=True)
        assert got.score == ref.score
        for s in range(5):
            assert np.array_equal(got_cv[s], ref_cv[s])


def mask_numpy(img, index):
    h, w = img.shape
    a = img.astype(np.int64)
    right = np.ones_like(a, bool)
    right[:, :-1] = a[:, :-1] == a[:, 1:]
    down = np.ones_like(a, bool)
    down[:-1, :] = a[:-1, :] == a[1:, :]
    flat = np.pad((right & down).astype(np.int64), 3)
    c = np.pad(flat.cumsum(0).cumsum(1), ((1, 0), (1, 0)))
    count = c[7:, 7:] - c[:-7, 7:] - c[7:, :-7] + c[:-7, :-7]
    return (count > index).astype(np.uint8)


def mode_numpy(img):
    a = img.astype(np.int64)
    h, w = a.shape

    def mode3(x, y, z):
        m = np.minimum(np.minimum(x, y), z)
        return np.where((x == y) | (x == z), x, np.where(y == z, y, m))

    hm = a.copy()
    hm[:, 1:-1] = mode3(a[:, :-2], a[:, 1:-1], a[:, 2:])
    out = a.copy()
    out[1:-1, :] = mode3(hm[:-2, :], hm[1:-1, :], hm[2:, :])
    return out.astype(np.uint16)


def pictures(es):
    rng = np.random.default_rng(1)
    return [staircase(40, 70), rng.integers(300, 304, (37, 53)).astype(np.uint16),
            banded_mosaic(es, 128, 96, 3, quant_step=4, quant_where=1), np.full((9, 5), 77, np.uint16),
            rng.integers(64, 941, (31, 47)).astype(np.uint16)]


def test_mask_mode_decimation_match_numpy(es, orc):
    for img in pictures(es):
        h, w = img.shape
        p = orc.params(w, h)
        assert np.array_equal(orc.spatial_mask(img, p.mask_index), mask_numpy(img, p.mask_index))
        assert np.array_equal(orc.mode_filter(img), mode_numpy(img))
        ims, mks = orc.pyramid(p, img)
        im, mk = mode_numpy(img), mask_numpy(img, p.mask_index)
        for s in range(5):
            if s:
                im, mk = mode_numpy(im[::2, ::2]), mk[::2, ::2]
            assert np.array_equal(ims[s], im) and np.array_equal(mks[s], mk)


def test_from8_matches_numpy(orc):
    rng = np.random.default_rng(2)
    img = rng.integers(0, 256, (23, 31)).astype(np.uint16)
    a = img.astype(np.int64) << 2
    exp = a.copy()
    exp[:-1, :-1] = (a[:-1, :-1] + a[:-1, 1:] + a[1:, :-1] + a[1:, 1:]) >> 2
    exp[:-1, -1] = (a[:-1, -1] + a[1:, -1]) >> 1
    exp[-1, :-1] = (a[-1, :-1] + a[-1, 1:]) >> 1
    assert np.array_equal(orc.from8(img), exp.astype(np.uint16))


def test_cvalues_sliding_equals_brute_force(es, orc):
    for img in pictures(es) + [banded_mosaic(es, 200, 120, 9, quant_step=2)]:
        h, w = img.shape
        p = orc.params(w, h)
        ims, mks = orc.pyramid(p, img)
        for s in range(5):
            a = orc.cvalues(p, ims[s], mks[s], brute=True)
            b = orc.cvalues(p, ims[s], mks[s], brute=False)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_cvalue_of_a_two_level_step_is_the_closed_form(orc):
    """Dark two-level step (diff 1 counts below code 178): at a boundary pixel c = p0 p1 / (p0 + p1) in single
    precision, p0 and p1 the two levels' counts in its window."""
    h, w = 40, 60
    img = np.where(np.arange(w)[None, :] < 30, 120, 121).repeat(h, 0).astype(np.uint16)
    p = orc.params(w, h)
    cv = orc.cvalues(p, img, np.ones((h, w), np.uint8))
    pad = p.pad
    lut = orc.recip_lut(p.lut_size)
    for y in (0, 7, 20, 39):
        for x in (27, 29, 30, 32):
            rows = min(y + pad, h - 1) - max(y - pad, 0) + 1
            cols0 = np.arange(max(x - pad, 0), min(x + pad, w - 1) + 1)
            p_low, p_high = rows * np.sum(cols0 < 30), rows * np.sum(cols0 >= 30)
            p0, p1 = (p_low, p_high) if x < 30 else (p_high, p_low)
            exp = np.float32(p0 * p1) * lut[p0 + p1] if p1 else np.float32(0)
            assert cv[y, x] == exp
    assert cv[:, : 30 - pad - 1].max() == 0


def test_reciprocal_table_is_correctly_rounded(orc):
    """The reciprocal table (src/kernels.cu:5, :200) holds 1/i by design: every entry up to the 4K window's
    (2 * 32 + 1)^2 + 1 = 4226 is the correctly rounded float32 quotient, checked against the exact rational 1/i.
    The device computes the same quotient with a correctly rounded division."""
    from fractions import Fraction
    n = 4226
    lut = orc.recip_lut(n)
    assert lut[0] == 0
    for i in range(1, n):
        v = float(lut[i])
        exact = Fraction(1, i)
        # no float32 lies strictly closer to 1/i than v: its neighbours are at least as far
        lo = float(np.nextafter(np.float32(v), np.float32(0)))
        hi = float(np.nextafter(np.float32(v), np.float32(2)))
        err = abs(Fraction(v) - exact)
        assert err <= abs(Fraction(lo) - exact) and err <= abs(Fraction(hi) - exact), i
    assert np.array_equal(lut[1:], np.float32(1) / np.arange(1, n, dtype=np.float32))


def test_flat_picture_has_full_mask_and_no_banding(orc):
    img = np.full((72, 128), 500, np.uint16)
    p = orc.params(128, 72)
    _, mks = orc.pyramid(p, img)
    assert all(m.all() for m in mks)
    r = orc.cambi(p, img)
    assert r.score == 0 and all(c == 0 for c in r.count)


def test_values_above_the_band_have_no_cvalue(orc):
    img = staircase(48, 96, base=600)  # every code above tvi[3] - 4
    p = orc.params(96, 48)
    assert orc.cambi(p, img).score == 0


@pytest.mark.parametrize("n,positive", [(1000, 0.3), (1000, 0.9), (5, 1.0), (77777, 0.75)])
def test_pool_matches_a_sort(orc, n, positive):
    rng = np.random.default_rng(n)
    c = (rng.integers(1, 5, n) * rng.integers(1, 500, n) / rng.integers(2, 600, n)).astype(np.float32)
    c = np.maximum(c, np.float32(0.5))
    c[rng.random(n) > positive] = 0
    k = min(max(int(0.6 * n), 1), n)
    r = orc.pool(c, k)
    pos = c[c > 0]
    assert r.count[0] == pos.size
    assert r.sum_fx[0] == int(np.sum(pos.astype(np.float64) * 2**24))
    top = np.sort(c)[::-1][:k].astype(np.float64)
    if pos.size > k:
        thr = np.sort(pos)[::-1][k - 1]
        assert r.thr_bits[0] == thr.view(np.uint32)
        assert r.above[0] == np.sum(pos > thr)
    assert math.isclose(r.pooled[0], math.fsum(top) / k, rel_tol=1e-12)


def test_fixed24_is_exact(orc):
    rng = np.random.default_rng(4)
    for c in np.concatenate([[0.5, 0.75, 1.0, 4225.0], rng.uniform(0.5, 8000, 1000)]).astype(np.float32):
        fx = orc.lib.orc_fixed24(float(c))
        assert fx / 2**24 == float(c)


def test_score_weights_and_cap(orc):
    img = staircase(72, 200)
    p = orc.params(200, 72)
    r = orc.cambi(p, img)
    total = sum(r.pooled[s] * (16, 8, 4, 2, 1)[s] for s in range(5))
    assert r.score == min(total / (p.window * p.window), 1000.0)
