"""Histogram, quantile, accumulator and score of the oracle against the reference's own src/stats.hpp (compiled from
the reference tree into oracle/_ref), numpy's linear quantile and the known answers of the reference's tests."""
import ctypes as C
import math

import numpy as np
import pytest


def test_histogram_matches_reference_stats_hpp(orc):
    ref = orc.ref_stats()
    if ref is None:
        pytest.skip("oracle/_ref was not built (the reference tree is absent)")
    rng = np.random.default_rng(17)
    for pcts in (rng.normal(90, 40, 10000), rng.uniform(-700, 700, 999), np.array([-500.025, 499.975, 0.025, -0.025]),
                 np.array([12.5]), np.array([])):
        pcts = np.ascontiguousarray(pcts, np.float64)
        bins = np.zeros(20001, np.uint64)
        count = C.c_uint64()
        qs = np.array([0.5, 0.1, 0.0, 1.0, 0.37])
        out = np.zeros(qs.size)
        ref.ref_hist(pcts.ctypes.data, pcts.size, bins.ctypes.data, C.byref(count), qs.ctypes.data, qs.size, out.ctypes.data)
        h = orc.hist_of(pcts)
        assert np.array_equal(np.frombuffer(h.bins, np.uint64), bins) and h.n == count.value
        for q, r in zip(qs, out):
            g = orc.hist_quantile(h, q)
            assert (math.isnan(g) and math.isnan(r)) or g == r


def accum(orc, frames=10, vmaf=None, cambi=0.0, cambi_src=0.0, fine_d=100.0, fine_r=100.0, lasting=()):
    a = orc.Accum()
    orc.lib.orc_accum_clear(C.byref(a))
    a.frames = frames
    a.cambi, a.cambi_source = cambi, cambi_src
    a.all_d[0], a.all_r[0] = fine_d, fine_r
    for p in lasting:
        orc.lib.orc_hist_add(C.byref(a.lasting), float(p))
    parts = orc.Parts()
    orc.lib.orc_accum_score(C.byref(a), float("nan") if vmaf is None else vmaf * frames, C.byref(parts))
    return parts


def test_score_known_answers(orc):
    """tests/test_reference_compare.py:171-189 restated on Accum::score (src/stats.hpp:70-82)."""
    full = [100.0] * 500
    p = accum(orc, vmaf=100.0, cambi=5.0, cambi_src=5.0, lasting=full)
    assert (p.fidelity, p.texture, p.banding, p.grain) == (100.0, 100.0, 100.0, 100.0) and abs(p.score - 100) < 1e-12
    p = accum(orc, vmaf=100.0, cambi=30.0, cambi_src=5.0, lasting=full)  # 2.5 CAMBI per frame above the source
    assert abs(p.banding - 50.0) < 1e-12 and abs(p.score - (100 - 50 * 0.25)) < 1e-12
    assert accum(orc, vmaf=100.0, cambi=0.0, cambi_src=5.0, lasting=full).banding == 100.0
    assert abs(accum(orc, vmaf=100.0, fine_d=16.0, fine_r=100.0, lasting=full).grain - 40.0) < 1e-12
    assert abs(accum(orc, vmaf=100.0, fine_d=169.0, fine_r=100.0, lasting=full).grain - 70.0) < 1e-12
    p = accum(orc, vmaf=100.0, lasting=[80.0] * 450 + [-20.0] * 50)
    assert abs(p.texture - 40.0) < 1e-12
    p = accum(orc, vmaf=90.0, lasting=[0.0] * 3)
    assert math.isnan(p.texture)
    rest = 1 - 0.30
    assert abs(p.score - (90 * 0.30 + 100 * (rest - 0.30)) / rest) < 1e-12


def test_score_matches_reference_stats_hpp(orc):
    ref = orc.ref_stats()
    if ref is None:
        pytest.skip("oracle/_ref was not built (the reference tree is absent)")
    rng = np.random.default_rng(19)
    for n in (0, 99, 100, 1234):
        lasting = np.ascontiguousarray(rng.normal(70, 25, n))
        frames, vmaf, cambi, src, fd, fr = 48, 4500.0, 71.5, 12.25, 8123.5, 9001.25
        parts = np.zeros(5)
        ref.ref_score(frames, vmaf, cambi, src, fd, fr, lasting.ctypes.data, n, parts.ctypes.data)
        p = accum(orc, frames=frames, vmaf=vmaf / frames, cambi=cambi, cambi_src=src, fine_d=fd, fine_r=fr, lasting=lasting)
        for g, r in zip((p.fidelity, p.texture, p.banding, p.grain, p.score), parts):
            assert (math.isnan(g) and math.isnan(r)) or g == r


@pytest.mark.parametrize("n", [1, 2, 7, 100, 8040])
def test_quantile_is_numpys_linear(orc, n):
    v = np.random.default_rng(n).gamma(2.0, 3.0, n)
    assert orc.quantile_linear(v, 0.8) == pytest.approx(np.quantile(v, 0.8), rel=1e-14, abs=1e-14)


@pytest.mark.parametrize("n", [2, 7, 100, 8040, 32400])
def test_quantile_f32_rounds_numpys_linear_quantile(orc, n):
    """On float keys whose neighbouring order statistics are within a factor 2 (b - a is then exact in float), the
    reference's evaluation is numpy's linear quantile rounded once to float."""
    v = np.random.default_rng(n).gamma(4.0, 300.0, n).astype(np.float32)
    c = np.sort(v)
    lo = int(0.8 * (n - 1))
    if lo + 1 < n:
        assert c[lo] >= c[lo + 1] / 2
    assert orc.quantile_f32(v, 0.8) == np.float32(np.quantile(v.astype(np.float64), 0.8))


def test_quantile_f32_near_tie():
    """Keys one float ulp apart (src/main.cpp:288 rounds the interpolated value to float): with 9 keys the threshold
    sits 0.4 ulp above the rank-6 key and rounds back onto it, so that window is in the mask; kept in double it would
    not be."""
    orc = __import__("conftest").load_oracle()
    one, next_up = np.float32(1.0), np.nextafter(np.float32(1.0), np.float32(2.0))
    v = np.array([0.1, 0.2, 0.3, 0.4, 0.5, 0.6, one, next_up, 2.0], np.float32)
    q = orc.quantile_f32(v, 0.8)
    assert q == 1.0
    assert int(((v >= q) & (v > 0)).sum()) == 3
    assert orc.quantile_linear(v.astype(np.float64), 0.8) > 1.0  # the FP64 threshold would leave it out
    # 12 keys: 0.8 ulp above, rounded up onto the next key
    v = np.array([0.1] * 8 + [one, next_up, 3.0, 4.0], np.float32)
    assert orc.quantile_f32(v, 0.8) == next_up


def test_quantile_f32_difference_in_float():
    """b - a is a float difference: 2^25 - 1 rounds to 2^25 before the interpolation (q = 1 + 0.4 * 2^25 rounded
    to float), where an exact difference would give 13421773."""
    orc = __import__("conftest").load_oracle()
    v = np.array([0.0] * 6 + [1.0, 2.0 ** 25, 2.0 ** 26], np.float32)
    assert orc.quantile_f32(v, 0.8) == np.float32(13421774.0)
    assert np.float32(1.0 + (2.0 ** 25 - 1.0) * 0.4) == np.float32(13421773.0)


def test_structure_mask_picks_the_top_windows(orc):
    """tests/test_reference_compare.py:59-68: the windows at or above the quantile and above zero."""
    grid = np.zeros((10, 10))
    grid[2:4, 3:6] = 5.0
    q = orc.quantile_linear(grid.ravel(), 1 - 0.06)
    mask = (grid >= q) & (grid > 0)
    assert mask.sum() == 6 and mask[2:4, 3:6].all()
    flat = np.zeros((5, 5))
    assert not ((flat >= orc.quantile_linear(flat.ravel(), 0.8)) & (flat > 0)).any()


def test_tree_sum(orc):
    rng = np.random.default_rng(23)
    for n in (1, 255, 256, 257, 32400):
        v = rng.gamma(1.5, 1e5, n)
        assert orc.tree_sum256(v) == pytest.approx(math.fsum(v), rel=1e-13)
        ints = np.floor(v)
        assert orc.tree_sum256(ints) == math.fsum(ints)


def test_hist_bins_round_half_away_from_zero(orc):
    """lround((clamp(p) + 500) / 0.05) (src/stats.hpp:16): the exact double quotient rounded half away from zero."""
    from decimal import ROUND_HALF_UP, Decimal
    rng = np.random.default_rng(29)
    for pct in [-500.0, 500.0, 0.0, 0.025, -0.025, 0.075, 1e9, -1e9, 12.475] + list(rng.uniform(-510, 510, 2000)):
        x = (min(max(pct, -500.0), 500.0) - -500.0) / 0.05
        exp = int(Decimal(x).quantize(Decimal(1), rounding=ROUND_HALF_UP))
        assert orc.lib.orc_hist_bin(pct) == exp
