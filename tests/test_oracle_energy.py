"""The window-energy and lasting-texture oracle pinned against a float64 numpy restatement of compare.py's
definitions (reference/encode_score/compare.py:77-108), the eigenvalues of the structure tensor, direct means, and the
reference's own invariant tests (tests/test_reference_compare.py) with their thresholds."""
import numpy as np
import pytest


def box64(a, k):
    """compare.box (compare.py:77-82) kept in float64: mean over k x k with the edges repeated."""
    r = k // 2
    c = np.pad(a.astype(np.float64), r, mode="edge").cumsum(0).cumsum(1)
    c = np.pad(c, ((1, 0), (1, 0)))
    return (c[k:, k:] - c[:-k, k:] - c[k:, :-k] + c[:-k, :-k]) / (k * k)


def window_sums(v, win):
    h, w = v.shape[0] // win * win, v.shape[1] // win * win
    return v[:h, :w].reshape(h // win, win, w // win, win).sum(axis=(1, 3))


def energies64(y, win):
    """compare.bands + window_energies (compare.py:91-108) in float64."""
    b3 = box64(y, 3)
    fine, mid = y - b3, b3 - box64(y, 9)
    gx, gy = np.zeros_like(b3), np.zeros_like(b3)
    gx[:, 1:-1], gy[1:-1, :] = b3[:, 2:] - b3[:, :-2], b3[2:, :] - b3[:-2, :]
    jxx, jyy, jxy = window_sums(gx * gx, win), window_sums(gy * gy, win), window_sums(gx * gy, win)
    return {"fine": fine, "mid": mid, "jxx": jxx, "jyy": jyy, "jxy": jxy,
            "E_fine": window_sums(fine**2, win), "E_mid": window_sums(mid**2, win),
            "oriented": np.sqrt((jxx - jyy) ** 2 + 4 * jxy**2)}


def texture(height=128, width=128, period=8):
    y, x = np.mgrid[0:height, 0:width]
    return (512 + 120 * np.sin((x + y) * 2 * np.pi / period)).astype(np.float32)


@pytest.mark.parametrize("win", [16, 32, 64])
def test_window_stats_match_float64_definitions(orc, win):
    rng = np.random.default_rng(win)
    r = np.clip(texture(200, 264) + rng.normal(0, 8, (200, 264)), 0, 1023).round().astype(np.uint16)
    d = np.clip(box64(r, 3).round(), 0, 1023).astype(np.uint16)
    sums, stats, mid_r, mid_d = orc.window_stats(r, d, 0, win)
    er, ed = energies64(r.astype(np.float64), win), energies64(d.astype(np.float64), win)
    grid = er["E_fine"].shape
    st = stats.reshape(grid[0], grid[1], 8)
    np.testing.assert_allclose(st[..., 0], er["E_fine"], rtol=1e-10)
    np.testing.assert_allclose(st[..., 1], er["E_mid"], rtol=1e-10)
    np.testing.assert_allclose(st[..., 2], ed["E_fine"], rtol=1e-10)
    np.testing.assert_allclose(st[..., 3], ed["E_mid"], rtol=1e-10)
    np.testing.assert_allclose(st[..., 4], window_sums(er["fine"] * ed["fine"], win), rtol=1e-9, atol=1e-6)
    np.testing.assert_allclose(st[..., 5], window_sums(er["mid"] * ed["mid"], win), rtol=1e-9, atol=1e-6)
    np.testing.assert_allclose(st[..., 6], er["oriented"], rtol=1e-9, atol=1e-6)
    np.testing.assert_allclose(st[..., 7], ed["oriented"], rtol=1e-9, atol=1e-6)
    H, W = grid[0] * win, grid[1] * win
    np.testing.assert_allclose(mid_r[:H, :W] / 81.0, er["mid"][:H, :W], atol=1e-9)
    # the scaled quantities are exact integers
    assert sums.dtype == np.int64


def test_oriented_is_the_eigenvalue_difference(orc):
    rng = np.random.default_rng(3)
    r = rng.integers(64, 941, (48, 64)).astype(np.uint16)
    sums, stats, _, _ = orc.window_stats(r, r, 0, 16)
    for i in range(sums.shape[0]):
        j = np.array([[sums[i, 6], sums[i, 8]], [sums[i, 8], sums[i, 7]]], np.float64) / 81.0
        lo, hi = np.linalg.eigvalsh(j)
        assert abs(stats[i, 6] - (hi - lo)) <= 1e-9 * max(1.0, hi)


def test_box_matches_a_direct_mean():
    """tests/test_reference_compare.py:33-38, on the float64 box the oracle's integers restate."""
    a = np.random.default_rng(7).random((20, 30))
    out = box64(a, 3)
    assert abs(out[5, 7] - a[4:7, 6:9].mean()) < 1e-12
    assert abs(out[0, 0] - np.pad(a, 1, mode="edge")[0:3, 0:3].mean()) < 1e-12


def stats_of(orc, y, win=16):
    y = np.clip(np.round(y), 0, 1023).astype(np.uint16)
    _, st, _, _ = orc.window_stats(y, y, 0, win, want_mid=False)
    return st


def test_grain_has_no_oriented_energy_and_stripes_have_a_lot(orc):
    """tests/test_reference_compare.py:42-46"""
    rng = np.random.default_rng(7)
    grain = stats_of(orc, 512 + rng.normal(0, 20, (128, 128)))
    stripes = stats_of(orc, texture())
    assert stripes[:, 6].mean() > 20 * grain[:, 6].mean()
    assert grain[:, 0].mean() > 0


def test_window_grid_shape(orc):
    """tests/test_reference_compare.py:48-50: a 96x160 picture has a 6 x 10 grid."""
    st = stats_of(orc, texture(96, 160))
    assert st.shape[0] == 60


def test_blurring_removes_fine_detail(orc):
    """tests/test_reference_compare.py:52-56"""
    rng = np.random.default_rng(8)
    sharp = texture() + rng.normal(0, 10, (128, 128))
    a, b = stats_of(orc, sharp), stats_of(orc, box64(sharp, 5))
    assert b[:, 0].sum() < 0.5 * a[:, 0].sum()


def test_persist_matches_direct_sums(orc):
    rng = np.random.default_rng(5)
    h, w = 64, 80
    ring = [rng.integers(-80000, 80000, (h, w)).astype(np.int32) for _ in range(13)]
    mid_d = rng.integers(-80000, 80000, (h, w)).astype(np.int32)
    got = orc.persist(ring, mid_d, 16)
    r = [x.astype(object) for x in ring]
    sa = r[0] + r[2] + r[4] + r[8] + r[10] + r[12]
    sb = r[1] + r[3] + r[5] + r[7] + r[9] + r[11]
    st = sa + sb
    exp = [window_sums(v, 16).astype(np.int64) for v in (r[6] * st, mid_d.astype(object) * st, sa * sb, sa * sa, sb * sb)]
    assert np.array_equal(got, np.stack([e.ravel() for e in exp], 1))
    vals = np.empty(5)
    import ctypes as C
    orc.lib.orc_persist_values(got[0].ctypes.data, vals.ctypes.data)
    # A = sa / (6 * 81), B = sb / (6 * 81), T = (A + B) / 2, mid = M / 81
    a, b = sa / 486.0, sb / 486.0
    t = (a + b) / 2
    assert abs(vals[0] - window_sums((r[6] / 81.0) * t, 16)[0, 0]) < 1e-6 * max(1, abs(vals[0]))
    assert abs(vals[2] - window_sums(a * b, 16)[0, 0]) < 1e-6 * max(1, abs(vals[2]))


def lasting_run(orc, make, n=20):
    """Runs the oracle over n frames of (source, encode) from make(f) and returns the merged accumulator."""
    run = orc.Run(400, 400, per_segment=1000, threads=8)
    for f in range(n):
        s, e = make(f)
        run.frame(np.clip(np.round(s), 64, 940).astype(np.uint16), np.clip(np.round(e), 64, 940).astype(np.uint16))
    return orc.accum_summary(run.merged())[0]


def test_a_clean_copy_keeps_the_texture_although_its_grain_is_gone(orc):
    """tests/test_reference_compare.py:141-144"""
    rng = np.random.default_rng(11)
    clean = texture(400, 400)
    got = lasting_run(orc, lambda f: (clean + rng.normal(0, 8, clean.shape), clean), n=14)
    assert got["n_lasting"] > 50
    assert got["median_lasting"] > 95


def test_fresh_grain_on_a_flattened_picture_keeps_none_of_it(orc):
    """tests/test_reference_compare.py:146-149"""
    rng = np.random.default_rng(12)
    clean = texture(400, 400)
    got = lasting_run(orc, lambda f: (clean + rng.normal(0, 8, clean.shape), 512 + rng.normal(0, 60, clean.shape)), n=14)
    assert abs(got["median_lasting"]) < 15


def test_a_picture_that_is_only_grain_has_no_persistent_windows(orc):
    """tests/test_reference_compare.py:151-155"""
    rng = np.random.default_rng(13)
    noise = lambda: 512 + rng.normal(0, 20, (400, 400))
    got = lasting_run(orc, lambda f: (noise(), noise()), n=14)
    assert got["n_lasting"] < 5 * 2  # two centre frames of 14


def test_without_later_frames_there_is_no_lasting_texture(orc):
    """tests/test_reference_compare.py:163-164: the first 12 frames give no centre frame."""
    clean = texture(400, 400)
    got = lasting_run(orc, lambda f: (clean, clean), n=12)
    assert got["n_lasting"] == 0 and np.isnan(got["median_lasting"])
