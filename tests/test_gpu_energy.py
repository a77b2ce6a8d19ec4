"""Window energies, lasting-texture sums and selects on the GPU against the CPU oracle (exact integers)."""
import numpy as np
import pytest

from conftest import banded_mosaic

pytestmark = pytest.mark.gpu


def pair(es, w, h, seed, frame=0):
    src = banded_mosaic(es, w, h, seed, frame=frame)
    d = es.SynthDesc(kind=1, seed=seed, full_w=w, full_h=h, w=w, h=h, depth=10, grain_q8=0, blur_q8=154,
                     quant_step=2, block=8, block_pct=50)
    return src, es.synth_frame_host(d, frame)


@pytest.mark.parametrize("win", [16, 32, 64])
@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("size", [(256, 144), (333, 211), (1920, 1080)])
def test_window_stats_exact(gpu, orc, win, fp64, size):
    w, h = size
    if w < win or h < win:
        pytest.skip("smaller than a window")
    print(f"case window_stats {w}x{h} win {win} fp64 {fp64}", flush=True)
    src, enc = pair(gpu, w, h, 40 + w)
    ref_sums, ref_stats, ref_mr, ref_md = orc.window_stats(src, enc, 0, win)
    windows = (w // win) * (h // win)
    dr, dd = gpu.DeviceArray.from_host(src), gpu.DeviceArray.from_host(enc)
    mr, md = gpu.DeviceArray((h, w), np.int32), gpu.DeviceArray((h, w), np.int32)
    mr.zero()
    md.zero()
    sums = gpu.DeviceArray((windows, 12), np.int64)
    gpu.window_stats(dr, dd, 0, w, h, win, mr, md, sums, fp64=fp64)
    got = sums.host()
    assert np.array_equal(got, ref_sums)
    H, W = (h // win) * win, (w // win) * win
    assert np.array_equal(mr.host()[:H, :W], ref_mr[:H, :W])
    assert np.array_equal(md.host()[:H, :W], ref_md[:H, :W])
    assert np.array_equal(gpu.window_values(got).view(np.uint64), ref_stats.view(np.uint64))


def test_window_stats_8bit_up2(gpu, orc):
    """An 8-bit pipeline's planes brought to the 10-bit scale (up = 2, src/main.cpp:446)."""
    rng = np.random.default_rng(5)
    src = rng.integers(16, 236, (96, 160)).astype(np.uint16)
    enc = np.clip(src.astype(int) + rng.integers(-3, 4, src.shape), 16, 235).astype(np.uint16)
    ref_sums, _, _, _ = orc.window_stats(src, enc, 2, 16, want_mid=False)
    sums = gpu.DeviceArray((ref_sums.shape[0], 12), np.int64)
    gpu.window_stats(gpu.DeviceArray.from_host(src), gpu.DeviceArray.from_host(enc), 2, 160, 96, 16, None, None, sums)
    assert np.array_equal(sums.host(), ref_sums)


@pytest.mark.parametrize("size", [(256, 144), (333, 211), (1920, 1080)])
def test_window_stats_f32_reference_outputs(gpu, orc, size):
    """es_window_stats_f32 writes what the reference's window_stats writes (src/kernels.cu:340-347): float mid planes
    and 8 floats per 16 x 16 window, each the exact value rounded once to float."""
    w, h = size
    print(f"case window_stats_f32 {w}x{h}", flush=True)
    src, enc = pair(gpu, w, h, 70 + w)
    _, ref_stats, ref_mr, ref_md = orc.window_stats(src, enc, 0, 16)
    H, W = (h // 16) * 16, (w // 16) * 16
    dr, dd = gpu.DeviceArray.from_host(src), gpu.DeviceArray.from_host(enc)
    mr, md = gpu.DeviceArray((h, w), np.float32), gpu.DeviceArray((h, w), np.float32)
    mr.zero()
    md.zero()
    out = gpu.DeviceArray((ref_stats.shape[0], 8), np.float32)
    gpu.window_stats_f32(dr, dd, 0, w, h, mr, md, out)
    assert np.array_equal(out.host().view(np.uint32), ref_stats.astype(np.float32).view(np.uint32))
    for got, ref in ((mr.host(), ref_mr), (md.host(), ref_md)):
        want = (ref[:H, :W].astype(np.float64) / 81.0).astype(np.float32)
        assert np.array_equal(got[:H, :W].view(np.uint32), want.view(np.uint32))
        assert not got[H:, :].any() and not got[:, W:].any()  # only the windows' area is written


def test_persist_f32_reference_outputs(gpu, orc):
    """es_persist_stats_f32 on float mid planes (mid = M / 81) against the oracle's exact sums on the M planes, scaled
    to the reference's units (src/kernels.cu:441-450): within float rounding."""
    rng = np.random.default_rng(11)
    h, w = 160, 240
    ring = [rng.integers(-90000, 90000, (h, w)).astype(np.int32) for _ in range(13)]
    mid_d = rng.integers(-90000, 90000, (h, w)).astype(np.int32)
    ref = orc.persist(ring, mid_d, 16).astype(np.float64) / np.array([78732.0, 78732.0, 236196.0, 236196.0, 236196.0])
    centre = 9
    dev_ring = np.empty((13, h, w), np.float32)
    for o in range(-6, 7):
        dev_ring[(centre + o + 13) % 13] = ring[o + 6] / 81.0
    out = gpu.DeviceArray((ref.shape[0], 5), np.float32)
    gpu.persist_stats_f32(gpu.DeviceArray.from_host(dev_ring), h * w, centre,
                          gpu.DeviceArray.from_host((mid_d / 81.0).astype(np.float32)), w, h, out)
    got = out.host().astype(np.float64)
    scale = np.abs(ref).max(axis=0)
    assert np.all(np.abs(got - ref) <= 2e-6 * scale + 1e-6 * np.abs(ref))


@pytest.mark.parametrize("win", [16, 64])
def test_persist_exact(gpu, orc, win):
    rng = np.random.default_rng(9)
    h, w = 192, 256
    ring = [rng.integers(-90000, 90000, (h, w)).astype(np.int32) for _ in range(13)]
    mid_d = rng.integers(-90000, 90000, (h, w)).astype(np.int32)
    ref = orc.persist(ring, mid_d, win)
    centre = 4  # slot of the centre frame in the device ring
    dev_ring = np.empty((13, h, w), np.int32)
    for o in range(-6, 7):
        dev_ring[(centre + o + 13) % 13] = ring[o + 6]
    dr = gpu.DeviceArray.from_host(dev_ring)
    dm = gpu.DeviceArray.from_host(mid_d)
    sums = gpu.DeviceArray(ref.shape, np.int64)
    gpu.persist_stats(dr, h * w, centre, dm, w, h, win, sums)
    assert np.array_equal(sums.host(), ref)


@pytest.mark.parametrize("n", [1, 2, 1000, 32400, 200003])
def test_select_rank(gpu, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 2**64 - 1, n, dtype=np.uint64)
    keys[: n // 3] = keys[0]  # ties
    d = gpu.DeviceArray.from_host(keys)
    s = np.sort(keys)
    for rank in {0, n // 2, int(0.8 * (n - 1)), n - 1}:
        assert gpu.select_rank(d, n, rank) == s[rank]
    k32 = rng.integers(0, 2**32 - 1, n, dtype=np.uint32)
    d32 = gpu.DeviceArray.from_host(k32)
    s32 = np.sort(k32)
    for rank in {0, n // 2, n - 1}:
        assert gpu.select_rank(d32, n, rank) == s32[rank]


def test_hist_quantiles(gpu, orc):
    rng = np.random.default_rng(11)
    pcts = [rng.normal(80, 30, 5000), rng.normal(0, 200, 77), np.array([]), np.full(10, 100.0), rng.uniform(-600, 600, 333)]
    hists = np.zeros((len(pcts), 20001), np.uint64)
    refs = []
    for i, p in enumerate(pcts):
        h = orc.hist_of(p)
        hists[i] = np.frombuffer(h.bins, np.uint64)
        refs.append([orc.hist_quantile(h, q) for q in (0.5, 0.1, 0.0, 1.0)])
    out, counts = gpu.hist_quantiles(gpu.DeviceArray.from_host(hists), len(pcts), [0.5, 0.1, 0.0, 1.0])
    for i in range(len(pcts)):
        assert counts[i] == len(pcts[i])
        for j in range(4):
            a, b = out[i, j], refs[i][j]
            assert (np.isnan(a) and np.isnan(b)) or a == b
