"""es_hist_quantiles_noalloc through the C ABI against es_hist_quantiles, the entry point it mirrors: two calls in a row
(the second smaller, on the same static scratch), empty histograms (NaN) and the size limits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def histograms(es, n, seed):
    rng = np.random.default_rng(seed)
    h = np.zeros((n, es.HIST_BINS), np.uint64)
    for i in range(n):
        if i % 5 == 3:
            continue  # empty: NaN quantiles, count 0
        centre = rng.integers(0, es.HIST_BINS)
        bins = np.clip(rng.normal(centre, rng.uniform(5, 3000), rng.integers(1, 4000)).astype(np.int64), 0,
                       es.HIST_BINS - 1)
        np.add.at(h[i], bins, 1)
    return h


def same(a, b):
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def test_noalloc_quantiles_equal_the_allocating_form(gpu):
    for n, qs, seed in ((300, [0.5, 0.1, 0.9, 0.0, 1.0, 0.25, 0.75, 0.999], 1), (7, [0.5, 0.1, 0.33], 2)):
        print(f"case hist_quantiles_noalloc {n} histograms, {len(qs)} quantiles", flush=True)
        h = histograms(gpu, n, seed)
        d = gpu.DeviceArray.from_host(h)
        want, want_n = gpu.hist_quantiles(d, n, qs)
        got, got_n = gpu.hist_quantiles_noalloc(d, n, qs)
        assert same(got, want) and np.array_equal(got_n, want_n)
        assert np.isnan(got[3]).all() and got_n[3] == 0
    # the largest call it holds
    n = 4096
    h = np.zeros((n, gpu.HIST_BINS), np.uint64)
    h[np.arange(n), np.arange(n) * 4] = np.arange(n, dtype=np.uint64) + 1
    d = gpu.DeviceArray.from_host(h)
    qs = np.linspace(0, 1, 16)
    want, want_n = gpu.hist_quantiles(d, n, qs)
    got, got_n = gpu.hist_quantiles_noalloc(d, n, qs)
    assert same(got, want) and np.array_equal(got_n, want_n)


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    d = gpu.DeviceArray((1, gpu.HIST_BINS), np.uint64)  # the sizes are rejected before any memory is read
    with pytest.raises(gpu.EsError, match="no-allocation"):
        gpu.hist_quantiles_noalloc(d, 4097, [0.5])
    with pytest.raises(gpu.EsError, match="no-allocation"):
        gpu.hist_quantiles_noalloc(d, 4, np.linspace(0, 1, 17))
