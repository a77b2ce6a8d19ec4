"""bench.py's host logic on the CPU: the frame pool and the autotuner's frames under any --steps / --warmup, and the
configurations' shapes and lengths."""
import importlib

import pytest

bench = importlib.import_module("bench")


@pytest.mark.parametrize("steps,warmup", [(20, 5), (1, 0), (8, 2), (240, 16), (1000, 50)])
def test_pool_holds_the_autotuners_frames(steps, warmup):
    # the autotuner times three batches of its largest batch on the pool's first frames, whatever the run's length
    pool = bench.frame_pool(steps, warmup, tune=1, batch=32, cap=256)
    assert pool >= bench.TUNE_FRAMES >= 3 * bench.TUNE_BATCH_MAX
    assert pool <= max(256, bench.TUNE_FRAMES)


def test_pool_covers_a_short_run_without_reuse():
    # a run shorter than the cap gets a distinct frame for every frame step
    assert bench.frame_pool(3, 1, tune=1, batch=32, cap=512) == 4 * bench.TUNE_BATCH_MAX
    assert bench.frame_pool(3, 1, tune=0, batch=8, cap=256) == 4 * 8  # no autotuner: the run's frames only
    assert bench.frame_pool(30, 2, tune=0, batch=8, cap=512) == 256


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("total", [96, 160, 800, 1440])
def test_configuration_lengths(es, n, total):
    src, encs, sbytes, shift, text = bench.config(n, es, total)
    lengths = [length for _, length in encs]
    # the longest encode runs the whole run, none runs past it, and every one reaches the lasting window (13 frames)
    assert max(lengths) == total
    assert all(13 <= length <= total for length in lengths)
    assert lengths == sorted(lengths, reverse=True)
    for d, _ in encs:
        assert d.w % 2 == 0 and d.h % 2 == 0 and d.w <= src.w and d.h <= src.h
    assert (sbytes, shift) == ((1, 2) if n == 4 else (2, 0))
    assert text.startswith(f"C{n}")


def test_config3_and_5_keep_the_reference_proportions(es):
    _, encs, _, _, _ = bench.config(3, es, 1440)
    assert [length for _, length in encs] == [1440, 1100, 720, 333]
    _, encs, _, _, _ = bench.config(5, es, 1440)
    assert [length for _, length in encs][0] == 1440 and [length for _, length in encs][-1] == 500


@pytest.mark.parametrize("cfg,w,h,stage", [(1, 1920, 1080, "cambi_cvalues"), (2, 3840, 2160, "cambi_cvalues"),
                                          (2, 3840, 2160, "window_stats"), (2, 3840, 2160, "persist"),
                                          (3, 1920, 1080, "cambi_cvalues"), (4, 1920, 1080, "cambi_cvalues")])
def test_pmc_summaries_hold_the_profiled_kernels(cfg, w, h, stage):
    """The committed rocprofv3 summaries the bench's roofline reads: HBM bytes, counters and the profiled run's input
    bytes per launch for each kernel it may name."""
    rec, b_in, src = bench.pmc_profile(cfg, w, h, bench.kernel_name(stage, 32, True))
    assert rec is not None, src
    assert rec["bytes"] > 0 and b_in > 0 and src.startswith("profiles/")
    if stage == "cambi_cvalues":
        d = rec["derived"]
        assert 0 < d["wait_any_per_wave_cycle"] < 1 and d["lds_bank_conflict_rate"] > 0 and rec["mfma_i8_ops"] > 0
