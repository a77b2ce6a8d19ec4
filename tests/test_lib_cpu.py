"""The C-ABI library without a GPU: it loads, exports every symbol include/encscore.h declares, and its host-only
entry points (CAMBI parameters, window figures, the synthetic generator) agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def test_library_exports_every_declared_symbol(es):
    header = open(os.path.join(ROOT, "include", "encscore.h")).read()
    names = set(re.findall(r"\b(es_\w+)\s*\(", header))
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(es.lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("size", [(1920, 1080), (3840, 2160), (1920, 800), (1800, 1080), (3840, 1600), (256, 144),
                                  (64, 64), (17, 9), (1, 1)])
def test_cambi_params_equal_the_oracles(es, orc, size):
    w, h = size
    a, b = es.cambi_params(w, h), orc.params(w, h)
    for f in ("window", "pad", "mask_index", "vlt", "band_base", "band_size"):
        assert getattr(a, f) == getattr(b, f), f
    assert list(a.tvi) == list(b.tvi) and list(a.k) == list(b.k) and list(a.sw) == list(b.sw)


def test_bad_arguments_are_errors_not_crashes(es):
    with pytest.raises(es.EsError):
        es.cambi_params(0, 10)
    with pytest.raises(es.EsError):
        es.cambi_params(100, 100, depth=12)
    with pytest.raises(es.EsError):
        es.cambi_params(8000, 4000)  # window above 65


def test_window_values_equal_the_oracles(es, orc):
    rng = np.random.default_rng(3)
    sums = rng.integers(-2**40, 2**40, (500, 12), dtype=np.int64)
    sums[:, [0, 1, 2, 3, 6, 7, 9, 10]] = np.abs(sums[:, [0, 1, 2, 3, 6, 7, 9, 10]])
    got = es.window_values(sums)
    ref = np.empty(8)
    for i in range(500):
        orc.lib.orc_window_values(sums[i].ctypes.data, ref.ctypes.data)
        assert np.array_equal(got[i].view(np.uint64), ref.view(np.uint64))


def test_synth_is_deterministic_and_in_range(es):
    d = es.SynthDesc(kind=1, seed=42, full_w=300, full_h=200, w=300, h=200, depth=10, grain_q8=256)
    a, b = es.synth_frame_host(d, 3, threads=1), es.synth_frame_host(d, 3, threads=4)
    assert np.array_equal(a, b) and a.min() >= 64 and a.max() <= 940
    assert not np.array_equal(a, es.synth_frame_host(d, 4))  # grain changes per frame
    d8 = es.SynthDesc(kind=1, seed=42, full_w=300, full_h=200, w=300, h=200, depth=8, grain_q8=256)
    a8 = es.synth_frame_host(d8, 3, bytes=1)
    assert a8.dtype == np.uint8 and a8.min() >= 16 and a8.max() <= 235
    c = es.SynthDesc(kind=1, seed=42, full_w=300, full_h=200, crop_x=20, crop_y=10, w=200, h=150, depth=10, grain_q8=256)
    assert np.array_equal(es.synth_frame_host(c, 3), a[10:160, 20:220])  # content is defined on the source's grid
    one = es.SynthDesc(kind=0, seed=0x5EED0001, full_w=192, full_h=108, w=192, h=108, depth=10, grain_q8=256)
    x = es.synth_frame_host(one, 0)
    assert 64 <= x.min() and x.max() <= 940


@pytest.mark.parametrize("name", ["dropin_check.cpp", "dropin_main.cpp"])
def test_dropin_headers_compile(name):
    """The C++ drop-in headers (gpu.hpp / cambi.hpp / cache.hpp over the C ABI) compile with the host compiler alone."""
    import subprocess
    src = os.path.join(ROOT, "encode-score_amd", "host", name)
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "encode-score_amd", "host"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_frame_view_follows_the_format_table(es, orc):
    """es_frame_view's (semi, bytes, shift, depth) per format is view_of's table (src/main.cpp:135-147)."""
    for name, code in es.FORMATS.items():
        f = es.frame_view(code, 0x1000, 0x2000, 0x3000, 4096, 2048)
        assert (f.semi, f.bytes, f.shift, f.depth) == orc.FORMATS[name], name
        assert (f.ypitch, f.cpitch) == (4096, 2048)
        assert (f.v is None) == bool(f.semi)
    with pytest.raises(es.EsError, match="unsupported pixel format"):
        es.frame_view(7, 0x1000, 0x2000, 0x3000, 64, 64)


def test_psnr_and_colour_means_equal_the_oracles(es, orc):
    rng = np.random.default_rng(11)
    for _ in range(200):
        n = int(rng.integers(1, 10**7))
        sse = int(rng.integers(0, 1023 ** 2)) * n // int(rng.integers(1, 1000))
        for depth in (8, 10):
            a, b = es.psnr(sse, n, depth), orc.psnr(sse, n, depth)
            assert a == b
    assert es.psnr(0, 10, 10) == 72.0
    sums = np.array([512 * 40, 600 * 40, 5 * 40 << 23], np.uint64)
    assert list(es.color_values(sums, 40)) == [512.0, 600.0, 5.0]


def test_frame_arguments_are_checked_before_any_device_work(es):
    f = es.frame_view("p010", 0x1000, 0x2000, 0, 4096, 4096)
    for bad in (dict(w=63), dict(h=0), dict(pd=9), dict(x0=-2)):
        kw = dict(x0=0, y0=0, pd=10, w=64, h=32)
        kw.update(bad)
        with pytest.raises(es.EsError):
            es.frame_planes(f, kw["x0"], kw["y0"], kw["pd"], 0x4000, 0x8000, 0xC000, kw["w"], kw["h"])
    f8 = es.frame_view("nv12", 0x1000, 0x2000, 0, 4096, 4096)
    with pytest.raises(es.EsError, match="depth"):
        es.frame_planes(es.frame_view("p010", 0x1000, 0x2000, 0, 4096, 4096), 0, 0, 8, 0x4000, 0x8000, 0xC000, 64, 32)
    with pytest.raises(es.EsError, match="aligned"):
        es.frame_planes(f8, 0, 0, 10, 0x4002, 0x8000, 0xC000, 64, 32)
    with pytest.raises(es.EsError):
        es.color_sums(0x1000, 0x2000, 10, 3, 0x3000)


def test_dropin_cambi_cache_class(orc, tmp_path):
    """The C++ drop-in CambiCache (host/cache.hpp, src/cache.hpp's interface) on the host: a score kept by one
    instance is found by the next, in the file the oracle's restatement of the naming names."""
    import subprocess
    src = tmp_path / "cache_check.cpp"
    src.write_text(r'''
#include "cache.hpp"
#include <cstdio>
int main(int argc, char **argv)
{
    try {
        {
            CambiCache c(argv[1], 0, 140, 1920, 800, 10);
            if (c.find(1001)) return 2;
            c.add(1001, 3.25);
            c.add(2002, 4.5);
        }
        CambiCache again(argv[1], 0, 140, 1920, 800, 10);
        const auto a = again.find(1001), b = again.find(2002), n = again.find(3003);
        std::printf("%s %g %g %d\n", again.file().c_str(), a ? *a : -1.0, b ? *b : -1.0, n ? 1 : 0);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
''')
    exe = tmp_path / "cache_check"
    lib_dir = os.path.join(ROOT, "encode-score_amd")
    r = subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(lib_dir, "host"),
                        str(src), "-L", lib_dir, "-lencscore_hip", "-Wl,-rpath," + lib_dir, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    source = tmp_path / "film.mkv"
    source.write_bytes(b"frames" * 5000)
    env = dict(os.environ, XDG_CACHE_HOME=str(tmp_path / "xdg"))
    r = subprocess.run([str(exe), str(source)], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr
    path, a, b, n = r.stdout.split()
    assert path == str(tmp_path / "xdg" / "encode-score" / orc.cambi_cache_name(source, 355, 0, 140, 1920, 800, 10))
    assert (float(a), float(b), n) == (3.25, 4.5, "0")
