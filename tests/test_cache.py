"""The source-CAMBI cache (src/cache.cpp) on the host: the oracle's FNV-1a against the published test vectors, the
library's file name against the oracle's restatement of the key, the 16-byte record format, flushing, a record cut
short by an interrupted run, timestamps-less frames and an inert cache."""
import os
import struct

import numpy as np
import pytest

VERSION = 355  # this library's CAMBI version in the cache key (the reference's is 1)


def test_fnv1a64_published_vectors(orc):
    # Fowler/Noll/Vo's published FNV-1a 64-bit test vectors
    assert orc.fnv1a64(b"") == 0xcbf29ce484222325
    assert orc.fnv1a64(b"a") == 0xaf63dc4c8601ec8c
    assert orc.fnv1a64(b"foobar") == 0x85944171f73967e8


@pytest.mark.parametrize("size", [0, 1000, (1 << 20) + 5, 3 << 20])
def test_cache_file_name_and_records(es, orc, tmp_path, size):
    src = tmp_path / "source.mkv"
    rng = np.random.default_rng(size)
    src.write_bytes(rng.integers(0, 256, size, dtype=np.uint8).tobytes())
    c = es.CambiCache(src, 0, 140, 1920, 800, 10, directory=tmp_path / "cache")
    name = orc.cambi_cache_name(src, VERSION, 0, 140, 1920, 800, 10)
    assert c.file() == str(tmp_path / "cache" / name)
    assert c.find(5) is None
    scores = [(i * 1001, float(rng.uniform(0, 20))) for i in range(300)]
    for pts, v in scores:
        c.add(pts, v)
    c.add(es.NOPTS, 1.0)  # a frame without a timestamp is not kept
    assert c.find(1001) == scores[1][1] and c.find(es.NOPTS) is None
    # 240 records reach the file at once, the rest on close
    assert orc.cambi_cache_records(c.file()) == scores[:240]
    c.close()
    assert orc.cambi_cache_records(tmp_path / "cache" / name) == scores
    # a later run reads them back; another crop or depth is another file
    again = es.CambiCache(src, 0, 140, 1920, 800, 10, directory=tmp_path / "cache")
    assert all(again.find(p) == v for p, v in scores)
    other = es.CambiCache(src, 0, 140, 1920, 800, 8, directory=tmp_path / "cache")
    assert other.file() != again.file() and other.find(0) is None


def test_partial_record_is_dropped(es, orc, tmp_path):
    src = tmp_path / "s.y4m"
    src.write_bytes(b"x" * 4096)
    name = tmp_path / "c" / orc.cambi_cache_name(src, VERSION, 2, 4, 64, 32, 10)
    os.makedirs(tmp_path / "c")
    name.write_bytes(struct.pack("<qd", 7, 1.5) + struct.pack("<qd", 8, 2.5) + b"\x01\x02\x03")
    c = es.CambiCache(src, 2, 4, 64, 32, 10, directory=tmp_path / "c")
    assert c.find(7) == 1.5 and c.find(8) == 2.5
    assert os.path.getsize(name) == 32  # truncated to whole records
    c.add(9, 3.5)
    c.close()
    assert orc.cambi_cache_records(name) == [(7, 1.5), (8, 2.5), (9, 3.5)]


def test_inert_without_a_source(es, tmp_path):
    c = es.CambiCache(tmp_path / "missing.mkv", 0, 0, 64, 64, 10, directory=tmp_path / "c")
    assert c.file() == ""
    c.add(1, 2.0)
    assert c.find(1) is None
    c.close()
