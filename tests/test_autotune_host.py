"""Autotuner host logic: cache keys, size buckets, nearest-bucket lookup,
JSON round trip and the candidate launch shapes (no GPU)."""
import json

from rigidmultiblobswall_amd import autotune


def test_keys_and_buckets():
    k1 = autotune.cache_key("gfx950:sramecc+:xnack-", 1, "fp32_blocked", 1, 10752.0 ** 2, False)
    assert k1.startswith("gfx950|mask1|fp32_blocked|k1|single|p")
    assert autotune.pairs_bucket(2.0 ** 20) == 20.0
    assert autotune.pairs_bucket(2.0 ** 20.2) == 20.0
    assert autotune.pairs_bucket(2.0 ** 20.3) == 20.5
    # fp32 and its long name share entries
    assert autotune.cache_key("gfx950", 15, "fp32_f64acc", 4, 1e9, True) == \
        autotune.cache_key("gfx950", 15, "fp32", 4, 1e9, True)


def test_cache_round_trip_and_nearest(tmp_path):
    path = str(tmp_path / "t.json")
    c = autotune.TuningCache(path)
    key = autotune.cache_key("gfx950", 1, "fp64", 1, 2.0 ** 26, False)
    c.store(key, {"chunk_len": 1024, "max_wgs": 0, "rhs_group": 0, "ms": 1.0})
    c.save()
    d = json.load(open(path))
    assert d["entries"][key]["chunk_len"] == 1024
    c2 = autotune.TuningCache(path)
    assert c2.lookup(key)["chunk_len"] == 1024
    # half an octave away: nearest entry; two octaves away: none
    near = autotune.cache_key("gfx950", 1, "fp64", 1, 2.0 ** 26.5, False)
    far = autotune.cache_key("gfx950", 1, "fp64", 1, 2.0 ** 28.5, False)
    other = autotune.cache_key("gfx950", 1, "fp32_blocked", 1, 2.0 ** 26, False)
    assert c2.lookup(near)["chunk_len"] == 1024
    assert c2.lookup(far) is None
    assert c2.lookup(other) is None


def test_candidates():
    ch = autotune.chunk_candidates(10752)
    assert 0 not in ch and 256 in ch and 10752 in ch and 448 in ch   # 448 = 10752 / 24
    assert 64 in ch and 128 in ch                                    # sub-tile symmetric items
    assert all(c % 64 == 0 for c in ch)
    assert autotune.chunk_candidates(100) == [64, 128]
    cs = autotune.candidates(nmax=10752, k=1, cu=256, ragged=False)
    assert {c["rhs_group"] for c in cs} == {0}
    assert {c["max_wgs"] for c in cs} == {0, 1024, 2048, 2560}
    cs4 = autotune.candidates(nmax=32768, k=4, cu=256, ragged=False)
    assert {c["rhs_group"] for c in cs4} == {0, 1, 2, 4}
    cs2 = autotune.candidates(nmax=100, k=2, cu=256, ragged=True)
    assert {c["rhs_group"] for c in cs2} == {0, 1}
    assert {c["max_wgs"] for c in cs2} == {0, 1024, 2048, 2560, 4096}
