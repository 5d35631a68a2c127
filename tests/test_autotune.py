"""The autotuner's host logic: the deterministic tie rule, the cache file and
its keys (no GPU needed)."""
import json

from partconv_amd import autotune as A


def test_tie_rule_is_deterministic():
    # within 5% of the fastest: the default wins
    assert A.choose([1, 2, 0], [10.0, 9.8, 12.0], 1) == 1
    # the default clearly slower: the first near-best candidate in order
    assert A.choose([1, 2, 0], [11.0, 10.0, 10.2], 1) == 2
    assert A.choose([2, 4, 8], [5.0, 4.0, 4.1], 2) == 4
    assert A.choose([2, 4, 8], [5.0, 4.1, 4.0], 2) == 4   # 4.1 is within 5% of 4.0, earlier in order
    # every choice is one of the candidates
    for costs in ([1, 1, 1], [3, 2, 1], [1, 2, 3]):
        assert A.choose([7, 8, 9], costs, 8) in (7, 8, 9)


def test_cache_round_trip(tmp_path):
    p = tmp_path / "sub" / "tuning.json"
    assert A.load_cache(p) == {"version": A.VERSION, "entries": {}}
    A.store_entry(p, "k1", {"workers": 2, "bins_per_thread": 4})
    A.store_entry(p, "k2", {"levels": [[32, 64]]})
    d = A.load_cache(p)
    assert d["entries"]["k1"]["workers"] == 2 and d["entries"]["k2"]["levels"] == [[32, 64]]
    p.write_text("not json")
    assert A.load_cache(p)["entries"] == {}                 # a damaged file is ignored
    p.write_text(json.dumps({"version": 999, "entries": {"x": 1}}))
    assert A.load_cache(p)["entries"] == {}                 # another version is ignored


def test_cache_path_from_environment(monkeypatch, tmp_path):
    monkeypatch.setenv("PARTCONV_AMD_TUNE_CACHE", str(tmp_path / "t.json"))
    assert A.default_cache_path() == tmp_path / "t.json"
    monkeypatch.delenv("PARTCONV_AMD_TUNE_CACHE")
    assert A.default_cache_path().name == "tuning.json"
