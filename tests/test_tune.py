"""The tuned launch tables (haissinskijulia_amd/tune.py): lookups by bin
count, the library's own rule (0) for counts it does not list, and the
options' default taking the table's block."""
import json

import haissinskijulia_amd as H
from haissinskijulia_amd import tune


def test_table_lookup(tmp_path, monkeypatch):
    path = tmp_path / "tuning.json"
    path.write_text(json.dumps({"device": "gfx950", "fft_block": {"1024": 512, "16384": 2048}}))
    monkeypatch.setattr(tune, "TABLE", str(path))
    monkeypatch.setattr(tune, "_table", None)
    assert tune.fft_block_for(1024) == 512
    assert tune.fft_block_for(16384) == 2048
    assert tune.fft_block_for(4096) == 0


def test_missing_or_bad_table_means_library_rule(tmp_path, monkeypatch):
    monkeypatch.setattr(tune, "TABLE", str(tmp_path / "absent.json"))
    monkeypatch.setattr(tune, "_table", None)
    assert tune.fft_block_for(1024) == 0
    bad = tmp_path / "bad.json"
    bad.write_text("{not json")
    monkeypatch.setattr(tune, "TABLE", str(bad))
    monkeypatch.setattr(tune, "_table", None)
    assert tune.fft_block_for(1024) == 0


def test_options_take_the_tuned_block(tmp_path, monkeypatch):
    path = tmp_path / "tuning.json"
    path.write_text(json.dumps({"fft_block": {"2048": 256}}))
    monkeypatch.setattr(tune, "TABLE", str(path))
    monkeypatch.setattr(tune, "_table", None)
    keep = []
    o = H.TrackingOptions()._c(keep, 1000, 10, 2048)
    assert o.fft_block == 256
    assert H.TrackingOptions(fft_block=1024)._c(keep, 1000, 10, 2048).fft_block == 1024
    assert H.TrackingOptions()._c(keep, 1000, 10, 512).fft_block == 0


def test_tile_and_run_tables(tmp_path, monkeypatch):
    path = tmp_path / "tuning.json"
    path.write_text(json.dumps({"fft_block": {}, "tile_pairs": {"16": 2, "26": 4},
                                "run_pairs": {"4096": 2048}}))
    monkeypatch.setattr(tune, "TABLE", str(path))
    monkeypatch.setattr(tune, "_table", None)
    assert tune.size_class(2 * 65536) == 16 and tune.size_class(1) == 0
    assert tune.tile_pairs_for(2 * 65536) == 2          # class 16
    assert tune.tile_pairs_for(2 * 2 ** 18) == 2         # class 18: 16 is two classes away
    assert tune.tile_pairs_for(2 * 2 ** 25) == 4         # class 25: 26 is the nearest
    assert tune.tile_pairs_for(2 * 2 ** 21) == 0         # class 21: both five classes away
    assert tune.tile_pairs_for(2 * 2 ** 20) == 0         # class 20: nothing within two classes
    assert tune.tile_pairs_for(0) == 0
    assert tune.run_pairs_for(4096) == 2048 and tune.run_pairs_for(1024) == 0
    keep = []
    o = H.TrackingOptions()._c(keep, 2 * 65536, 10, 4096)
    assert (o.tile_pairs, o.run_pairs) == (2, 2048)
    o = H.TrackingOptions(tile_pairs=8, run_pairs=100)._c(keep, 2 * 65536, 10, 4096)
    assert (o.tile_pairs, o.run_pairs) == (8, 100)


def test_rule_tile_pairs():
    # about 1024 tiles of 1024 threads, at least one pair per thread
    assert tune.rule_tile_pairs(10_000) == 1
    assert tune.rule_tile_pairs(10_000_000) == 5      # 5e6 pairs: 4883 per tile
    assert tune.rule_tile_pairs(100_000_000) == 48    # 5e7 pairs: 48829 per tile
    assert tune.rule_tile_pairs(2 * 1024 * 1024 * 7) == 7
    assert tune.rule_tile_pairs(2 * 1024 * 1024 * 7 + 2) == 8


def test_tile_sweep_method(monkeypatch):
    """The tile sweep on a stand-in device (no GPU): one untimed call first,
    then the candidates interleaved within each repeat; the rule runs as its
    explicit value (0 in TrackingOptions would take the table's entry), and
    no candidate repeats the rule's tiling."""
    calls = []

    class FakeBunch:
        def __init__(self, n, dev):
            self.n = n

        def generate(self, *a):
            pass

        def evolve(self, params, nbins, opts):
            calls.append(opts.tile_pairs)
            return type("R", (), {"ms_total": 10.0 - (opts.tile_pairs == 4)})()

        def close(self):
            pass

    monkeypatch.setattr(H, "DeviceBunch", FakeBunch)
    monkeypatch.setattr(H.MI355X, "default", classmethod(lambda cls: None))
    monkeypatch.setattr(tune, "_table", {"fft_block": {}, "tile_pairs": {22: 64}, "run_pairs": {}})
    out = tune.tune_tiles([10_000_000], reps=2)
    rule = tune.rule_tile_pairs(10_000_000)                  # 5
    cands = [1, 2, 4, 8, 16, 32, 64, 128]
    assert 0 not in calls and rule not in cands
    assert calls == [rule] + 2 * ([rule] + cands)            # warm-up, then two interleaved repeats
    assert out == {22: {0: 10.0, **{c: 10.0 - (c == 4) for c in cands}}}
    # 4 pairs beat the rule by 10%: the merged table takes it over the old 64
    assert tune.merge_tables(tune._table, tiles=out)["tile_pairs"] == {"22": 4}


def test_committed_table_is_well_formed():
    with open(tune.TABLE) as f:
        table = json.load(f)
    for name in ("fft_block", "tile_pairs", "run_pairs"):
        for k, v in table.get(name, {}).items():
            assert int(k) >= 0 and int(v) > 0
    for k, v in table.get("fft_block", {}).items():
        assert int(v) & (int(v) - 1) == 0                # powers of two


def test_a_sweep_merges_into_the_table():
    base = {"fft_block": {256: 256, 1024: 64, 65536: 512},
            "tile_pairs": {17: 8, 19: 4, 22: 4, 26: 8, 28: 2}, "run_pairs": {4096: 8192}}
    # timings by size class, 0 = the rule: class 19 has a winner (2, 2% under
    # the rule); 22's entry is neutral (0.5% over the rule) and stays; 26's
    # entry is 5% slower than the rule and goes; 28's entry was not timed and
    # goes; 17's the sweep did not measure and it stays; 12 gets no entry
    tiles = {12: {0: 10.0, 2: 10.05, 4: 10.2},
             19: {0: 10.0, 2: 9.8, 4: 9.95},
             22: {0: 10.0, 4: 10.05, 8: 10.5},
             26: {0: 10.0, 8: 10.5, 64: 9.95},
             28: {0: 10.0, 64: 10.2}}
    out = tune.merge_tables(base, fft={256: 64, 2048: 128}, tiles=tiles)
    assert out["fft_block"] == {"256": 64, "1024": 64, "2048": 128, "65536": 512}
    assert out["tile_pairs"] == {"17": 8, "19": 2, "22": 4}
    assert out["run_pairs"] == {"4096": 8192}
    assert base["fft_block"][256] == 256   # the input is left as it was
    assert base["tile_pairs"][26] == 8
