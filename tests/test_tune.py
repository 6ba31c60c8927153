"""The convolution-block table (haissinskijulia_amd/tune.py): lookups by bin
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


def test_committed_table_is_well_formed():
    with open(tune.TABLE) as f:
        table = json.load(f)
    for name in ("fft_block", "tile_pairs", "run_pairs"):
        for k, v in table.get(name, {}).items():
            assert int(k) >= 0 and int(v) > 0
    for k, v in table.get("fft_block", {}).items():
        assert int(v) & (int(v) - 1) == 0                # powers of two


def test_a_sweep_merges_into_the_table():
    base = {"fft_block": {256: 256, 1024: 64, 65536: 512}, "tile_pairs": {19: 4, 22: 4, 26: 8},
            "run_pairs": {4096: 8192}}
    # an FFT sweep of two bin counts keeps the third; a tile sweep of 1e6 and
    # 1e8 (classes 19 and 26) drops 26 where the rule won and keeps 22
    out = tune.merge_tables(base, fft={256: 64, 2048: 128}, tiles={19: 2}, tile_sizes=[1e6, 1e8])
    assert out["fft_block"] == {"256": 64, "1024": 64, "2048": 128, "65536": 512}
    assert out["tile_pairs"] == {"19": 2, "22": 4}
    assert out["run_pairs"] == {"4096": 8192}
    assert base["fft_block"][256] == 256   # the input is left as it was
