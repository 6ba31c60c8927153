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
