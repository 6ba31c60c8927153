"""The GPU integration benchmark (haissinskijulia_amd/benchlog.py) end to end:
a small run, its log read back by the restated reference parser, and the
fresh-copy semantics of the samples (every sample starts from the same
initial bunch, so every sample ends in the same state)."""
import statistics

import numpy as np
import pytest

import haissinskijulia_amd as H
from haissinskijulia_amd import benchlog as BL
from haissinskijulia_amd import workloads as W
from test_benchlog import parse_log

pytestmark = pytest.mark.gpu


def test_benchmark_run_and_log(tmp_path):
    n, turns = 20_000, 30
    prm, times, r = BL.run(n, turns, samples=3)
    assert len(times) == 3 and all(t > 0 for t in times)
    med = statistics.median(times)
    text = BL.format_log(prm, med)
    (tmp_path / BL.log_filename(turns, n)).write_text(text)
    row = parse_log(text)
    assert row["particles"] == n and row["turns"] == turns
    assert row["time_sec"] == pytest.approx(med, rel=2e-3)
    assert np.isfinite(r.sigma_E) and np.isfinite(r.sigma_z) and r.turns_done == turns


def test_samples_start_from_the_same_bunch():
    # two evolve calls from fresh copies of one initial bunch are identical
    n, turns = 8_000, 12
    prm = BL.parameters(n, turns)
    lat = H.SimulationParameters.from_dict(W.reference_benchmark_lattice(turns))
    b = H.DeviceBunch(n)
    b.generate_particles(1, 0.0, 0.0, prm["σ_z0"], prm["σ_E0"])
    z0, e0 = (np.array(a, copy=True) for a in b.download())
    out = []
    for _ in range(2):
        b.upload(z0, e0)
        r = b.evolve(lat, prm["nbins_calculated"], H.TrackingOptions(seed=1))
        out.append((r.sigma_E, r.sigma_z, r.E0))
    b.close()
    assert out[0] == out[1]
