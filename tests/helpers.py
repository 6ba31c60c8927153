"""Shared helpers of the test-suite (lattices, tolerances, device gate)."""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from haissinskijulia_amd import workloads as W  # noqa: E402

# Parity tolerance of coordinates against the oracle: max |diff| over the
# bunch divided by the rms of the oracle's coordinate ("normwise relative").
# 1e-12 after 1 and 100 turns with identical injected normals.
COORD_RTOL = 1e-12
# wake potential against the oracle's FFT convolution (norm-wise)
POTENTIAL_RTOL = 1e-10


def normwise(a: np.ndarray, ref: np.ndarray) -> float:
    scale = float(np.sqrt(np.mean(ref * ref))) if ref.size else 1.0
    if scale == 0.0:
        scale = 1.0
    return float(np.max(np.abs(a - ref))) / scale if ref.size else 0.0


def params_dict(p: dict) -> dict:
    return dict(p)


def to_sim_params(p: dict):
    from haissinskijulia_amd import SimulationParameters
    return SimulationParameters.from_dict(p)


def bench_lattice(turns: int, flags=(True, True, True, True, True)) -> dict:
    return W.reference_benchmark_lattice(turns, flags)


def ring_lattice(turns: int, flags=(True, True, True, True, True), voltage: float = 3.2e6) -> dict:
    return W.ring_3gev_lattice(turns, voltage=voltage, flags=flags)
