"""Offline tuner of the convolution block size (bins per FFT block).

``python -m haissinskijulia_amd.tune`` times the wake convolution of one
turn (``TrackingOptions(timing=True)``, HIP events around the convolution
launches) for every admissible block size at each bin count and writes the
fastest into ``tuning.json`` beside this file.  ``TrackingOptions`` with
``fft_block=0`` then takes the block from that table; bin counts it does not
list keep the library's own rule.  The table is read once and depends only on
the bin count, so a bunch gets the same block, and the same bits, alone, in a
batch and in a served mix.
"""
from __future__ import annotations

import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")

_table = None


def fft_block_for(nbins: int) -> int:
    """The tuned convolution block for `nbins`, 0 if the table has none."""
    global _table
    if _table is None:
        try:
            with open(TABLE) as f:
                _table = {int(k): int(v) for k, v in json.load(f).get("fft_block", {}).items()}
        except (OSError, ValueError):
            _table = {}
    return _table.get(int(nbins), 0)


def tune(nbins_list, n: int = 200_000, turns: int = 60, reps: int = 3) -> dict:
    import haissinskijulia_amd as H
    from haissinskijulia_amd import workloads as W
    dev = H.MI355X.default()
    maxb = 4096
    out = {}
    for nb in nbins_list:
        best = None
        # blocks of at least nbins/128: the blocked convolution's work grows
        # with the square of the block count
        for B in [64 << k for k in range(7)
                  if max(64, nb // 128) <= (64 << k) <= min(maxb, max(nb, 64))]:
            ms = []
            for r in range(reps):
                b = H.DeviceBunch(n, dev)
                b.generate(1 + r, W.SIGMA_Z, W.SIGMA_E)
                res = b.evolve(H.SimulationParameters.from_dict(W.ring_3gev_lattice(turns)), nb,
                               H.TrackingOptions(seed=2, resonators=[W.RESONATOR_BB],
                                                 charge=5e-9, bin_half_width_sigmas=8.0,
                                                 timing=True, fft_block=B))
                b.close()
                ms.append(res.ms_conv_avg)
            t = sorted(ms)[len(ms) // 2]
            print(f"nbins {nb:6d} block {B:5d}: convolution {t * 1e3:8.2f} us", flush=True)
            if best is None or t < best[1]:
                best = (B, t)
        out[nb] = best[0]
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--nbins", type=int, nargs="*",
                    default=[256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536, 131072])
    ap.add_argument("--write", action="store_true", help="write tuning.json")
    ap.add_argument("--out", default=TABLE, help="where --write puts the table")
    args = ap.parse_args()
    table = tune(args.nbins)
    print(json.dumps({"fft_block": table}))
    if args.write:
        import haissinskijulia_amd as H
        with open(args.out, "w") as f:
            json.dump({"device": H.MI355X.default().info()["arch"],
                       "fft_block": {str(k): v for k, v in sorted(table.items())}}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
