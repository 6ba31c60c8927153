"""Offline tuner of the library's launch knobs, written to ``tuning.json``.

Three tables, each timed on the device with HIP events:

- ``fft_block``: bins per FFT block of the wake convolution, by bin count
  (the convolution's own time per turn, ``TrackingOptions(timing=True)``);
- ``tile_pairs``: particle pairs per thread in each tile of the turn pass,
  by bunch size class (round(log2 of the pair count)), from the whole
  evolve call's device time (the tiling sets the pass's work per
  workgroup, the number of per-wave partial sums and the reductions over
  them);
- ``run_pairs``: the cost, in pairs, that the pass's workgroup balancing
  charges to each run of one bunch's tiles in a workgroup (staging and
  flushing its LDS), by bin count, from the batched 256-bunch scan's device
  time.

The threads per block of the turn pass are not a knob: its workgroup holds
the LDS histogram and the window of the potential and takes the whole LDS
of a CU, so it is one 1024-thread workgroup per CU.

``python -m haissinskijulia_amd.tune --write`` runs all three and writes
the table beside this file; ``TrackingOptions`` fields left at 0 take their
values from it, and keys it does not list keep the library's own rules.
Each value depends only on the bin count or the bunch size, so a bunch gets
the same tiling, and the same bits, alone, in a batch and in a served mix.
The run cost changes no result; the FFT block changes the potential by
rounding only (a different blocking of the same convolution), the tiling
the order of the per-tile sums.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")

_table = None


def _tables() -> dict:
    global _table
    if _table is None:
        try:
            with open(TABLE) as f:
                raw = json.load(f)
            _table = {name: {int(k): int(v) for k, v in raw.get(name, {}).items()}
                      for name in ("fft_block", "tile_pairs", "run_pairs")}
        except (OSError, ValueError, AttributeError):
            _table = {"fft_block": {}, "tile_pairs": {}, "run_pairs": {}}
    return _table


def fft_block_for(nbins: int) -> int:
    """The tuned convolution block for `nbins`, 0 if the table has none."""
    return _tables()["fft_block"].get(int(nbins), 0)


def size_class(n: int) -> int:
    """The bunch-size class of the tile table: round(log2(pairs))."""
    pairs = (int(n) + 1) // 2
    return int(round(math.log2(pairs))) if pairs > 0 else 0


def rule_tile_pairs(n: int, threads: int = 1024) -> int:
    """The pairs per thread per tile that the library's own rule gives a bunch
    of n particles: about 1024 tiles, at least one pair per thread."""
    pairs_per_tile = -(-((int(n) + 1) // 2) // 1024)       # ceil(pairs / 1024)
    return max(1, -(-pairs_per_tile // threads))            # ceil(... / threads)


def tile_pairs_for(n: int) -> int:
    """The tuned pairs per thread per tile for a bunch of n particles: the
    entry of the nearest tuned size class (at most two classes away), 0 if
    there is none."""
    tab = _tables()["tile_pairs"]
    if not tab or n <= 0:
        return 0
    c = size_class(n)
    best = min(tab, key=lambda k: (abs(k - c), k))
    return tab[best] if abs(best - c) <= 2 else 0


def run_pairs_for(nbins: int) -> int:
    """The tuned run cost for bunches of `nbins` bins, 0 if the table has none."""
    return _tables()["run_pairs"].get(int(nbins), 0)


def _median(xs):
    xs = sorted(xs)
    return xs[len(xs) // 2]


def tune_fft(nbins_list, n: int = 200_000, turns: int = 60, reps: int = 3) -> dict:
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
            t = _median(ms)
            print(f"nbins {nb:6d} block {B:5d}: convolution {t * 1e3:8.2f} us", flush=True)
            if best is None or t < best[1]:
                best = (B, t)
        out[nb] = best[0]
    return out


def tune_tiles(sizes, nbins: int = 4096, turns: int = 40, reps: int = 3) -> dict:
    """Pairs per thread per tile by bunch size class: the whole evolve call's
    device time (set-up, turns and epilogue) of one bunch for each candidate,
    the library's rule (0) among them and no candidate that tiles as the rule
    does.  The rule runs as its explicit value, since 0 in TrackingOptions
    would take the table's own entry.  Returns {size class: {candidate: ms}};
    merge_tables turns it into entries."""
    import haissinskijulia_amd as H
    from haissinskijulia_amd import workloads as W
    dev = H.MI355X.default()
    out = {}
    for n in sizes:
        n = int(n)
        pairs = (n + 1) // 2
        cands = [0] + [p for p in (1, 2, 4, 8, 16, 32, 64, 128)
                       if p * 1024 <= max(pairs, 1024) and p != rule_tile_pairs(n)]
        b = H.DeviceBunch(n, dev)

        def run(tp, seed):
            b.generate(seed, W.SIGMA_Z, W.SIGMA_E)
            return b.evolve(H.SimulationParameters.from_dict(W.ring_3gev_lattice(turns)), nbins,
                            H.TrackingOptions(seed=2, resonators=[W.RESONATOR_BB], charge=5e-9,
                                              bin_half_width_sigmas=8.0,
                                              tile_pairs=tp or rule_tile_pairs(n),
                                              fft_block=fft_block_for(nbins))).ms_total

        run(0, 1)   # untimed: the bunch's first call pays for first touches
        ms = {tp: [] for tp in cands}
        for r in range(reps):           # candidates interleaved within each repeat
            for tp in cands:
                ms[tp].append(run(tp, 1 + r))
        times = {tp: _median(v) for tp, v in ms.items()}
        for tp in cands:
            print(f"n {n:10d} tile_pairs {tp or 'rule':>4}: {times[tp] / turns * 1e3:9.2f} us/turn",
                  flush=True)
        b.close()
        out[size_class(n)] = times
    return out


def tune_runs(candidates=(256, 1280, 2048, 4096, 8192, 12288, 16384, 24576), turns: int = 40,
              reps: int = 3) -> dict:
    """Run cost for the scan's 4096-bin bunches: the batched 256-bunch scan's
    device time per candidate (the library's rule, 4 nbins = 16384, among
    them), the candidates interleaved in each repeat after an untimed call; the
    rule is kept unless a candidate beats it by more than 1%."""
    import haissinskijulia_amd as H
    from haissinskijulia_amd import workloads as W
    dev = H.MI355X.default()
    specs = W.scan_bunches()
    nbins = specs[0]["nbins"]
    rule = 4 * nbins
    candidates = sorted(set(candidates) | {rule})
    bunches = [H.DeviceBunch(s["n"], dev) for s in specs]
    params = [H.SimulationParameters.from_dict(W.ring_3gev_lattice(turns, voltage=s["V"]))
              for s in specs]

    def run(c):
        for s, b in zip(specs, bunches):
            b.generate(s["seed"], W.SIGMA_Z, W.SIGMA_E)
        opts = [H.TrackingOptions(seed=s["seed"] + 1, resonators=s["resonators"],
                                  charge=s["charge"], bin_half_width_sigmas=s["half_width"],
                                  run_pairs=c) for s in specs]
        return H.batch_evolve(bunches, [s["nbins"] for s in specs], params, opts)[0].ms_total

    run(candidates[0])   # untimed: the bunches' first call pays for first touches
    ms = {c: [] for c in candidates}
    for r in range(reps):              # candidates interleaved within each repeat
        for c in candidates:
            ms[c].append(run(c))
    times = {c: _median(v) for c, v in ms.items()}
    for c in candidates:
        print(f"scan, {nbins} bins, run_pairs {c:5d}: {times[c] / turns * 1e3:9.2f} us/turn",
              flush=True)
    for b in bunches:
        b.close()
    best = min(times, key=times.get)
    if times[best] >= 0.99 * times[rule]:     # within 1% of the rule: the rule
        best = rule
    return {nbins: best}


def merge_tables(table: dict, fft=None, tiles=None, runs=None) -> dict:
    """The table after sweeps: a sweep replaces the entries it measured and
    keeps the others.  For each size class the tile sweep timed (`tiles`, as
    tune_tiles returns it), a candidate that beats the rule by more than 1%
    becomes the entry; otherwise the class keeps its entry if the sweep timed
    that entry within 1% of the rule, and drops it if not.  A neutral entry
    stays because the sweep times one bunch alone, and the tiling also sets
    how a batch of many bunches spreads over the workgroups (on the 256-bunch
    scan, 4 pairs per thread for the bunches near 1e6 particles, neutral
    alone, takes 14% off the batch's time).  Keys as strings, in ascending
    order, ready for JSON."""
    out = {name: dict(t) for name, t in table.items()}
    if fft is not None:
        out["fft_block"].update(fft)
    for c, times in (tiles or {}).items():
        best = min(times, key=times.get)
        old = out["tile_pairs"].pop(c, None)
        if best != 0 and times[best] < 0.99 * times[0]:
            out["tile_pairs"][c] = best
        elif old is not None and old in times and times[old] <= 1.01 * times[0]:
            out["tile_pairs"][c] = old
    if runs is not None:
        out["run_pairs"].update(runs)
    return {name: {str(k): v for k, v in sorted(t.items())} for name, t in out.items()}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--what", nargs="*", default=["fft", "tiles", "runs"],
                    choices=["fft", "tiles", "runs"])
    ap.add_argument("--nbins", type=int, nargs="*",
                    default=[256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536, 131072])
    ap.add_argument("--sizes", type=float, nargs="*", default=[1e4, 1e5, 1e6, 1e7, 1e8],
                    help="bunch sizes of the tile sweep")
    ap.add_argument("--write", action="store_true", help="write the table")
    ap.add_argument("--out", default=TABLE, help="where --write puts the table")
    args = ap.parse_args()
    table = merge_tables(
        _tables(),
        fft=tune_fft(args.nbins) if "fft" in args.what else None,
        tiles=tune_tiles(args.sizes) if "tiles" in args.what else None,
        runs=tune_runs() if "runs" in args.what else None)
    print(json.dumps(table))
    if args.write:
        import haissinskijulia_amd as H
        with open(args.out, "w") as f:
            json.dump({"device": H.MI355X.default().info()["arch"], **table}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
