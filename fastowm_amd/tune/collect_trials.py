"""Training data of the statistics-driven tuner (run on the GPU).

For every homogeneous tile set below, times the batch chain under each of the
48 configurations the tuning table can hold (leaf size x stage-1 window block
x sort key width; best of `--reps` runs after one untimed run, host wall time
of owm_batch_run_device including the seed copy) and records the leaf
histogram of one sample tile of the set, computed with the settings the tuner
uses (owm_qtree_stats, maxNumber policy, OWM_TUNE_STATS_MAXNUMBER points,
level OWM_TUNE_STATS_LEVEL), with the tile's derived window grid.

    python -m fastowm_amd.tune.collect_trials out.json [--quick]

fit_model.py fits the tuner's cost model to the file this writes.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from fastowm_amd import owm, synth

LEAVES = (16, 32, 64, 128)
BLOCKS = ((8, 8, 4, 0), (16, 8, 4, 0), (8, 8, 2, 0), (16, 8, 4, 1), (16, 16, 4, 0), (32, 16, 4, 0))
SORTS = (32, 24, 16, 8)
SET_POINTS = 8 << 20  # points per homogeneous set of small tiles


def tile_sets():
    """(n per tile, density pts/m2, Wsize, Overlap, cluster fraction) of each set."""
    s = []
    for n in (25_000, 120_000, 500_000, 2_000_000, 6_000_000):
        for dens in (1.5, 6.0, 18.0):
            for ws in (8, 12, 16):
                s.append((n, dens, ws, 0.8, 0.0))
    for n in (120_000, 2_000_000):
        for dens in (3.0, 12.0):
            s.append((n, dens, 12, 0.5, 0.0))
            s.append((n, dens, 12, 0.8, 0.3))
    for ws, ov in ((5, 0.8), (20, 0.8), (12, 0.6), (12, 0.9), (8, 0.5), (16, 0.7)):
        s.append((500_000, 6.0, ws, ov, 0.0))
    for dens in (2.0, 6.0, 18.0):
        for ws in (8, 16):
            s.append((20_000_000, dens, ws, 0.8, 0.0))
    s.append((60_000_000, 3.0, 12, 0.8, 0.0))
    s.append((60_000_000, 12.0, 12, 0.8, 0.0))
    return s


def set_table(ctx, base, leaf, block, sort_bits):
    ents = [dict(e) for e in base]
    for e in ents:
        e["leaf_size"], e["sort_bits"] = leaf, sort_bits
        e["block_x"], e["block_y"], e["block_k"], e["summaries"] = block
    ctx.set_tune_table(ents)


def main():
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--quick", action="store_true", help="every seventh set only")
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    sets = tile_sets()[::7] if args.quick else tile_sets()
    res = []
    with owm.Context(0) as ctx:
        base = ctx.tune_table()
        for si, (n, dens, ws, ov, cf) in enumerate(sets):
            k = max(1, SET_POINTS // n)
            side = math.sqrt(n / dens)
            tiles = [synth.Tile(seed=7, stream=50_000 + 100 * si + t, n=n, width=side, height=side,
                                x0=400_000.0 + 10_000.0 * t, y0=4_600_000.0, cluster_frac=cf,
                                n_clusters=max(1, n // 2000) if cf else 0, cluster_sigma=3.0,
                                wsize=ws, bsize=20, overlap=ov) for t in range(k)]
            arrs = [t.generate() for t in tiles]
            x, y, z = (torch.from_numpy(np.concatenate([a[i] for a in arrs])).to("cuda:0")
                       for i in range(3))
            del arrs
            desc = owm.batch_desc([(n, dict(wsize=ws, bsize=20, overlap=ov, leaf_size=0))] * k)
            times = {}
            for leaf in LEAVES:
                for block in BLOCKS:
                    for sb in SORTS:
                        set_table(ctx, base, leaf, block, sb)
                        best = 1e300
                        for r in range(args.reps + 1):
                            torch.cuda.synchronize()
                            t0 = time.perf_counter()
                            ctx.batch_run(x, y, z, desc, copy=False)
                            dt = (time.perf_counter() - t0) * 1e3
                            if r:
                                best = min(best, dt)
                        times["%d,%d,%d,%d,%d,%d" % ((leaf,) + block + (sb,))] = round(best, 4)
            t0 = time.perf_counter()
            bins, sm = ctx.qtree_stats(x[:n], y[:n], policy=owm.QT_MAXNUMBER,
                                       max_number=owm.TUNE_STATS_MAXNUMBER,
                                       level=owm.TUNE_STATS_LEVEL)
            stats_ms = (time.perf_counter() - t0) * 1e3
            d = owm.derive(ws, 20, ov, n, sm["bbox"])
            res.append(dict(n=n, density=dens, wsize=ws, overlap=ov, cluster_frac=cf, tiles=k,
                            sample=dict(bins=bins, bbox=list(sm["bbox"]), max_radius=sm["max_radius"],
                                        depth=sm["depth"], stats_ms=round(stats_ms, 3),
                                        displace=d.displace, ncols=d.ncols, nrows=d.nrows),
                            times=times))
            bk = min(times, key=times.get)
            print(f"[set {si + 1}/{len(sets)}] n={n} density={dens} W={ws} O={ov} clusters={cf} "
                  f"tiles={k}: best {bk} {times[bk]:.3f} ms, stats {stats_ms:.2f} ms", flush=True)
            del x, y, z
            with open(args.out, "w") as f:
                json.dump(res, f)
    print("done", flush=True)


if __name__ == "__main__":
    sys.exit(main())
