#!/usr/bin/env python3
"""OWM ground-seed benchmark on one MI355X.

Headline workload (the metric's batched configuration): 512 independent tiles
with log-normal point counts (median 400k, clipped 20k-6M, 329M points in all),
densities 1-20 pts/m2, Wsize from {8, 12, 16}, Overlap 0.8, Bsize 20, fp64
coordinates, synthetic data (fastowm_amd.synth, seed 0), already resident in
HBM.  One step = the whole pipeline over the batch in one launch chain: tree,
stage 1, sort + stage 2, stage 3, merge, seeds copied to the host.

Prints ONE JSON line: batched Mpoints/s (higher is better; eager steps, with
HIP events around the dominant kernel family only), the roofline of that
family, the CPU baseline (the oracle on a bounded sample of the same batch,
all of this host's CPU share), and beside them: the same batch replayed from
a captured graph and run one tile per launch chain (plain), the single-tile
configurations (20M, 120M and 1M points; the 1M one also on the CPU), the
.xyz reader, the quadtree statistics and the serving loop.
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# kernel family behind each timed phase, and the algorithmic bytes per unit the
# library's phase timers count (the per-launch figure is that times the units)
PHASE_KERNEL = {"stage1": "owm::k_stage1", "leaf_aggs": "owm::k_leaf_aggs",
                "sort_tree_pass": "owm::k_sort_records", "morton": "owm::k_morton",
                "bbox": "owm::k_bbox", "agg_levels": "owm::k_agg_levels",
                "stage1_join": "owm::k_s1_bin+k_scan_excl"}
PHASE_BYTES = {"stage1": "24 B x N x Wsize/Displace + 4 B x windows",
               "leaf_aggs": "32 B x N + 48 B x leaves",
               "sort_tree_pass": "56 B x N for the first pass (x, y, z read; a 32-byte point written), 64 B x N for later passes (a 32-byte point read and written)",
               "morton": "16 B x N", "bbox": "16 B x N", "agg_levels": "48 B x (leaves + level entries)",
               "karras": "24 B x leaves + 56 B x nodes", "stage2_mark": "12 B x windows",
               "stage2_compact": "4 B x N / 32 (the LLP bitmap read)",
               "merge_seeds": "4 B x N / 32 + 8 B x stage-3 cells",
               "stage1_join": "72 B x leaves + 12 B x blocks"}
TRAFFIC_BATCH = os.path.join(ROOT, "profiles", "pmc_traffic_batch_512.json")


def measured_traffic(kernel, path):
    """HBM bytes per launch of `kernel` on this workload from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE passes (scaled by the copy-kernel factor),
    or None when no such measurement is committed."""
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    k = t.get("kernels", {}).get(kernel)
    if k is None:
        return None, None
    return int(k["bytes_per_launch"]), os.path.relpath(path, ROOT)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_cores():
    """The host cores this process may use: OMP_NUM_THREADS when set (the
    CPU share a one-GPU box gives a process), else the affinity set."""
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except Exception:
        aff = None
    return {"nproc": os.cpu_count(), "affinity": aff, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "cpu_model": model}


def oracle_tiles(jobs, workers):
    """The oracle on many tiles at once: one tile per worker thread (the
    oracle call releases the GIL), each single-threaded; seed arrays."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import concurrent.futures as cf
    import pyoracle
    saved = pyoracle.threads()
    pyoracle.set_threads(1)
    try:
        with cf.ThreadPoolExecutor(workers) as ex:
            return list(ex.map(lambda j: pyoracle.run(*j, 32)[0], jobs))
    finally:
        pyoracle.set_threads(saved)


def cpu_baseline_batch(tiles, arrs, gpu_seeds, sample_points, reps):
    """The oracle (a C restatement of the reference TBB algorithm: quadtree,
    exact stage 1, sort + stage 2, stage 3) on the first tiles of the same
    batch up to sample_points points, tiles in parallel on every core of this
    host's CPU share; median of reps runs.  Its seeds must equal the GPU's."""
    import numpy as np
    cores = cpu_cores()
    k, pts = 0, 0
    while k < len(tiles) and pts < sample_points:
        pts += tiles[k].n
        k += 1
    jobs = [(arrs[i][0], arrs[i][1], arrs[i][2], tiles[i].wsize, tiles[i].bsize, tiles[i].overlap)
            for i in range(k)]
    times, same = [], True
    for _ in range(reps):
        t0 = time.perf_counter()
        got = oracle_tiles(jobs, cores)
        times.append(time.perf_counter() - t0)
        same = same and all(np.array_equal(a, b) for a, b in zip(got, gpu_seeds[:k]))
    med = statistics.median(times)
    return {"value": round(pts / med / 1e6, 2), "unit": "Mpoints/s", "cores": cores, "kind": "port",
            "sample": f"first {k} tiles ({pts} points) of the same 512-tile batch, tiles in parallel "
                      f"(one oracle thread each) on {cores} cores, median of {reps} runs (tree + stages 1-3)",
            "equal_to_gpu": bool(same), "host": host_info()}


def cpu_single(x, y, z, want, reps, name):
    """The oracle on one whole tile with all of this host's CPU share (its
    OpenMP threads), median of reps runs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    cores = cpu_cores()
    pyoracle.set_threads(cores)
    times, same = [], True
    for _ in range(reps):
        t0 = time.perf_counter()
        seeds, _, _ = pyoracle.run(x, y, z, 12, 20, 0.8, 32)
        times.append(time.perf_counter() - t0)
        same = same and bool(np.array_equal(seeds, want))
    med = statistics.median(times)
    return {"mpts": round(len(x) / med / 1e6, 2), "cores": cores, "kind": "port",
            "sample": f"{name}, median of {reps} runs (tree + stages 1-3), {cores} OpenMP threads",
            "equal_to_gpu": same}


def cpu_window_split(x, y, z, gpu_phases, factors=(0.02, 0.05, 0.1), reps=3):
    """The CPU side of a column split of the stage-1 windows
    (octree_traverse_heter: nCols*factor columns to one device, the rest to
    the other): the oracle's stage 1 (the reference's exact per-window query)
    over the first factor*nCols columns with this host's cores, against the
    GPU's stage 1 of the whole tile.  A split cannot finish before
    max(GPU stage 1 x (1 - factor), CPU share); the tree's copy to the host and
    the merge of the two halves would only add."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    cores = cpu_cores()
    pyoracle.set_threads(cores)
    box = (float(x.min()), float(x.max()), float(y.min()), float(y.max()))
    p = pyoracle.params(12, 20, 0.8, box, len(x), 32)
    t0 = time.perf_counter()
    tree = pyoracle.Tree(x, y, z, p)
    t_tree = time.perf_counter() - t0
    gpu_ms = sum(gpu_phases[k]["ms_per_step"] for k in ("stage1", "stage1_join") if k in gpu_phases)
    rows = []
    for f in factors:
        cols = max(1, int(round(f * p.ncols)))
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            tree.stage1_cols(0, cols)
            times.append(time.perf_counter() - t0)
        cpu_ms = statistics.median(times) * 1e3
        bound = max(gpu_ms * (1.0 - cols / p.ncols), cpu_ms)
        rows.append({"factor": f, "cpu_columns": cols, "cpu_ms": round(cpu_ms, 3),
                     "split_ms_at_least": round(bound, 3),
                     "vs_gpu_alone": round(gpu_ms / bound, 4)})
    del tree
    return {"gpu_stage1_ms": round(gpu_ms, 4), "columns": int(p.ncols), "rows": int(p.nrows),
            "cpu_tree_s": round(t_tree, 2), "cores": cores, "kind": "port", "splits": rows,
            "note": "vs_gpu_alone < 1: the split is slower than the GPU alone"}


def phase_table(full):
    out = {}
    for name, p in full.items():
        if p["calls"] == 0:
            continue
        out[name] = {"ms_per_step": round(p["ms"], 4), "calls_per_step": p["calls"],
                     "gbs": round(p["bytes"] / (p["ms"] * 1e-3) / 1e9, 1) if p["ms"] > 0 else None}
    return out


def roofline_of(prof, dname, steps, n_points, traffic_file):
    """Achieved bandwidth of the dominant kernel family from the HIP events of
    the timed steps: its algorithmic bytes per launch / its mean launch time;
    for the tree sort also on the survey's per-point bytes (a key/index sort of
    24 B per point and pass plus a 52 B gather)."""
    dp = prof[dname]
    per_launch_bytes = dp["bytes"] / dp["calls"]
    per_launch_s = dp["ms"] * 1e-3 / dp["calls"]
    achieved = per_launch_bytes / per_launch_s / 1e9
    kernel = PHASE_KERNEL.get(dname, dname)
    traffic, src = measured_traffic(kernel, traffic_file)
    r = {"bound": "hbm", "kernel": kernel, "phase": dname, "achieved": round(achieved, 1),
         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
         "traffic": traffic, "traffic_source": src, "launch_ms": round(per_launch_s * 1e3, 4),
         "launches_per_step": dp["calls"] // steps,
         "algorithmic_bytes_per_launch": int(per_launch_bytes), "bytes_per_unit": PHASE_BYTES.get(dname)}
    if dname == "sort_tree_pass":
        passes = dp["calls"] // steps
        survey = (24 * passes + 52) * n_points
        step_s = dp["ms"] * 1e-3 / steps
        r["survey_bytes_per_step"] = int(survey)
        r["frac_survey_bytes"] = round(survey / step_s / 1e9 / HBM_PEAK_GBS, 4)
        r["survey_bytes_rule"] = ("24 B x N per pass (a key/index sort) + 52 B x N (gather), per step "
                                  "over the passes' summed launch time")
    return r


def run_headline(ctx, torch, synth, args):
    """The 512-tile batch: tuning (untimed), warmup, one fully phase-timed
    step, the timed eager steps (events on the dominant kernel family only),
    graph replays, and the same tiles one launch chain each (plain)."""
    import numpy as np
    from fastowm_amd import owm
    tiles = synth.batch_tiles(args.batch_tiles, seed=0)
    arrs = [t.generate() for t in tiles]
    n_total = sum(t.n for t in tiles)
    dev = [torch.from_numpy(np.concatenate([a[k] for a in arrs])).to("cuda:0") for k in range(3)]
    # the tile table is built once; results are views of the context's
    # page-locked seed buffer (copied out after the timed loops)
    desc = owm.batch_desc([(t.n, dict(wsize=t.wsize, bsize=t.bsize, overlap=t.overlap, leaf_size=0))
                           for t in tiles])
    torch.cuda.synchronize()
    tuning = {}
    if args.tune_table:
        with open(args.tune_table) as f:
            ctx.set_tune_table(json.load(f))
        t_tune, table = 0.0, ctx.tune_table()
        tuning["table"] = args.tune_table
    else:
        # the statistics model picks the table (the headline runs with it);
        # with --tuner both the timed-trial tuner runs too and each table's
        # batch time is measured, alternating, for the comparison
        log(f"autotune ({args.tuner})")
        t0 = time.perf_counter()
        table = ctx.autotune(*dev, desc, method="stats" if args.tuner != "trials" else "trials")
        t_tune = time.perf_counter() - t0
        tuning[("stats" if args.tuner != "trials" else "trials") + "_s"] = round(t_tune, 3)
        if args.tuner == "both":
            t0 = time.perf_counter()
            trial_table = ctx.autotune(*dev, desc, method="trials")
            tuning["trials_s"] = round(time.perf_counter() - t0, 3)
            best = {"stats": 1e300, "trials": 1e300}
            for _ in range(3):
                for name, tab in (("stats", table), ("trials", trial_table)):
                    ctx.set_tune_table(tab)
                    ctx.batch_run(*dev, desc, copy=False, split=False)
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    for _ in range(3):
                        ctx.batch_run(*dev, desc, copy=False, split=False)
                    torch.cuda.synchronize()
                    best[name] = min(best[name], (time.perf_counter() - t0) / 3 * 1e3)
            tuning["stats_table_ms"] = round(best["stats"], 3)
            tuning["trials_table_ms"] = round(best["trials"], 3)
            tuning["trials_table"] = [{k: e[k] for k in ("max_n", "leaf_size", "block_x", "block_y",
                                                         "block_k", "summaries", "sort_bits")}
                                      for e in trial_table]
            ctx.set_tune_table(table)
    for _ in range(args.warmup):
        got = ctx.batch_run(*dev, desc)
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.batch_run(*dev, desc)
    full = ctx.profile_read()
    ctx.profile(False)
    dname = max((k for k in full if full[k]["calls"]), key=lambda k: full[k]["ms"])
    ev_phases = sorted({dname, "sort_tree_pass"})
    ctx.profile(True, phases=ev_phases)
    log(f"timing {args.steps} batched steps (events on {', '.join(ev_phases)})")
    # a step ends with the seeds in host memory as the C ABI returns them (one
    # array, per-tile offsets); the per-tile arrays are made after the clock
    t0 = time.perf_counter()
    for _ in range(args.steps):
        flat, offs = ctx.batch_run(*dev, desc, copy=False, split=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    got = [a.copy() for a in np.split(flat, offs[1:-1].astype(np.intp))]
    prof = ctx.profile_read()
    ctx.profile(False)
    ms = (t1 - t0) / args.steps * 1e3
    roofline = roofline_of(prof, dname, args.steps, n_total, TRAFFIC_BATCH)
    roofline_sort = roofline_of(prof, "sort_tree_pass", args.steps, n_total, TRAFFIC_BATCH)
    if args.save_tune_table:
        with open(args.save_tune_table, "w") as f:
            json.dump(table, f, indent=1)
    res = {"tiles": len(tiles), "points": int(n_total), "ms": ms, "phases": phase_table(full),
           "roofline": roofline, "roofline_sort": roofline_sort,
           "batch": {"batched_eager_mpts": round(n_total / ms * 1e-3, 1), "batched_eager_ms": round(ms, 3),
                     "autotune_s": round(t_tune, 3), "tuning": tuning,
                     "tuned": [{k: e[k] for k in ("max_n", "leaf_size", "block_x", "block_y", "block_k",
                                                  "summaries", "sort_bits")} for e in table]}}
    if args.headline_only:
        return res, got, tiles, arrs
    log("graph replays")
    ctx.batch_run(*dev, desc, graph=True)  # capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        flat, offs = ctx.batch_run(*dev, desc, graph=True, copy=False, split=False)
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / args.steps
    gg = [a.copy() for a in np.split(flat, offs[1:-1].astype(np.intp))]
    log("plain: one launch chain per tile")
    offs = np.cumsum([0] + [t.n for t in tiles])
    views = [[d[offs[i]:offs[i + 1]] for d in dev] for i in range(len(tiles))]
    for i in range(min(4, len(tiles))):
        ctx.run_tile(*views[i], tiles[i].wsize, tiles[i].bsize, tiles[i].overlap)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain = [ctx.run_tile(*views[i], t.wsize, t.bsize, t.overlap) for i, t in enumerate(tiles)]
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    same = all(np.array_equal(a, b) for a, b in zip(plain, got)) and \
        all(np.array_equal(a, b) for a, b in zip(gg, got))
    log("served: the same tiles as device requests, all at once")
    served = run_served_batch(owm, ctx, tiles, views, got)
    served["vs_plain"] = round(served["served_mpts"] / (n_total / t_plain / 1e6), 2)
    served["vs_batched"] = round(served["served_mpts"] / (n_total / ms * 1e-3), 2)
    del dev, views
    torch.cuda.empty_cache()
    res["batch"].update({"batched_graph_mpts": round(n_total / t_graph / 1e6, 1),
                         "plain_mpts": round(n_total / t_plain / 1e6, 1),
                         "batched_graph_ms": round(t_graph * 1e3, 3), "plain_ms": round(t_plain * 1e3, 3),
                         "speedup_eager_vs_plain": round(t_plain * 1e3 / ms, 2),
                         "speedup_graph_vs_plain": round(t_plain / t_graph, 2),
                         "batched_equals_plain": same, "served": served})
    return res, got, tiles, arrs


def run_served_batch(owm, ctx, tiles, views, want, reps=5, slots=2, max_batch=64):
    """The batch's tiles submitted to the serving loop at once, one device
    request per tile (slices of the resident batch cloud: groups of consecutive
    slices are read in place), from a submitting thread while the caller
    waits for each result in order; every result must equal the batched
    chain's.  The first burst warms the slots and is not counted; p50/p99 are
    the tiles' submit-to-seeds latencies in the fastest burst (a burst queues
    every tile at once, so they span up to its wall time)."""
    import threading
    import numpy as np
    n_total = sum(t.n for t in tiles)
    srv = owm.Server(ctx, max_inflight=slots, max_batch=max_batch)
    walls, lats, same = [], [], True
    for rep in range(reps + 1):
        ids = [None] * len(tiles)
        res = [None] * len(tiles)
        lat = [0.0] * len(tiles)

        def submitter():
            for i, t in enumerate(tiles):
                ids[i] = srv.submit(*views[i], t.wsize, t.bsize, t.overlap)

        t0 = time.perf_counter()
        th = threading.Thread(target=submitter)
        th.start()
        for i in range(len(tiles)):
            while ids[i] is None:
                time.sleep(0.0002)
            res[i], lat[i] = srv.wait(ids[i])
        th.join()
        wall = time.perf_counter() - t0
        if rep:
            walls.append(wall)
            lats.append(lat)
        same = same and all(np.array_equal(a, b) for a, b in zip(res, want))
    srv.stop()
    best = min(walls)
    lb = np.array(lats[walls.index(best)]) * 1e3  # submit -> seeds ready, best burst
    return {"requests": len(tiles), "slots": slots, "max_batch": max_batch,
            "served_mpts": round(n_total / best / 1e6, 1),
            "served_median_mpts": round(n_total / statistics.median(walls) / 1e6, 1),
            "walls_ms": [round(w * 1e3, 2) for w in walls],
            "p50_ms": round(float(np.percentile(lb, 50)), 2),
            "p99_ms": round(float(np.percentile(lb, 99)), 2),
            "served_equals_batched": same}


def run_single(ctx, torch, dev, n, steps, name, cpu=None):
    """One resident tile: repeated calls (a repeat replays the captured chain),
    and eager steps with events on the dominant kernel family."""
    for _ in range(3):
        seeds = ctx.run_tile(*dev, 12, 20, 0.8, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        seeds = ctx.run_tile(*dev, 12, 20, 0.8, 0)
    torch.cuda.synchronize()
    t_rep = (time.perf_counter() - t0) / steps
    ctx.profile(True)
    ctx.run_tile(*dev, 12, 20, 0.8, 0)
    full = ctx.profile_read()
    ctx.profile(False)
    dname = max((k for k in full if full[k]["calls"]), key=lambda k: full[k]["ms"])
    ctx.profile(True, phases=[dname])
    t0 = time.perf_counter()
    for _ in range(steps):
        seeds = ctx.run_tile(*dev, 12, 20, 0.8, 0)
    torch.cuda.synchronize()
    t_eager = (time.perf_counter() - t0) / steps
    prof = ctx.profile_read()
    ctx.profile(False)
    out = {"workload": name, "points": n, "seeds": int(len(seeds)),
           "repeat_ms": round(t_rep * 1e3, 3), "repeat_mpts": round(n / t_rep / 1e6, 1),
           "eager_ms": round(t_eager * 1e3, 3), "eager_mpts": round(n / t_eager / 1e6, 1),
           "dominant": {"phase": dname, "launch_ms": round(prof[dname]["ms"] / prof[dname]["calls"], 4),
                        "gbs": round(prof[dname]["bytes"] / (prof[dname]["ms"] * 1e-3) / 1e9, 1)},
           "phases": phase_table(full)}
    return out, seeds


def run_reader_bench(ctx, synth, x, y, z, reps=3):
    """The .xyz reader on the C2 cloud written as LiDAR text (two decimals,
    no header): wall time per parse from host bytes (PCIe upload of the text
    included) and the device time of its kernels (HIP events)."""
    text = synth.format_xyz(x, y, z, 2)
    c = ctx.parse_xyz(text)  # allocations
    ctx.profile(True, phases=["xyz_parse"])
    t0 = time.perf_counter()
    for _ in range(reps):
        c = ctx.parse_xyz(text)
    wall = (time.perf_counter() - t0) / reps
    p = ctx.profile_read()["xyz_parse"]
    ctx.profile(False)
    dev_ms = p["ms"] / reps
    return {"bytes": len(text), "points": int(c.n), "host_lines": int(c.host_lines),
            "wall_ms": round(wall * 1e3, 2), "mpts_wall": round(c.n / wall / 1e6, 1),
            "device_ms": round(dev_ms, 3), "launches": p["calls"] // reps,
            "device_gbs_text": round(len(text) / (dev_ms * 1e-3) / 1e9, 1),
            "algorithmic_bytes": int(p["bytes"] / reps),
            "roofline_frac": round(p["bytes"] / reps / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "bytes_per_unit": "text read once + 24 B per point written"}, text, c


def reader_cpu(text, cloud, sample_bytes=100_000_000):
    """The reference reader's loop (the oracle's strtod restatement, one
    thread) on the first ~100 MB of the same text, and the device's values
    against it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    s = text[:sample_bytes]
    s = s[:s.rindex(b"\n") + 1]
    t0 = time.perf_counter()
    ox, oy, oz = pyoracle.parse_xyz(s, s.count(b"\n") + 1)
    dt = time.perf_counter() - t0
    gx, gy, gz = cloud.to_numpy()
    k = len(ox)
    same = all(np.array_equal(a[:k].view(np.uint64), b.view(np.uint64))
               for a, b in ((gx, ox), (gy, oy), (gz, oz)))
    return {"mpts": round(k / dt / 1e6, 2), "cores": 1, "kind": "port",
            "sample": f"first {len(s)} bytes ({k} points) of the same text, strtod per number",
            "bit_equal_to_device": bool(same)}


def run_qstats_bench(ctx, synth, dx, dy, x, y, reps=5):
    """Quadtree leaf statistics (the reference's statsmaxnum -s 32 -L 5 and
    stats -r 0.1 -L 5) of the resident C2 cloud: device time per call (HIP
    events over the path keys, their sort and the depth loop) and its
    algorithmic bytes, then the oracle on a bounded sample of the same cloud:
    its classes must equal the device's on that sample, and its time is the
    CPU baseline."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    box = synth.bbox(x, y)
    out = {}
    for name, pol, kw in (("statsmaxnum", 0, {"max_number": 32}), ("stats", 1, {"min_radius": 0.1})):
        ctx.qtree_stats(dx, dy, policy=pol, level=5, bbox=box, **kw)  # allocations
        ctx.profile(True, phases=["qtree_stats"])
        for _ in range(reps):
            cls, sm = ctx.qtree_stats(dx, dy, policy=pol, level=5, bbox=box, **kw)
        p = ctx.profile_read()["qtree_stats"]
        ctx.profile(False)
        ms = p["ms"] / reps
        # oracle on the first 2M points (a few seconds of one core)
        k = 2_000_000
        sub, _ = ctx.qtree_stats(dx[:k].contiguous(), dy[:k].contiguous(), policy=pol, level=5,
                                 bbox=box, **kw)
        t0 = time.perf_counter()
        want = pyoracle.qtree_stats(x[:k], y[:k], box, pol, level=5, **kw)
        dt = time.perf_counter() - t0
        out[name] = {"params": dict(kw, level=5), "points": len(x), "depth": sm["depth"],
                     "leaves": sm["leaves"], "bins": len(cls), "device_ms": round(ms, 3),
                     "mpts": round(len(x) / (ms * 1e-3) / 1e6, 1), "launches": p["calls"] // reps,
                     "algorithmic_bytes": int(p["bytes"] / reps),
                     "gbs": round(p["bytes"] / reps / (ms * 1e-3) / 1e9, 1),
                     "roofline_frac": round(p["bytes"] / reps / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "cpu_baseline": {"mpts": round(k / dt / 1e6, 2), "cores": 1, "kind": "port",
                                      "sample": f"first {k} points of the same cloud, the oracle's "
                                                "sequential parallel_qtree restatement"},
                     "equal_to_oracle_on_sample": sub == want}
    return out


def run_serve_bench(owm, synth, rates, n_rate, pool_size, slots, device=False):
    """Poisson request streams against the serving loop: N log-uniform
    10k-8M points, Wsize 5-20, Bsize 10-30, Overlap 0.5-0.9; one request in
    four re-queries a resident tile (same key, new parameters).  device: the
    pool's clouds are resident in HBM and submitted as device arrays (no
    PCIe transfer per request); a rate of 0 submits four times as many
    requests at once (the loop's saturated throughput, over a burst long
    enough that its start and end weigh little)."""
    import threading
    import numpy as np
    pool = []
    for _, t in synth.serve_requests(pool_size, rate=1.0, seed=11):
        arrs = t.generate()
        if device:
            import torch
            arrs = tuple(torch.from_numpy(a).to("cuda:0") for a in arrs)
        pool.append((t, arrs))
    if device:
        import torch
        torch.cuda.synchronize()
    out = []
    with owm.Context(0) as ctx:
        srv = owm.Server(ctx, max_inflight=slots, max_batch=32)
        # warm the slots once (allocation of arenas)
        for t, (x, y, z) in pool[:slots * 2]:
            srv.wait(srv.submit(x, y, z, t.wsize, t.bsize, t.overlap))
        for rate in rates:
            n_req = n_rate if rate > 0 else 4 * n_rate
            reqs = synth.serve_requests(n_req, rate=rate if rate > 0 else 1e9, seed=int(rate) or 1)
            ids = [None] * n_req
            npts = [0] * n_req
            t0 = time.perf_counter()

            def submitter():
                for i, (t_arr, tr) in enumerate(reqs):
                    now = time.perf_counter() - t0
                    if t_arr > now:
                        time.sleep(t_arr - now)
                    k = i % pool_size
                    t, (x, y, z) = pool[k]
                    key = k + 1 if i % 4 == 3 else 0
                    ids[i] = srv.submit(x, y, z, tr.wsize, tr.bsize, tr.overlap, key=key)
                    npts[i] = t.n

            th = threading.Thread(target=submitter)
            th.start()
            lats = []
            for i in range(n_req):
                while ids[i] is None:
                    time.sleep(0.0002)
                _, lat = srv.wait(ids[i])
                lats.append(lat)
            th.join()
            wall = time.perf_counter() - t0
            lats = np.array(lats) * 1e3
            out.append({"inputs": "device" if device else "host",
                        "rate_tiles_s": rate if rate > 0 else "burst", "requests": n_req,
                        "p50_ms": round(float(np.percentile(lats, 50)), 2),
                        "p99_ms": round(float(np.percentile(lats, 99)), 2),
                        "mpts": round(sum(npts) / wall / 1e6, 1),
                        "offered_mpts": round(sum(npts) / reqs[-1][0] / 1e6, 1) if rate > 0 else None})
            log("serve", out[-1])
        srv.stop()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch-tiles", type=int, default=512)
    ap.add_argument("--tune-table", default=None,
                    help="load this tuning table (JSON, as owm_tune_get) instead of autotuning")
    ap.add_argument("--save-tune-table", default=None, help="write the tuning table used (JSON)")
    ap.add_argument("--tuner", choices=("stats", "trials", "both"), default="both",
                    help="headline table from the statistics model (stats) or timed trials; "
                         "both: the model's table, with the trials' table timed beside it")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the batched eager steps (no graph, plain or other legs: for profiling them)")
    ap.add_argument("--single-only", choices=("c1", "c2", "c3"), default=None,
                    help="only this single-tile leg (for profiling it); prints its object")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU legs")
    ap.add_argument("--cpu-sample", type=int, default=110_000_000,
                    help="points of the batch the CPU baseline runs on")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-single", action="store_true", help="skip the 20M / 120M / 1M tiles")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-reader", action="store_true", help="skip the .xyz reader leg")
    ap.add_argument("--no-qstats", action="store_true", help="skip the quadtree statistics leg")
    ap.add_argument("--serve-rates", default="50,500,2000")
    ap.add_argument("--serve-device-rates", default="2000,8000,0",
                    help="rates (tiles/s) of the device-input serving runs; 0 = all at once")
    ap.add_argument("--serve-requests", type=int, default=160)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times "
                         "(the known-size kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe << 20
        a = torch.ones(n // 4, dtype=torch.int32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the batch and serving paths "
                                   "shard tiles across GPUs by design but no multi-GPU run is "
                                   "measured here", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from fastowm_amd import owm, synth

    ctx = owm.Context(0)
    if args.single_only:
        cfg = {"c1": synth.config_c1, "c2": synth.config_c2, "c3": synth.config_c3}[args.single_only](0)
        if args.tune_table:
            with open(args.tune_table) as f:
                ctx.set_tune_table(json.load(f))
        x, y, z = cfg.generate()
        dev = [torch.from_numpy(a).to("cuda:0") for a in (x, y, z)]
        torch.cuda.synchronize()
        out, _ = run_single(ctx, torch, dev, cfg.n, args.steps, cfg.name)
        ctx.close()
        print(json.dumps({"single": out}), flush=True)
        return
    log("generating the 512-tile batch")
    head, batch_seeds, tiles, arrs = run_headline(ctx, torch, synth, args)
    ms = head["ms"]
    line = {"metric": "Mpoints/s end-to-end OWM (batched vs plain GPU vs TBB CPU); p99 tile latency",
            "value": round(head["points"] / ms * 1e-3, 1), "unit": "Mpoints/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "batched: 512 independent tiles, log-normal point counts (median "
                                   "400k, clipped 20k-6M), densities 1-20 pts/m2, Wsize {8,12,16}, "
                                   "Overlap 0.8, Bsize 20, one launch chain, resident in HBM",
                       "tiles": head["tiles"], "points": head["points"],
                       "step": "the whole batch, eager launch chain (the graph replay is beside it)"},
            "roofline": head["roofline"], "roofline_sort": head["roofline_sort"],
            "phases": head["phases"],
            "phases_source": "one untimed step with every phase timed by HIP events; the timed "
                             "steps carry events around the roofline kernel family and the "
                             "tree sort only",
            "batch": head["batch"]}
    if not args.headline_only:
        if not args.no_cpu:
            log("cpu baseline (oracle on a sample of the batch)")
            line["cpu_baseline"] = cpu_baseline_batch(tiles, arrs, batch_seeds, args.cpu_sample,
                                                      args.cpu_reps)
        del arrs, batch_seeds
        single = {}
        x = y = z = dx = dy = None
        if not args.no_single:
            for key, cfg, steps in (("c1", synth.config_c1(0), args.steps),
                                    ("c3", synth.config_c3(0), max(3, args.steps // 4)),
                                    ("c2", synth.config_c2(0), args.steps)):
                log("single tile", key)
                x, y, z = cfg.generate()
                dev = [torch.from_numpy(a).to("cuda:0") for a in (x, y, z)]
                torch.cuda.synchronize()
                single[key], seeds = run_single(ctx, torch, dev, cfg.n, steps, cfg.name)
                if key == "c1" and not args.no_cpu:
                    single[key]["cpu"] = cpu_single(x, y, z, seeds, 5, "the same 1M-point tile")
                    single[key]["cpu"]["gpu_over_cpu"] = round(single[key]["repeat_mpts"] /
                                                              single[key]["cpu"]["mpts"], 1)
                if key == "c2" and not args.no_cpu:
                    log("window split: CPU share of the stage-1 columns")
                    single[key]["window_split"] = cpu_window_split(x, y, z, single[key]["phases"])
                if key != "c2":
                    del dev
                else:
                    dx, dy = dev[0], dev[1]
            line["single"] = single
        if not args.no_reader and x is not None:
            log("reader leg")
            line["reader"], reader_text, reader_cloud = run_reader_bench(ctx, synth, x, y, z)
            if not args.no_cpu:
                line["reader"]["cpu_baseline"] = reader_cpu(reader_text, reader_cloud)
        if not args.no_qstats and dx is not None:
            log("quadtree statistics leg")
            line["qtree_stats"] = run_qstats_bench(ctx, synth, dx, dy, x, y)
        if not args.no_serve:
            log("serve leg")
            rates = [float(r) for r in args.serve_rates.split(",") if r]
            line["serve"] = run_serve_bench(owm, synth, rates, args.serve_requests, 48, 3)
            drates = [float(r) for r in args.serve_device_rates.split(",") if r]
            line["serve"] += run_serve_bench(owm, synth, drates, args.serve_requests, 48, 3, device=True)
            best = max(r["mpts"] for r in line["serve"])
            sb = head["batch"]["served"]
            line["serve_summary"] = {"best_served_mpts": best,
                                     "vs_plain": round(best / head["batch"]["plain_mpts"], 2),
                                     "vs_batched": round(best / line["value"], 2),
                                     "served_batch_mpts": sb["served_mpts"],
                                     "served_batch_vs_plain": sb["vs_plain"],
                                     "host_input_bound": "24 B per point over PCIe"}
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
