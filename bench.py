#!/usr/bin/env python3
"""OWM ground-seed benchmark on one MI355X.

Default workload: the 20M-point 2x2 km tile (5 pts/m2, 10% of points in
Gaussian clusters), Wsize 12, Overlap 0.8, Bsize 20, leaf 32, fp64 coordinates,
synthetic data (fastowm_amd.synth, seed 0).  One step = the whole pipeline on
that tile with the cloud already resident in HBM: spatial tree, stage 1,
sort + stage 2, stage 3, merge, seeds copied back to the host.

Prints ONE JSON line: Mpoints/s (higher is better), the roofline of the
dominant kernel family (HIP-event time on the library's stream, algorithmic
bytes), the CPU baseline (the oracle timed on this host's cores on the same
tile) and, unless --no-batch, the 512-tile uneven batch (plain one-tile-per-
launch-chain vs batched) on the same device.
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
               "leaf_aggs": "28 B x N + 48 B x leaves",
               "sort_tree_pass": "60 B x N per pass with 16-bit keys (first pass: key + x, y, z read, key + x, y, z, index written; last: key + x, y, z, index read, x, y, z, index written)",
               "morton": "20 B x N", "bbox": "16 B x N", "agg_levels": "48 B x (leaves + level entries)",
               "karras": "24 B x leaves + 76 B x nodes", "sort_windows_pass": "12 B x windows per pass",
               "stage2_compact": "8 B x windows", "stage1_join": "72 B x leaves + 12 B x blocks"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_c2.json")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE passes (scaled by the copy-kernel factor),
    or None when no such measurement is committed."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    k = t.get("kernels", {}).get(kernel)
    if k is None:
        return None, None
    return int(k["bytes_per_launch"]), os.path.relpath(TRAFFIC_FILE, ROOT)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(x, y, z, want_seeds, reps):
    """The oracle (a C restatement of the reference TBB algorithm, OpenMP) on
    the same tile; median of `reps` runs of tree + stages 1-3."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    cores = cpu_cores()
    pyoracle.set_threads(cores)
    times, parity = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        seeds, _, _ = pyoracle.run(x, y, z, 12, 20, 0.8, 32)
        times.append(time.perf_counter() - t0)
        parity = bool(np.array_equal(seeds, want_seeds))
    med = statistics.median(times)
    return {"value": round(len(x) / med / 1e6, 3), "unit": "Mpoints/s", "cores": cores,
            "kind": "port",
            "sample": f"full 20M-point tile, median of {reps} runs (tree + stages 1-3), "
                      f"{cores} OpenMP threads"}, parity


def run_batch_bench(ctx, torch, synth, n_tiles, steps):
    """Autotune on the batch, then plain (one run_tile launch chain per tile)
    vs batched (all tiles in one chain, eager and replayed from a hipGraph)."""
    import numpy as np
    tiles = synth.batch_tiles(n_tiles, seed=0)
    arrs = [t.generate() for t in tiles]
    n_total = sum(t.n for t in tiles)
    dev = [torch.from_numpy(np.concatenate([a[k] for a in arrs])).to("cuda:0") for k in range(3)]
    desc = [(t.n, dict(wsize=t.wsize, bsize=t.bsize, overlap=t.overlap, leaf_size=0)) for t in tiles]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    table = ctx.autotune(*dev, desc)
    t_tune = time.perf_counter() - t0
    # plain: one run_tile per tile on device-resident slices
    offs = np.cumsum([0] + [t.n for t in tiles])
    views = [[d[offs[i]:offs[i + 1]] for d in dev] for i in range(len(tiles))]
    for i in range(min(4, len(tiles))):
        ctx.run_tile(*views[i], tiles[i].wsize, tiles[i].bsize, tiles[i].overlap)
    t0 = time.perf_counter()
    plain = []
    for i, t in enumerate(tiles):
        plain.append(ctx.run_tile(*views[i], t.wsize, t.bsize, t.overlap))
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    ctx.batch_run(*dev, desc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eager = ctx.batch_run(*dev, desc)
    torch.cuda.synchronize()
    t_eager = (time.perf_counter() - t0) / steps
    ctx.batch_run(*dev, desc, graph=True)  # capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        got = ctx.batch_run(*dev, desc, graph=True)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / steps
    same = all(np.array_equal(a, b) for a, b in zip(plain, got)) and \
        all(np.array_equal(a, b) for a, b in zip(plain, eager))
    return {"tiles": len(tiles), "points": int(n_total),
            "plain_mpts": round(n_total / t_plain / 1e6, 1),
            "batched_mpts": round(n_total / t_batch / 1e6, 1),
            "batched_eager_mpts": round(n_total / t_eager / 1e6, 1),
            "batched_ms": round(t_batch * 1e3, 3), "batched_eager_ms": round(t_eager * 1e3, 3),
            "plain_ms": round(t_plain * 1e3, 3),
            "speedup": round(t_plain / t_batch, 2), "batched_equals_plain": same,
            "autotune_s": round(t_tune, 2),
            "tuned": [{k: e[k] for k in ("max_n", "leaf_size", "block_x", "block_y", "block_k", "summaries",
                                         "sort_bits")} for e in table if e["ms"] > 0]}


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


def run_serve_bench(owm, synth, rates, n_req, pool_size, slots):
    """Poisson request streams against the serving loop: N log-uniform
    10k-8M points, Wsize 5-20, Bsize 10-30, Overlap 0.5-0.9; one request in
    four re-queries a resident tile (same key, new parameters)."""
    import threading
    import numpy as np
    pool = []
    for _, t in synth.serve_requests(pool_size, rate=1.0, seed=11):
        pool.append((t, t.generate()))
    out = []
    with owm.Context(0) as ctx:
        srv = owm.Server(ctx, max_inflight=slots, max_batch=32)
        # warm the slots once (allocation of arenas)
        for t, (x, y, z) in pool[:slots * 2]:
            srv.wait(srv.submit(x, y, z, t.wsize, t.bsize, t.overlap))
        for rate in rates:
            reqs = synth.serve_requests(n_req, rate=rate, seed=int(rate))
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
            out.append({"rate_tiles_s": rate, "requests": n_req,
                        "p50_ms": round(float(np.percentile(lats, 50)), 2),
                        "p99_ms": round(float(np.percentile(lats, 99)), 2),
                        "mpts": round(sum(npts) / wall / 1e6, 1),
                        "offered_mpts": round(sum(npts) / reqs[-1][0] / 1e6, 1)})
            log("serve", out[-1])
        srv.stop()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-tile batch leg")
    ap.add_argument("--batch-tiles", type=int, default=512)
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-reader", action="store_true", help="skip the .xyz reader leg")
    ap.add_argument("--no-qstats", action="store_true", help="skip the quadtree statistics leg")
    ap.add_argument("--serve-rates", default="50,500,2000")
    ap.add_argument("--serve-requests", type=int, default=160)
    ap.add_argument("--batch-only", action="store_true",
                    help="only the 512-tile batch leg (for profiling it)")
    ap.add_argument("--leaf", type=int, default=0,
                    help="leaf size (maxNumber) of the timed tile; 0 = autotuned")
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

    import numpy as np
    import torch
    from fastowm_amd import owm, synth

    if args.batch_only:
        with owm.Context(0) as bctx:
            print(json.dumps({"batch": run_batch_bench(bctx, torch, synth, args.batch_tiles,
                                                       max(1, args.steps // 4))}), flush=True)
        return

    tile = synth.config_c2(seed=0)
    log("generating", tile.name)
    x, y, z = tile.generate()
    dx, dy, dz = (torch.from_numpy(a).to("cuda:0") for a in (x, y, z))
    torch.cuda.synchronize()
    ctx = owm.Context(0)
    leaf = args.leaf
    tuned = None
    if leaf == 0:
        # leaf size, window block and sort width from the autotuner (untimed)
        log("autotune")
        tuned = ctx.autotune(dx, dy, dz, [(tile.n, dict(wsize=12, bsize=20, overlap=0.8))])
        tuned = [e for e in tuned if e["ms"] > 0][-1]
    for _ in range(args.warmup):
        seeds = ctx.run_tile(dx, dy, dz, 12, 20, 0.8, leaf)
    torch.cuda.synchronize()
    # one untimed step with every phase timed (two events per phase cost the
    # stream a few microseconds each): the phase breakdown and the dominant
    # kernel family
    ctx.profile(True)
    seeds = ctx.run_tile(dx, dy, dz, 12, 20, 0.8, leaf)
    full = ctx.profile_read()
    ctx.profile(False)
    phases = {}
    for name, p in full.items():
        if p["calls"] == 0:
            continue
        phases[name] = {"ms_per_step": round(p["ms"], 4), "calls_per_step": p["calls"],
                        "gbs": round(p["bytes"] / (p["ms"] * 1e-3) / 1e9, 1) if p["ms"] > 0 else None}
    dname = max((k for k in full if full[k]["calls"]), key=lambda k: full[k]["ms"])
    # the timed steps record events around the dominant kernel family only
    ctx.profile(True, phases=[dname])
    log(f"timing {args.steps} steps (events on {dname})")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        seeds = ctx.run_tile(dx, dy, dz, 12, 20, 0.8, leaf)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = ctx.profile_read()
    ctx.profile(False)
    ms = (t1 - t0) / args.steps * 1e3
    value = tile.n / (ms / 1e3) / 1e6
    dp = prof[dname]
    per_launch_bytes = dp["bytes"] / dp["calls"]
    per_launch_s = dp["ms"] * 1e-3 / dp["calls"]
    achieved = per_launch_bytes / per_launch_s / 1e9
    kernel = PHASE_KERNEL.get(dname, dname)
    if dname == "stage1" and tuned:
        kernel = "owm::k_stage1<%d, %d, %d, %s>" % (tuned["block_x"], tuned["block_y"],
                                                    tuned["block_k"],
                                                    "true" if tuned["summaries"] else "false")
    traffic, traffic_src = measured_traffic(kernel)
    roofline = {"bound": "hbm", "kernel": kernel, "phase": dname, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "traffic_source": traffic_src,
                "launch_ms": round(per_launch_s * 1e3, 4),
                "algorithmic_bytes_per_launch": int(per_launch_bytes),
                "bytes_per_unit": PHASE_BYTES.get(dname)}

    line = {"metric": "Mpoints/s end-to-end OWM (tree + stages 1-3), plain GPU, 20M-point tile",
            "value": round(value, 1), "unit": "Mpoints/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "c2: 20M points, 2x2 km, 5 pts/m2, 10% in Gaussian clusters "
                                   "(sigma 3 m), Wsize 12, Overlap 0.8, Bsize 20",
                       "points": tile.n, "seeds": int(len(seeds)),
                       "leaf_size": leaf if leaf else (tuned["leaf_size"] if tuned else None),
                       "tuned": {k: tuned[k] for k in ("leaf_size", "block_x", "block_y", "block_k", "summaries",
                                                       "sort_bits")} if tuned else None},
            "roofline": roofline, "phases": phases,
            "phases_source": "one untimed step with every phase timed by HIP events; the timed "
                             "steps carry events around the roofline kernel only"}
    if not args.no_batch:
        log("batch leg")
        line["batch"] = run_batch_bench(ctx, torch, synth, args.batch_tiles, max(1, args.steps // 4))
    if not args.no_reader:
        log("reader leg")
        line["reader"], reader_text, reader_cloud = run_reader_bench(ctx, synth, x, y, z)
    if not args.no_qstats:
        log("quadtree statistics leg")
        line["qtree_stats"] = run_qstats_bench(ctx, synth, dx, dy, x, y)
    if not args.no_serve:
        log("serve leg")
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        line["serve"] = run_serve_bench(owm, synth, rates, args.serve_requests, 48, 3)
    if not args.no_cpu:
        log("cpu baseline leg")
        cb, parity = cpu_baseline(x, y, z, seeds, args.cpu_reps)
        line["cpu_baseline"] = cb
        line["parity_vs_oracle"] = parity
        if not args.no_reader:
            line["reader"]["cpu_baseline"] = reader_cpu(reader_text, reader_cloud)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
