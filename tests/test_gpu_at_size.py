"""The batched and served configurations at their full sizes against the
oracle, tile by tile (bit-exact seed sets).

- segmented record sort with 16384-point sort tiles: a batch of more than
  2^21 points whose tiles average at least 64K points sorts every tile as its
  own segment of 16384-point sort tiles (the path the 512-tile batch takes);
- the 512-tile uneven batch itself (log-normal 20k-6M points per tile, 329M in
  all), graph-replayed twice;
- a served stream of 10k-8M-point requests with mixed parameters, admitted
  while others run, plus keyed re-queries of a resident tree.

The oracle (oracle/owm_oracle.c) restates CPP-TBB-CPU/src/optim3_func.cpp:
541-662 and application/owm_optim3.cpp:143-230; it runs one tile per worker
thread here."""
import concurrent.futures as cf
import os
import time

import numpy as np
import pytest

import pyoracle as O
from fastowm_amd import owm, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _workers():
    return max(1, min(16, os.cpu_count() or 1))


def _oracle_many(jobs):
    """jobs: (x, y, z, wsize, bsize, overlap); one tile per worker thread
    (the oracle call releases the GIL), each single-threaded."""
    saved = O.threads()
    O.set_threads(1)
    try:
        with cf.ThreadPoolExecutor(_workers()) as ex:
            return list(ex.map(lambda j: O.run(*j, 32)[0], jobs))
    finally:
        O.set_threads(saved)


def _gen(tiles):
    arrs = [t.generate() for t in tiles]
    return arrs, [np.concatenate([a[k] for a in arrs]) for k in range(3)]


def test_segmented_16k_record_sort_tiles():
    import torch
    sizes = [400_000, 1_500_000, 650_000, 900_000, 1_200_000, 480_000]
    tiles = []
    for i, n in enumerate(sizes):
        t = synth.scaled(synth.config_c2(seed=20 + i), n)
        t.wsize, t.bsize, t.overlap = (12, 8, 16)[i % 3], (20, 15, 21)[i % 3], (0.8, 0.5, 0.9)[i % 3]
        tiles.append(t)
    assert sum(sizes) > 2 ** 21 and sum(sizes) / len(sizes) >= 65536
    arrs, (x, y, z) = _gen(tiles)
    desc = [(t.n, dict(wsize=t.wsize, bsize=t.bsize, overlap=t.overlap, leaf_size=32)) for t in tiles]
    say("oracle", len(tiles), "tiles", sum(sizes), "points")
    wants = _oracle_many([(a[0], a[1], a[2], t.wsize, t.bsize, t.overlap) for a, t in zip(arrs, tiles)])
    dev = [torch.from_numpy(v).to("cuda:0") for v in (x, y, z)]
    with owm.Context(0) as ctx:
        for mode in ("eager", "graph", "graph replay"):
            say("segmented 16k sort tiles", mode)
            got = ctx.batch_run(*dev, desc, graph=mode != "eager")
            for i, (g, w) in enumerate(zip(got, wants)):
                assert np.array_equal(g, w), f"{mode}: tile {i}"


def test_512_tile_batch_graph_replayed():
    import torch
    tiles = synth.batch_tiles(512)
    t0 = time.time()
    arrs, (x, y, z) = _gen(tiles)
    say("512 tiles generated", len(x), "points", round(time.time() - t0, 1), "s")
    desc = [(t.n, dict(wsize=t.wsize, bsize=t.bsize, overlap=t.overlap, leaf_size=0)) for t in tiles]
    dev = [torch.from_numpy(v).to("cuda:0") for v in (x, y, z)]
    del x, y, z
    with owm.Context(0) as ctx:
        say("512 tiles: capture + two replays")
        runs = [ctx.batch_run(*dev, desc, graph=True) for _ in range(3)]
    del dev
    torch.cuda.empty_cache()
    t0 = time.time()
    wants = _oracle_many([(a[0], a[1], a[2], t.wsize, t.bsize, t.overlap) for a, t in zip(arrs, tiles)])
    say("oracle on 512 tiles", round(time.time() - t0, 1), "s")
    for r, got in enumerate(runs):
        assert len(got) == 512
        bad = [i for i, (g, w) in enumerate(zip(got, wants)) if not np.array_equal(g, w)]
        assert not bad, f"run {r}: tiles {bad[:10]} differ"


def test_served_stream_10k_to_8m_against_oracle():
    reqs = synth.serve_requests(20, rate=2000.0, seed=11)
    data = [(t, t.generate()) for _, t in reqs]
    sizes = [t.n for t, _ in data]
    say("served stream", len(data), "requests", min(sizes), "-", max(sizes), "points")
    assert max(sizes) > 2_000_000 and min(sizes) < 100_000
    jobs = [(a[0], a[1], a[2], t.wsize, t.bsize, t.overlap) for t, a in data]
    # keyed re-queries of the largest tile with other parameters
    kt, (kx, ky, kz) = max(data, key=lambda d: d[0].n)
    requery = [(12, 20, 0.8), (7, 13, 0.55), (20, 30, 0.9)]
    jobs += [(kx, ky, kz, ws, bs, ov) for ws, bs, ov in requery]
    wants = _oracle_many(jobs)
    with owm.Context(0) as ctx:
        srv = owm.Server(ctx, max_inflight=3, max_batch=8)
        ids = []
        for i, (t, (x, y, z)) in enumerate(data):
            ids.append(srv.submit(x, y, z, t.wsize, t.bsize, t.overlap))
            if i == len(data) // 2:  # re-queries admitted while the others run
                ids += [srv.submit(kx, ky, kz, ws, bs, ov, key=77) for ws, bs, ov in requery]
        order = list(range(len(data) // 2 + 1)) + list(range(len(data), len(data) + 3)) + \
            list(range(len(data) // 2 + 1, len(data)))
        for k, rid in enumerate(ids):
            seeds, lat = srv.wait(rid)
            want = wants[order[k]]
            assert lat > 0 and np.array_equal(seeds, want), f"request {k}"
        say("served: alone equals in the mix")
        assert np.array_equal(ctx.run_tile(kx, ky, kz, 7, 13, 0.55), wants[len(data) + 1])
        srv.stop()
