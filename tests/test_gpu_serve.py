"""Serving loop (continuous batching): every result equals the same tile run
alone, for requests admitted together, while others run, and as re-queries of
resident trees; cancellation and budget rejection."""
import time

import numpy as np
import pytest

import pyoracle as O
from fastowm_amd import owm, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def test_mixed_stream_equals_alone():
    reqs = synth.serve_requests(24, rate=500.0, seed=7, lo=5_000, hi=300_000)
    data = []
    for t_arr, t in reqs:
        x, y, z = t.generate()
        data.append((x, y, z, t))
    with owm.Context(0) as ctx:
        say("alone")
        alone = [ctx.run_tile(x, y, z, t.wsize, t.bsize, t.overlap) for x, y, z, t in data]
        srv = owm.Server(ctx, max_inflight=3, max_batch=8)
        say("mixed: submit all, then a keyed re-query wave while they run")
        ids = [srv.submit(x, y, z, t.wsize, t.bsize, t.overlap) for x, y, z, t in data]
        # keyed tiles: same cloud, different parameters -> tree kept resident
        kx, ky, kz, kt = data[0]
        keyed = []
        for ws, bs, ov in [(12, 20, 0.8), (8, 15, 0.5), (16, 21, 0.9), (12, 20, 0.8)]:
            keyed.append(((ws, bs, ov), srv.submit(kx, ky, kz, ws, bs, ov, key=42)))
        for i, rid in enumerate(ids):
            seeds, lat = srv.wait(rid)
            assert lat > 0
            assert np.array_equal(seeds, alone[i]), f"request {i}"
        for (ws, bs, ov), rid in keyed:
            seeds, _ = srv.wait(rid)
            want, _, _ = O.run(kx, ky, kz, ws, bs, ov, 32)
            assert np.array_equal(seeds, want), (ws, bs, ov)
        say("evict + cancel")
        srv.evict(42)
        with pytest.raises(owm.OWMError):
            srv.evict(42)
        big = [srv.submit(x, y, z, t.wsize, t.bsize, t.overlap) for x, y, z, t in data[:6]]
        srv.cancel(big[-1])
        done = 0
        for rid in big:
            try:
                srv.wait(rid)
                done += 1
            except owm.OWMError as e:
                assert e.name == "E_CANCELLED"
        assert done >= len(big) - 1
        srv.stop()


def test_budget_rejection():
    say("budget")
    t = synth.scaled(synth.config_c1(), 400_000)
    x, y, z = t.generate()
    with owm.Context(0, hbm_budget=32 << 20) as ctx:
        srv = owm.Server(ctx, max_inflight=1)
        rid = srv.submit(x, y, z, 12, 20, 0.8)
        with pytest.raises(owm.OWMError) as e:
            srv.wait(rid)
        assert e.value.name == "E_BUDGET"
        small = synth.scaled(synth.config_c1(), 5_000)
        xs, ys, zs = small.generate()
        rid = srv.submit(xs, ys, zs, 12, 20, 0.8)
        seeds, _ = srv.wait(rid)
        assert np.array_equal(seeds, O.run(xs, ys, zs, 12, 20, 0.8, 32)[0])
        srv.stop()


def test_poll_until_done():
    say("poll")
    t = synth.scaled(synth.config_c1(), 100_000)
    x, y, z = t.generate()
    with owm.Context(0) as ctx:
        srv = owm.Server(ctx, max_inflight=2)
        rid = srv.submit(x, y, z, 12, 20, 0.8)
        deadline = time.time() + 60
        res = None
        while res is None and time.time() < deadline:
            res = srv.poll(rid)
            time.sleep(0.001)
        assert res is not None
        assert np.array_equal(res[0], O.run(x, y, z, 12, 20, 0.8, 32)[0])
        srv.stop()


def test_device_resident_requests():
    """Requests whose clouds are already in HBM (owm_serve_submit_device),
    mixed with host requests and keyed re-queries of a device-resident tile:
    every result equals the oracle's."""
    torch = pytest.importorskip("torch")
    reqs = synth.serve_requests(12, rate=1000.0, seed=21, lo=10_000, hi=600_000)
    data = [(t, t.generate()) for _, t in reqs]
    wants = [O.run(x, y, z, t.wsize, t.bsize, t.overlap, 32)[0] for t, (x, y, z) in data]
    devs = [[torch.from_numpy(a).to("cuda:0") for a in arrs] for _, arrs in data]
    kt, (kx, ky, kz) = data[0]
    kdev = devs[0]
    requery = [(12, 20, 0.8), (9, 17, 0.6)]
    kwant = [O.run(kx, ky, kz, ws, bs, ov, 32)[0] for ws, bs, ov in requery]
    with owm.Context(0) as ctx:
        srv = owm.Server(ctx, max_inflight=2, max_batch=8)
        say("device + host submits")
        ids = []
        for i, ((t, (x, y, z)), dev) in enumerate(zip(data, devs)):
            src = dev if i % 3 else (x, y, z)
            ids.append(srv.submit(*src, t.wsize, t.bsize, t.overlap))
        kids = [srv.submit(*kdev, ws, bs, ov, key=5) for ws, bs, ov in requery]
        for i, rid in enumerate(ids):
            seeds, _ = srv.wait(rid)
            assert np.array_equal(seeds, wants[i]), f"request {i}"
        for w, rid in zip(kwant, kids):
            seeds, _ = srv.wait(rid)
            assert np.array_equal(seeds, w)
        srv.stop()


def test_device_slices_read_in_place():
    """Tiles of one resident cloud submitted as consecutive device slices
    (groups of them are read in place, without a copy), then the same slices
    out of order and with an empty slice between them (copied groups): every
    result equals the oracle's, and the empty request fails alone, as it does
    when run by itself, without failing the requests batched with it."""
    torch = pytest.importorskip("torch")
    reqs = synth.serve_requests(10, rate=1000.0, seed=33, lo=10_000, hi=400_000)
    data = [(t, t.generate()) for _, t in reqs]
    wants = [O.run(x, y, z, t.wsize, t.bsize, t.overlap, 32)[0] for t, (x, y, z) in data]
    dev = [torch.from_numpy(np.concatenate([a[k] for _, a in data])).to("cuda:0") for k in range(3)]
    offs = np.concatenate([[0], np.cumsum([t.n for t, _ in data])])
    views = [[d[offs[i]:offs[i + 1]] for d in dev] for i in range(len(data))]
    empty = [d[offs[3]:offs[3]] for d in dev]
    orders = [list(range(len(data))), [1, 0, 2, 4, 3, "e", 5, 6, 9, 8, 7]]
    with owm.Context(0) as ctx:
        srv = owm.Server(ctx, max_inflight=1, max_batch=4)
        for order in orders:
            say("slices in order", order)
            ids = []
            for i in order:
                if i == "e":
                    ids.append((i, srv.submit(*empty, 12, 20, 0.8)))
                else:
                    t = data[i][0]
                    ids.append((i, srv.submit(*views[i], t.wsize, t.bsize, t.overlap)))
            for i, rid in ids:
                if i == "e":
                    with pytest.raises(owm.OWMError):
                        srv.wait(rid)
                else:
                    seeds, _ = srv.wait(rid)
                    assert np.array_equal(seeds, wants[i]), f"tile {i}"
        srv.stop()
        say("the empty slice alone")
        with pytest.raises(owm.OWMError):
            ctx.run_tile(*empty, 12, 20, 0.8)
