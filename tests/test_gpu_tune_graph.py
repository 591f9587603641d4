"""Captured batch chains (hipGraph replay) and the size-class autotuner: every
replay and every tuned configuration gives the eager, untuned seed sets."""
import numpy as np
import pytest

import pyoracle as O
from fastowm_amd import owm, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _batch(sizes, seed):
    tiles, arrs = [], []
    for i, n in enumerate(sizes):
        t = synth.scaled(synth.config_c1(seed + i), n)
        t.wsize = (8, 12, 16)[i % 3]
        t.bsize = (20, 15, 21)[i % 3]
        t.overlap = (0.8, 0.5, 0.9)[i % 3]
        tiles.append(t)
        arrs.append(t.generate())
    x, y, z = (np.concatenate([a[k] for a in arrs]) for k in range(3))
    desc = [(t.n, dict(wsize=t.wsize, bsize=t.bsize, overlap=t.overlap, leaf_size=0)) for t in tiles]
    return x, y, z, desc, tiles, arrs


def _same(a, b):
    return len(a) == len(b) and all(np.array_equal(u, v) for u, v in zip(a, b))


def test_graph_replays_equal_eager():
    torch = pytest.importorskip("torch")
    xa, ya, za, da, _, _ = _batch([30_000, 5_000, 120_000, 60_000, 9_000, 250_000], seed=100)
    xb, yb, zb, db, _, _ = _batch([70_000, 40_000, 3_000], seed=200)
    dev_a = [torch.from_numpy(v).to("cuda:0") for v in (xa, ya, za)]
    dev_b = [torch.from_numpy(v).to("cuda:0") for v in (xb, yb, zb)]
    with owm.Context(0) as ctx:
        say("eager references")
        ref_a = ctx.batch_run(*dev_a, da)
        ref_b = ctx.batch_run(*dev_b, db)
        for rep in range(3):
            say("graph replay of batch a", rep)
            assert _same(ctx.batch_run(*dev_a, da, graph=True), ref_a)
            # eager work and a synchronisation between replays
            assert _same(ctx.batch_run(*dev_b, db), ref_b)
            torch.cuda.synchronize()
        say("second layout captured beside the first")
        assert _same(ctx.batch_run(*dev_b, db, graph=True), ref_b)
        assert _same(ctx.batch_run(*dev_a, da, graph=True), ref_a)
        assert _same(ctx.batch_run(*dev_b, db, graph=True), ref_b)


def test_graph_from_host_arrays_matches_oracle():
    x, y, z, desc, tiles, arrs = _batch([20_000, 45_000], seed=300)
    with owm.Context(0) as ctx:
        for rep in range(2):
            say("host-input graph", rep)
            got = ctx.batch_run(x, y, z, desc, graph=True)
            for t, (tx, ty, tz), g in zip(tiles, arrs, got):
                want, _, _ = O.run(tx, ty, tz, t.wsize, t.bsize, t.overlap, 32)
                assert np.array_equal(g, want)


def test_autotune_keeps_results_and_table_round_trips():
    x, y, z, desc, tiles, arrs = _batch([30_000, 80_000, 300_000, 1_200_000, 50_000], seed=400)
    with owm.Context(0) as ctx:
        say("untuned batch")
        before = ctx.batch_run(x, y, z, desc)
        say("autotune")
        table = ctx.autotune(x, y, z, desc)
        assert len(table) == owm.TUNE_CLASSES
        assert [e["max_n"] for e in table] == [1 << 16, 1 << 18, 1 << 20, 1 << 22, 1 << 24, 2 ** 64 - 1]
        for e in table:
            assert e["leaf_size"] in (16, 32, 64, 128)
            assert e["sort_bits"] in (8, 16, 24, 32)
            assert (e["block_x"], e["block_y"], e["block_k"], e["summaries"]) in (
                (8, 8, 4, 0), (16, 8, 4, 0), (8, 8, 2, 0), (16, 8, 4, 1), (16, 16, 4, 0), (32, 16, 4, 0))
        # the classes that held tiles were timed
        assert all(table[c]["ms"] > 0 for c in (0, 1, 2, 3))
        say("tuned batch equals untuned")
        assert _same(ctx.batch_run(x, y, z, desc), before)
        say("table round trip")
        assert ctx.tune_table() == table
        ctx.set_tune_table(table)
        assert ctx.tune_table() == table
        bad = dict(table[0])
        bad["sort_bits"] = 20
        with pytest.raises(owm.OWMError) as e:
            ctx.set_tune_table([bad])
        assert e.value.name == "E_ARG"


@pytest.mark.parametrize("sort_bits,leaf,block", [(24, 16, (8, 8, 2, 0)), (24, 128, (16, 8, 4, 0)),
                                                  (32, 64, (16, 8, 4, 0)), (24, 32, (16, 8, 4, 1)),
                                                  (32, 16, (8, 8, 4, 1)), (16, 32, (16, 8, 4, 0)),
                                                  (16, 64, (16, 16, 4, 0)), (16, 128, (32, 8, 4, 0)),
                                                  (16, 64, (32, 16, 4, 0)), (24, 16, (32, 16, 4, 0))])
def test_forced_table_entries_match_oracle(sort_bits, leaf, block):
    say("table entry", sort_bits, leaf, block)
    t = synth.scaled(synth.config_c2(seed=5), 700_000)
    x, y, z = t.generate()
    want, _, _ = O.run(x, y, z, 12, 20, 0.8, 32)
    x2, y2, z2, desc, tiles, arrs = _batch([40_000, 90_000, 20_000], seed=500)
    with owm.Context(0) as ctx:
        entries = ctx.tune_table()
        for e in entries:
            e["sort_bits"], e["leaf_size"] = sort_bits, leaf
            e["block_x"], e["block_y"], e["block_k"], e["summaries"] = block
        ctx.set_tune_table(entries)
        assert np.array_equal(ctx.run_tile(x, y, z, 12, 20, 0.8, 0), want)
        got = ctx.batch_run(x2, y2, z2, desc)
        for tt, (tx, ty, tz), g in zip(tiles, arrs, got):
            w, _, _ = O.run(tx, ty, tz, tt.wsize, tt.bsize, tt.overlap, 32)
            assert np.array_equal(g, w)


def test_many_small_tiles_over_two_million_points():
    """A batch above 2M points whose tiles average under 64K points sorts with
    4096-point tiles (every tile of the batch ends in a partial sort tile);
    each tile's seeds equal the oracle's, eager and replayed."""
    sizes = [40_000 + 3_000 * (i % 11) for i in range(44)]
    assert sum(sizes) > 2 ** 21 and sum(sizes) / len(sizes) < 65536
    x, y, z, desc, tiles, arrs = _batch(sizes, seed=600)
    with owm.Context(0) as ctx:
        say("many small tiles", len(sizes), sum(sizes))
        got = ctx.batch_run(x, y, z, desc)
        for t, (tx, ty, tz), g in zip(tiles, arrs, got):
            want, _, _ = O.run(tx, ty, tz, t.wsize, t.bsize, t.overlap, 32)
            assert np.array_equal(g, want)
        assert _same(ctx.batch_run(x, y, z, desc, graph=True), got)


def test_repeated_single_tile_calls_replay_a_graph():
    """owm_run_tile* replays the chain from a graph when a call repeats the
    previous call's layout exactly (captured on the first repeat): repeats of
    one tile, with other work between them, and host-input calls whose clouds
    differ but share the upload buffer, size and parameters, all give the
    oracle's seeds."""
    torch = pytest.importorskip("torch")
    t = synth.scaled(synth.config_c2(seed=31), 400_000)
    x, y, z = t.generate()
    box = synth.bbox(x, y)
    want, _, _ = O.run(x, y, z, 12, 20, 0.8, 32)
    other = synth.scaled(synth.config_c1(seed=32), 50_000)
    ox, oy, oz = other.generate()
    owant, _, _ = O.run(ox, oy, oz, 10, 15, 0.6, 16)
    dev = [torch.from_numpy(a).to("cuda:0") for a in (x, y, z)]
    with owm.Context(0) as ctx:
        for rep in range(4):
            say("device repeat", rep)
            assert np.array_equal(ctx.run_tile(*dev, 12, 20, 0.8, 32), want)
            assert np.array_equal(ctx.run_tile(*dev, 12, 20, 0.8, 32), want)
            if rep % 2:
                assert np.array_equal(ctx.run_tile(ox, oy, oz, 10, 15, 0.6, 16), owant)
            torch.cuda.synchronize()
        # same size, bbox and parameters, other points: the replayed chain
        # reads the new upload
        rng = np.random.default_rng(5)
        perm = rng.permutation(len(x))
        x2, y2, z2 = x[perm], y[perm], z[perm] + rng.normal(0, 0.5, len(x))
        want2, _, _ = O.run(x2, y2, z2, 12, 20, 0.8, 32, box)
        for rep in range(3):
            say("host repeat", rep)
            assert np.array_equal(ctx.run_tile(x, y, z, 12, 20, 0.8, 32, bbox=box), want)
            assert np.array_equal(ctx.run_tile(x2, y2, z2, 12, 20, 0.8, 32, bbox=box), want2)


def test_batch_desc_and_buffer_views():
    """A BatchDesc built once gives the same seeds as the tile list; with
    copy=False the results are views of the context's seed buffer (valid until
    the next call), equal to the copies; split=False gives the flat array and
    per-tile offsets."""
    x, y, z, desc, tiles, arrs = _batch([30_000, 70_000, 5_000], seed=700)
    with owm.Context(0) as ctx:
        say("batch desc / views")
        want = ctx.batch_run(x, y, z, desc)
        bd = owm.batch_desc(desc)
        assert len(bd) == 3 and bd.points == len(x)
        views = ctx.batch_run(x, y, z, bd, copy=False)
        assert _same([v.copy() for v in views], want)
        flat, offs = ctx.batch_run(x, y, z, bd, split=False)
        assert len(offs) == 4 and offs[0] == 0 and offs[-1] == len(flat)
        assert _same([flat[offs[i]:offs[i + 1]] for i in range(3)], want)
        assert _same(ctx.batch_run(x, y, z, bd, graph=True), want)
        for t, (tx, ty, tz), g in zip(tiles, arrs, want):
            w, _, _ = O.run(tx, ty, tz, t.wsize, t.bsize, t.overlap, 32)
            assert np.array_equal(g, w)


def _stats_samples(ctx, x, y, desc_list):
    """The samples owm_autotune_stats takes: up to OWM_TUNE_SAMPLES tiles per
    class, spread over the batch, each standing for its class's share."""
    offs = np.cumsum([0] + [n for n, _ in desc_list])
    members = {}
    for t, (n, _) in enumerate(desc_list):
        if n:
            members.setdefault(next(i for i, m in enumerate(CLASS_MAX) if n <= m), []).append(t)
    samples = []
    for c, mem in sorted(members.items()):
        k = min(len(mem), 16)
        chosen = [mem[(2 * j + 1) * len(mem) // (2 * k)] for j in range(k)]
        got = float(sum(desc_list[t][0] for t in chosen))
        cls_pts = float(sum(desc_list[t][0] for t in mem))
        for t in chosen:
            n, prm = desc_list[t]
            bins, sm = ctx.qtree_stats(x[offs[t]:offs[t] + n], y[offs[t]:offs[t] + n],
                                       policy=owm.QT_MAXNUMBER, max_number=owm.TUNE_STATS_MAXNUMBER,
                                       level=owm.TUNE_STATS_LEVEL)
            samples.append((n, n * cls_pts / got, prm, bins, sm))
    return samples


CLASS_MAX = [1 << 16, 1 << 18, 1 << 20, 1 << 22, 1 << 24, 2 ** 64 - 1]


def test_stats_autotune_keeps_results_and_matches_the_model():
    """owm_autotune_stats: the table is the model's pick from the sampled
    tiles' histograms (as owm_tune_model gives it), host and device inputs give
    the same table, every class present gets a predicted time, and the tuned
    batch gives the untuned seeds, equal to the oracle's."""
    torch = pytest.importorskip("torch")
    sizes = [30_000, 80_000, 300_000, 1_200_000, 50_000] + [20_000 + 7_000 * i for i in range(20)]
    x, y, z, desc, tiles, arrs = _batch(sizes, seed=410)
    with owm.Context(0) as ctx:
        say("untuned batch")
        before = ctx.batch_run(x, y, z, desc)
        say("statistics autotune, host inputs")
        table = ctx.autotune(x, y, z, desc, method="stats")
        assert len(table) == owm.TUNE_CLASSES
        for e in table:
            assert e["leaf_size"] in (16, 32, 64, 128) and e["sort_bits"] in (8, 16, 24, 32)
        assert all(table[c]["ms"] > 0 for c in (0, 1, 2, 3))
        say("device inputs")
        dev = [torch.from_numpy(a).to("cuda:0") for a in (x, y, z)]
        assert ctx.autotune(*dev, desc, method="stats") == table
        say("model on the same samples")
        want = owm.tune_model(_stats_samples(ctx, x, y, desc))
        for a, b in zip(table, want):
            assert {k: v for k, v in a.items() if k != "ms"} == {k: v for k, v in b.items() if k != "ms"}
            assert a["ms"] == pytest.approx(b["ms"], rel=1e-9)
        say("tuned batch equals untuned and the oracle")
        after = ctx.batch_run(x, y, z, desc)
        assert _same(after, before)
        for t, (tx, ty, tz), g in list(zip(tiles, arrs, after))[:6]:
            w, _, _ = O.run(tx, ty, tz, t.wsize, t.bsize, t.overlap, 32)
            assert np.array_equal(g, w)


def test_stats_autotune_skips_a_tile_it_cannot_histogram():
    """A tile with more coincident points than a histogram leaf may hold has no
    histogram; the tuner samples the others and the batch still matches the
    oracle."""
    rng = np.random.default_rng(9)
    x1 = np.full(4_000, 500_100.0) + np.repeat(rng.uniform(0, 40, 4), 1_000)
    y1 = np.full(4_000, 4_200_100.0) + np.repeat(rng.uniform(0, 40, 4), 1_000)
    z1 = rng.uniform(0, 5, 4_000)
    t2 = synth.scaled(synth.config_c1(seed=11), 60_000)
    x2, y2, z2 = t2.generate()
    x, y, z = (np.concatenate(a) for a in ((x1, x2), (y1, y2), (z1, z2)))
    desc = [(4_000, dict(wsize=12, bsize=20, overlap=0.8, leaf_size=0)),
            (60_000, dict(wsize=12, bsize=20, overlap=0.8, leaf_size=0))]
    with owm.Context(0) as ctx:
        say("coincident tile in the stats tuner")
        with pytest.raises(owm.OWMError):
            ctx.qtree_stats(x1, y1, policy=owm.QT_MAXNUMBER, max_number=owm.TUNE_STATS_MAXNUMBER,
                            level=owm.TUNE_STATS_LEVEL)
        table = ctx.autotune(x, y, z, desc, method="stats")
        assert table[0]["ms"] > 0
        got = ctx.batch_run(x, y, z, desc)
        for (tx, ty, tz), g in (((x1, y1, z1), got[0]), ((x2, y2, z2), got[1])):
            w, _, _ = O.run(tx, ty, tz, 12, 20, 0.8, 32)
            assert np.array_equal(g, w)
