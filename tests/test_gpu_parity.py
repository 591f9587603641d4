"""HIP path vs the CPU oracle (bit-exact seed sets and per-window minima).

Every case prints its name before touching the device, so a failing run names it.
"""
import json
import os
import sys

import numpy as np
import pytest

import pyoracle as O
from fastowm_amd import owm, synth

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = owm.Context(0)
    yield c
    c.close()


def _tile(n, seed=0, **kw):
    t = synth.scaled(synth.config_c1(seed), n)
    for k, v in kw.items():
        setattr(t, k, v)
    return t


SMALL = [
    # n, wsize, bsize, overlap, leaf, generator overrides
    (3000, 12, 20, 0.8, 32, {}),
    (3000, 12, 20, 0.8, 1, {}),
    (5000, 13, 21, 0.8, 8, {"cluster_frac": 0.5, "n_clusters": 5}),
    (5000, 7, 11, 0.5, 16, {}),
    (4000, 10, 20, 0.9, 128, {"cluster_frac": 0.3, "n_clusters": 3}),
    (4000, 8, 15, 0.0, 32, {}),
    (6000, 16, 30, 0.75, 64, {"roof_frac": 0.3, "n_roofs": 4}),
    (2500, 5, 10, 0.6, 3, {}),
    (50_000, 12, 20, 0.8, 32, {"cluster_frac": 0.7, "n_clusters": 6, "cluster_sigma": 8.0}),
    (200_000, 12, 20, 0.8, 0, {"cluster_frac": 0.1, "n_clusters": 20}),
]


@pytest.mark.parametrize("case", SMALL, ids=[f"n{c[0]}_w{c[1]}_b{c[2]}_o{c[3]}_l{c[4]}" for c in SMALL])
def test_small_tiles_match_oracle(ctx, case):
    n, ws, bs, ov, leaf, over = case
    say("small", case)
    x, y, z = _tile(n, seed=n + ws, **over).generate()
    want, _, _ = O.run(x, y, z, ws, bs, ov, leaf or 32)
    got = ctx.run_tile(x, y, z, ws, bs, ov, leaf)
    assert np.array_equal(got, want)


def test_stage_api_matches_oracle(ctx):
    say("stage api 30k")
    x, y, z = _tile(30_000, seed=4, cluster_frac=0.6, n_clusters=4).generate()
    box = synth.bbox(x, y)
    seeds, mins, llp, adds, p = O.seeds_from_stages(x, y, z, 12, 20, 0.8, 16, box)
    tree = ctx.parallel_qtree(x, y, z, leaf_size=16)
    info = tree.info()
    assert info["root_count"] == len(x)
    assert info["root_bbox"] == box
    assert info["leaves"] == (len(x) + 15) // 16 and info["internal"] == info["leaves"] - 1
    gmins, d = ctx.stage1(tree, 12, 0.8, 20)
    assert (d.ncols, d.nrows) == (p.ncols, p.nrows)
    assert np.array_equal(gmins, mins), "stage-1 window minima differ"
    gllp = ctx.stage2(tree, gmins, 12, 0.8, 20)
    assert np.array_equal(gllp, llp)
    gadds = ctx.stage3(tree, gllp, 12, 0.8, 20)
    assert np.array_equal(gadds, adds), "stage-3 additions differ (cell order)"
    assert np.array_equal(ctx.tree_run(tree, 12, 20, 0.8), seeds)
    # re-query the resident tree with other parameters
    for ws, bs, ov in [(8, 15, 0.5), (16, 21, 0.9), (13, 30, 0.0)]:
        want, _, _ = O.run(x, y, z, ws, bs, ov, 16, box)
        assert np.array_equal(ctx.tree_run(tree, ws, bs, ov), want)
    tree.free()


def test_golden_reference_fixtures(ctx):
    gold = json.load(open(os.path.join(HERE, "golden", "ref_o1qtree.json")))
    for c in gold["cases"]:
        say("golden", c["name"])
        t = synth.scaled(synth.config_c1(c["generator"].get("seed", 0)), c["n"])
        for k, v in c["generator"].items():
            setattr(t, k, v)
        x, y, z = t.generate()
        got = ctx.run_tile(x, y, z, c["wsize"], c["bsize"], c["overlap"], 32, bbox=c["bbox"])
        assert got.tolist() == c["ref_seeds"], c["name"]


def test_edge_planted_and_ties(ctx):
    sys.path.insert(0, HERE)
    from test_oracle import _edge_planted
    say("edge planted")
    x, y, z, p = _edge_planted()
    box = (p.minx, p.maxx, p.miny, p.maxy)
    tree = ctx.parallel_qtree(x, y, z, leaf_size=8)
    gm, _ = ctx.stage1(tree, 12, 0.8, 20, bbox=box)
    assert np.array_equal(gm, O.stage1_brute(x, y, z, p))
    tree.free()
    say("ties")
    x, y, z = _tile(20_000, seed=11).generate()
    z = np.minimum(np.round(z), 300.0)
    want, _, _ = O.run(x, y, z, 10, 20, 0.8, 8)
    assert np.array_equal(ctx.run_tile(x, y, z, 10, 20, 0.8, 8), want)


def test_nan_inf_signed_zero(ctx):
    say("nan/inf/-0")
    x, y, z = _tile(3000, seed=5).generate()
    x[10], y[20], z[30], z[40] = np.nan, np.nan, np.nan, np.inf
    z[50], z[51] = 0.0, -0.0
    x[51], y[51] = x[50], y[50]
    box = synth.bbox(x, y)
    p = O.params(12, 20, 0.8, box, len(x), 4)
    tree = ctx.parallel_qtree(x, y, z, leaf_size=4)
    gm, _ = ctx.stage1(tree, 12, 0.8, 20)
    assert np.array_equal(gm, O.stage1_brute(x, y, z, p))
    assert 50 in gm and 51 not in gm
    tree.free()


def test_duplicates_beyond_leaf(ctx):
    say("coincident points")
    x, y, z = _tile(5000, seed=9).generate()
    x[100:1400] = x[0]
    y[100:1400] = y[0]
    want, _, _ = O.run(x, y, z, 12, 20, 0.8, 8)
    assert np.array_equal(ctx.run_tile(x, y, z, 12, 20, 0.8, 8), want)


def test_batch_matches_per_tile_oracle(ctx):
    say("batch 24 tiles")
    tiles = synth.batch_tiles(24, seed=3, median=20_000, lo=2_000, hi=120_000)
    xs, ys, zs, desc, wants = [], [], [], [], []
    for i, t in enumerate(tiles):
        x, y, z = t.generate()
        ov = (0.8, 0.5, 0.0, 0.9)[i % 4]
        bs = (20, 21, 0, 15)[i % 4]
        xs.append(x)
        ys.append(y)
        zs.append(z)
        desc.append((t.n, dict(wsize=t.wsize, bsize=bs, overlap=ov, leaf_size=(0, 16, 64)[i % 3])))
        wants.append(O.run(x, y, z, t.wsize, bs, ov, 32)[0])
    got = ctx.batch_run(np.concatenate(xs), np.concatenate(ys), np.concatenate(zs), desc)
    for i, (g, w) in enumerate(zip(got, wants)):
        assert np.array_equal(g, w), f"tile {i}"


def test_stage1_join_rerun_and_empty_blocks(ctx):
    """A leaf/block join list too small for the data forces the re-run path;
    blocks left without leaves must still publish their window class ranges.
    (Regression: Overlap 0.9 tile whose join exceeded the first guess.)"""
    tiles = synth.batch_tiles(24, seed=3, median=20_000, lo=2_000, hi=120_000)
    t = tiles[3]
    x, y, z = t.generate()
    want, _, _ = O.run(x, y, z, t.wsize, 15, 0.9, 32)
    say("overlap 0.9 tile, default join size")
    assert np.array_equal(ctx.run_tile(x, y, z, t.wsize, 15, 0.9, 0), want)
    say("overlap 0.9 tile, join list forced to 64 entries")
    os.environ["OWM_S1_LIST_CAP"] = "64"
    try:
        got = ctx.run_tile(x, y, z, t.wsize, 15, 0.9, 0)
        tree = ctx.parallel_qtree(x, y, z, leaf_size=32)
        gm, _ = ctx.stage1(tree, t.wsize, 0.9, 15)
        tree.free()
    finally:
        del os.environ["OWM_S1_LIST_CAP"]
    assert np.array_equal(got, want)
    p = O.params(t.wsize, 15, 0.9, synth.bbox(x, y), len(x))
    assert np.array_equal(gm, O.Tree(x, y, z, p).stage1())


def test_device_resident_input(ctx):
    import torch
    say("device input c1")
    t = synth.config_c1()
    x, y, z = t.generate()
    want, _, _ = O.run(x, y, z, 12, 20, 0.8, 32)
    dx, dy, dz = (torch.from_numpy(a).cuda() for a in (x, y, z))
    got = ctx.run_tile(dx, dy, dz, 12, 20, 0.8, 32)
    assert np.array_equal(got, want)


def test_c2_full_size(ctx):
    say("c2 20M")
    t = synth.config_c2()
    x, y, z = t.generate()
    want, _, p = O.run(x, y, z, 12, 20, 0.8, 32)
    got = ctx.run_tile(x, y, z, 12, 20, 0.8, 32)
    assert np.array_equal(got, want)


def test_c3_full_size_roofs_with_ties(ctx):
    """120M points, 3x3 km, 30% on flat roofs with 1-cm z: deep trees and exact
    z ties (lowest index wins), bit-exact against the oracle; the stage-1
    minima of the resident tree also equal the oracle's per window."""
    say("c3 120M")
    t = synth.config_c3()
    x, y, z = t.generate()
    assert len(np.unique(z[: 200_000])) < 200_000  # ties are present
    want, _, p = O.run(x, y, z, 12, 20, 0.8, 32)
    got = ctx.run_tile(x, y, z, 12, 20, 0.8, 0)
    assert np.array_equal(got, want)
    say("c3 120M, default (untuned) leaf 32 and a re-query with other parameters")
    tree = ctx.parallel_qtree(x, y, z, leaf_size=32)
    seeds = ctx.tree_run(tree, 16, 30, 0.75)
    tree.free()
    want2, _, _ = O.run(x, y, z, 16, 30, 0.75, 32)
    assert np.array_equal(seeds, want2)


def test_profile_phase_mask(ctx):
    """Phase timing limited to one phase records that phase only (one event
    pair per multi-pass sort, counted as its passes) and leaves the seeds as
    they are; with no mask every phase of the chain is timed."""
    say("profile phase mask")
    t = _tile(60_000, seed=41)
    x, y, z = t.generate()
    want, _, _ = O.run(x, y, z, 12, 20, 0.8, 32)
    ctx.profile(True, phases=["sort_tree_pass"])
    got = ctx.run_tile(x, y, z, 12, 20, 0.8, 32)
    only = ctx.profile_read()
    ctx.profile(True)
    ctx.run_tile(x, y, z, 12, 20, 0.8, 32)
    full = ctx.profile_read()
    ctx.profile(False)
    assert np.array_equal(got, want)
    timed = {k for k, v in only.items() if v["calls"]}
    assert timed == {"sort_tree_pass"}
    entry = next(e for e in ctx.tune_table() if len(x) <= e["max_n"])
    passes = entry["sort_bits"] // 8
    assert only["sort_tree_pass"]["calls"] == passes and only["sort_tree_pass"]["ms"] > 0
    assert only["sort_tree_pass"]["calls"] == full["sort_tree_pass"]["calls"]
    for k in ("bbox", "leaf_aggs", "karras", "agg_levels", "stage1", "stage3"):
        assert full[k]["calls"] >= 1, k
    # the bbox pass takes the key digit counts too (no Morton pass); with the
    # bbox given there is no bbox pass and the Morton pass takes them
    assert full["morton"]["calls"] == 0
    ctx.profile(True)
    assert np.array_equal(ctx.run_tile(x, y, z, 12, 20, 0.8, 32, bbox=synth.bbox(x, y)), want)
    given = ctx.profile_read()
    ctx.profile(False)
    assert given["morton"]["calls"] == 1 and given["bbox"]["calls"] == 0


def test_errors(ctx):
    say("errors")
    x, y, z = _tile(2000).generate()
    with pytest.raises(owm.OWMError) as e:
        ctx.run_tile(x, y, z, 12, 20, 1.0)
    assert e.value.name == "E_ARG"
    with pytest.raises(owm.OWMError) as e:
        ctx.run_tile(x, y, z, 12, 20, 0.8, cap=1)
    assert e.value.name == "E_CAPACITY"
    # the context stays usable after errors
    want, _, _ = O.run(x, y, z, 12, 20, 0.8, 32)
    assert np.array_equal(ctx.run_tile(x, y, z, 12, 20, 0.8, 32), want)


def _lattice_cloud(n, seed, side=200):
    """Integer coordinates over [0, side]: with Wsize 10, Overlap 0.9 (step 1 m)
    every window edge is an integer, so many points lie exactly on edges."""
    rng = np.random.default_rng(seed)
    x = rng.integers(0, side + 1, n).astype(np.float64)
    y = rng.integers(0, side + 1, n).astype(np.float64)
    z = np.round(rng.normal(50.0, 3.0, n), 1)  # 0.1 m steps: many equal z
    x[0], y[0], x[1], y[1] = 0.0, 0.0, float(side), float(side)  # pin the bbox
    return x, y, z


def _stage1_table(ctx, summaries, block=(16, 8, 4)):
    entries = ctx.tune_table()
    for e in entries:
        e["block_x"], e["block_y"], e["block_k"] = block
        e["summaries"] = 1 if summaries == "on" else 0
    ctx.set_tune_table(entries)


@pytest.fixture
def restore_table(ctx):
    saved = ctx.tune_table()
    yield
    ctx.set_tune_table(saved)


@pytest.mark.parametrize("summaries", ["on", "off"])
@pytest.mark.parametrize("ws,bs,ov,leaf", [(10, 20, 0.9, 16), (10, 15, 0.5, 32), (12, 20, 0.8, 64),
                                           (7, 11, 0.5, 8)])
def test_points_on_edges_with_and_without_summaries(ctx, restore_table, summaries, ws, bs, ov, leaf):
    """Stage 1 with leaf cell summaries (classes from the global edge multiset)
    and point by point must agree with the brute-force window scan, also where
    points sit exactly on window edges and z ties."""
    _stage1_table(ctx, summaries, (16, 8, 4) if leaf % 2 else (8, 8, 4))
    say("lattice", summaries, ws, bs, ov, leaf)
    x, y, z = _lattice_cloud(40_000, seed=ws * 100 + leaf)
    box = synth.bbox(x, y)
    p = O.params(ws, bs, ov, box, len(x), leaf)
    tree = ctx.parallel_qtree(x, y, z, leaf_size=leaf)
    gm, _ = ctx.stage1(tree, ws, ov, bs)
    assert np.array_equal(gm, O.stage1_brute(x, y, z, p))
    tree.free()
    want, _, _ = O.run(x, y, z, ws, bs, ov, leaf)
    assert np.array_equal(ctx.run_tile(x, y, z, ws, bs, ov, leaf), want)


def _near_edge_cloud(n, seed, side, origin):
    """Integer lattice (window edges are integers with Wsize 10, Overlap 0.9)
    moved by 0, 1-3 ulps, 1e-9, 1e-6 or 0.5 either way, at the origin or at
    projected-grid magnitudes: points just inside and just outside the edges,
    where the arithmetic class estimate must hand over to exact compares."""
    rng = np.random.default_rng(seed)
    x0 = origin[0] + rng.integers(0, side + 1, n).astype(np.float64)
    y0 = origin[1] + rng.integers(0, side + 1, n).astype(np.float64)

    def nudge(v):
        kind = rng.integers(0, 9, len(v))
        out = v.copy()
        for k, steps in ((1, 1), (2, 2), (3, 3)):
            m = kind == k
            sgn = np.where(rng.integers(0, 2, m.sum()) == 1, np.inf, -np.inf)
            w = out[m]
            for _ in range(steps):
                w = np.nextafter(w, sgn)
            out[m] = w
        for k, d in ((4, 1e-9), (5, 1e-6), (6, 0.5)):
            m = kind == k
            out[m] += d * np.where(rng.integers(0, 2, m.sum()) == 1, 1.0, -1.0)
        return out

    x, y = nudge(x0), nudge(y0)
    z = np.round(rng.normal(50.0, 3.0, n), 1)
    x[0], y[0], x[1], y[1] = origin[0], origin[1], origin[0] + side, origin[1] + side
    return x, y, z


@pytest.mark.parametrize("block", [(16, 16, 4), (32, 8, 4), (32, 16, 4)])
@pytest.mark.parametrize("ws,bs,ov,leaf", [(10, 20, 0.9, 16), (7, 11, 0.5, 8), (12, 20, 0.8, 64)])
def test_points_on_edges_large_blocks(ctx, restore_table, block, ws, bs, ov, leaf):
    """The larger window blocks (512- and 1024-thread workgroups) give the
    brute-force window scan on a lattice whose points sit on window edges."""
    _stage1_table(ctx, "off", block)
    say("lattice, large block", block, ws, bs, ov, leaf)
    x, y, z = _lattice_cloud(60_000, seed=ws * 31 + leaf)
    box = synth.bbox(x, y)
    p = O.params(ws, bs, ov, box, len(x), leaf)
    tree = ctx.parallel_qtree(x, y, z, leaf_size=leaf)
    gm, _ = ctx.stage1(tree, ws, ov, bs)
    assert np.array_equal(gm, O.stage1_brute(x, y, z, p))
    tree.free()


@pytest.mark.parametrize("block", [(16, 16, 4), (32, 16, 4)])
@pytest.mark.parametrize("origin", [(0.0, 0.0), (612_345.0, 4_512_345.0)])
def test_stamped_classes_near_edges(ctx, restore_table, block, origin):
    """Whole runs build the tree and plan the query in one chain, so the tree
    sort stamps every point's global window classes and stage 1 bins by them;
    points a few ulps from an edge are not stamped and are binned by their
    coordinates.  Seeds against the oracle, with the larger window blocks."""
    _stage1_table(ctx, "off", block)
    for ws, bs, ov, leaf in ((10, 20, 0.9, 16), (7, 11, 0.5, 64), (12, 20, 0.7, 32)):
        say("stamped, near edges", block, origin, ws, bs, ov, leaf)
        x, y, z = _near_edge_cloud(30_000, seed=ws * 7 + leaf, side=200, origin=origin)
        want, _, _ = O.run(x, y, z, ws, bs, ov, leaf)
        assert np.array_equal(ctx.run_tile(x, y, z, ws, bs, ov, leaf), want)


def test_stamped_classes_high_overlap_and_wide_grid(ctx, restore_table):
    """Blocks whose class range exceeds the class tables (overlap 0.995 with
    32x16 blocks) read coordinates; a tile with more than 16383 window columns
    is not stamped at all.  Both in one batch with an ordinary tile."""
    _stage1_table(ctx, "off", (32, 16, 4))
    rng = np.random.default_rng(5)
    clouds = []
    # 2 m windows at overlap 0.995: a step of 1 cm, 200 steps a window
    clouds.append(((rng.random(20_000) * 10.0, rng.random(20_000) * 10.0,
                    np.round(rng.normal(50, 2, 20_000), 2)), (2, 10, 0.995)))
    # Wsize 2 at overlap 0.9 over 3.5 km: 17500 window columns
    n = 30_000
    clouds.append(((rng.random(n) * 3500.0, rng.random(n) * 6.0, np.round(rng.normal(50, 2, n), 2)),
                   (2, 10, 0.9)))
    x, y, z = _tile(40_000, seed=9).generate()
    clouds.append(((x, y, z), (12, 20, 0.8)))
    wants = []
    for (cx, cy, cz), (ws, bs, ov) in clouds:
        say("stamp fallback tile", len(cx), ws, bs, ov)
        wants.append(O.run(cx, cy, cz, ws, bs, ov, 32)[0])
        assert np.array_equal(ctx.run_tile(cx, cy, cz, ws, bs, ov, 32), wants[-1])
    desc = [(len(c[0][0]), dict(wsize=c[1][0], bsize=c[1][1], overlap=c[1][2], leaf_size=32))
            for c in clouds]
    got = ctx.batch_run(*(np.concatenate([c[0][k] for c in clouds]) for k in range(3)), desc)
    for i, (g, w) in enumerate(zip(got, wants)):
        assert np.array_equal(g, w), f"tile {i}"


@pytest.mark.parametrize("origin", [(0.0, 0.0), (612_345.0, 4_512_345.0)])
def test_window_columns_near_edges(ctx, origin):
    """A window side of a whole number of steps (m = 2, 4, 5 and 7 here) runs
    stage 1 on window columns when the tree is built with its query; points on
    an edge (the integer lattice, where these grids put integer edges) or a few
    ulps from one carry no column and take their windows by exact compares,
    many per block (more than one pass of exceptions per round).  Seeds
    against the oracle, tile by tile and as one batch."""
    cfgs = [(8, 20, 0.5, 16), (10, 20, 0.75, 32), (12, 20, 0.8, 64), (14, 20, 1.0 - 1.0 / 7.0, 32)]
    clouds, wants = [], []
    for i, (ws, bs, ov, leaf) in enumerate(cfgs):
        if i % 2:
            x, y, z = _lattice_cloud(30_000, seed=41 + i)
            x, y = x + origin[0], y + origin[1]
        else:
            x, y, z = _near_edge_cloud(30_000, seed=43 + i, side=200, origin=origin)
        say("window columns", origin, ws, bs, ov, leaf)
        want, _, _ = O.run(x, y, z, ws, bs, ov, leaf)
        assert np.array_equal(ctx.run_tile(x, y, z, ws, bs, ov, leaf), want)
        clouds.append((x, y, z))
        wants.append(want)
    say("window columns, one batch", origin)
    desc = [(len(c[0]), dict(wsize=ws, bsize=bs, overlap=ov, leaf_size=leaf))
            for c, (ws, bs, ov, leaf) in zip(clouds, cfgs)]
    got = ctx.batch_run(*(np.concatenate([c[k] for c in clouds]) for k in range(3)), desc)
    for i, (g, w) in enumerate(zip(got, wants)):
        assert np.array_equal(g, w), f"tile {i}"


def test_class_cells_and_window_columns_agree(ctx, monkeypatch):
    """The same batch (every tile's window side 5 steps) through stage 1 on
    window columns and, with OWM_NO_PAIRS=1, on class cells: both give the
    oracle's seeds for every tile."""
    tiles = synth.batch_tiles(6, seed=12, median=300_000, lo=100_000, hi=900_000)
    clouds = [t.generate() for t in tiles]
    wants = [O.run(*c, t.wsize, t.bsize, t.overlap, 32)[0] for c, t in zip(clouds, tiles)]
    desc = [(t.n, dict(wsize=t.wsize, bsize=t.bsize, overlap=t.overlap, leaf_size=0)) for t in tiles]
    xyz = [np.concatenate([c[k] for c in clouds]) for k in range(3)]
    for mode in ("columns", "classes"):
        if mode == "classes":
            monkeypatch.setenv("OWM_NO_PAIRS", "1")
        say("batch of 6 through", mode)
        got = ctx.batch_run(*xyz, desc)
        for i, (g, w) in enumerate(zip(got, wants)):
            assert np.array_equal(g, w), f"{mode}: tile {i}"


def test_window_columns_large_blocks_on_edges(ctx):
    """A launch with more 32x32-window blocks than workgroup slots runs the
    32x32 column blocks: one tile of 1000 x 1000 windows (Wsize 8, Overlap
    0.5, integer edges every 4 m) whose integer-lattice points lie on an edge
    along x or y about half the time, so nearly every round of every block
    takes several passes of exceptions.  Seeds against the oracle."""
    rng = np.random.default_rng(77)
    n = 1_000_000
    x = rng.integers(0, 4001, n).astype(np.float64)
    y = rng.integers(0, 4001, n).astype(np.float64)
    z = np.round(rng.normal(50.0, 3.0, n), 1)
    say("large column blocks, edge lattice", n)
    want, _, _ = O.run(x, y, z, 8, 20, 0.5, 32)
    assert np.array_equal(ctx.run_tile(x, y, z, 8, 20, 0.5, 32), want)


@pytest.mark.parametrize("origin", [(0.0, 0.0), (612_345.0, 4_512_345.0)])
def test_points_ulps_from_edges(ctx, restore_table, origin):
    """Stage 1 classifies by arithmetic away from edges and by exact compares
    near them; points a few ulps either side of an edge, at small and at
    projected-grid coordinates, must give the brute-force window scan."""
    _stage1_table(ctx, "off")
    say("near edges", origin)
    x, y, z = _near_edge_cloud(40_000, seed=int(origin[0]) % 1000 + 7, side=200, origin=origin)
    box = synth.bbox(x, y)
    p = O.params(10, 20, 0.9, box, len(x), 16)
    tree = ctx.parallel_qtree(x, y, z, leaf_size=16)
    gm, _ = ctx.stage1(tree, 10, 0.9, 20)
    assert np.array_equal(gm, O.stage1_brute(x, y, z, p))
    tree.free()
    want, _, _ = O.run(x, y, z, 10, 20, 0.9, 16)
    assert np.array_equal(ctx.run_tile(x, y, z, 10, 20, 0.9, 16), want)


def test_stage3_many_crossed_leaves(ctx):
    """Stage 3 defers the leaves a cell's edge crosses to a wave-wide scan (256
    per cell at most).  With no LLPs every 60 m cell is queried, and its edges
    cross more than 256 two-point leaves at 5 points/m2, so the lanes past the
    list scan their leaves themselves; the result is the brute-force grid scan."""
    say("stage 3 crossed leaves")
    t = synth.scaled(synth.config_c2(seed=9), 2_400_000)
    x, y, z = t.generate()
    box = synth.bbox(x, y)
    p = O.params(12, 60, 0.8, box, len(x), 2)
    none = np.zeros(0, np.uint32)
    want = np.sort(O.stage3_brute(x, y, z, p, none))
    tree = ctx.parallel_qtree(x, y, z, leaf_size=2)
    got = ctx.stage3(tree, none, 12, 0.8, 60)
    tree.free()
    assert len(want) > 100 and np.array_equal(np.sort(got), want)


@pytest.mark.parametrize("summaries", ["on", "off"])
def test_c2_scaled_with_and_without_summaries(ctx, restore_table, summaries):
    _stage1_table(ctx, summaries)
    say("c2 scaled", summaries)
    t = synth.scaled(synth.config_c2(seed=3), 2_000_000)
    x, y, z = t.generate()
    want, _, _ = O.run(x, y, z, 12, 20, 0.8, 32)
    for leaf in (16, 32, 64):
        assert np.array_equal(ctx.run_tile(x, y, z, 12, 20, 0.8, leaf), want), leaf
