"""Quadtree leaf statistics on the GPU (owm_qtree_stats) against the oracle's
restatement of parallel_qtree and against the reference's own quadtree code
(tests/golden/ref_qstats.json, minRadius policy, level 0): the (depth,
points, leaves) classes must be identical, for both node policies, forced
levels, points on quadrant centre lines, leaves beyond the exact count bins,
host and device inputs, a parsed cloud, and the error cases."""
import json
import os
import sys

import numpy as np
import pytest

import pyoracle as O
from fastowm_amd import owm, synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import tile_for  # noqa: E402
from make_qstats_golden import cloud  # noqa: E402

GOLD = json.load(open(os.path.join(HERE, "golden", "ref_qstats.json")))


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = owm.Context()
    yield c
    c.close()


def both(ctx, x, y, box, policy, **kw):
    got, sm = ctx.qtree_stats(x, y, policy=policy, bbox=box, **kw)
    want = O.qtree_stats(x, y, box, policy, **kw)
    assert want is not None
    assert got == want
    assert sm["leaves"] == sum(k for _, _, k in want)
    assert sm["depth"] == max(d for d, _, _ in want)
    return got


@pytest.mark.parametrize("case", GOLD["cases"], ids=[c["name"] for c in GOLD["cases"]])
def test_reference_quadtree_fixtures(ctx, case):
    say("ref fixture", case["name"])
    x, y, box = cloud(case["cloud"])
    got, _ = ctx.qtree_stats(x, y, policy=owm.QT_MINRADIUS, min_radius=case["min_radius"],
                             level=0, bbox=box)
    assert [list(c) for c in got] == case["ref_classes"]


MAXNUM = [
    ("uniform_200k_m32_L5", {}, 200_000, 32, 5),
    ("uniform_100k_m1_L0", {"seed": 2}, 100_000, 1, 0),
    ("clusters_150k_m8_L3", {"cluster_frac": 0.7, "n_clusters": 6, "cluster_sigma": 3.0}, 150_000, 8, 3),
    ("c2style_300k_m100_L2", {"cluster_frac": 0.1, "n_clusters": 5}, 300_000, 100, 2),
    ("c2style_300k_m2000_L1", {"cluster_frac": 0.1, "n_clusters": 5, "seed": 4}, 300_000, 2000, 1),
]


@pytest.mark.parametrize("name,over,n,maxnum,level", MAXNUM, ids=[c[0] for c in MAXNUM])
def test_maxnumber_vs_oracle(ctx, name, over, n, maxnum, level):
    say("maxnumber", name)
    x, y, _ = tile_for(over, n).generate()
    got = both(ctx, x, y, synth.bbox(x, y), owm.QT_MAXNUMBER, max_number=maxnum, level=level)
    if maxnum >= 1024:
        assert any(c >= 1024 for _, c, _ in got)  # leaves past the exact bins


MINRAD = [
    ("uniform_200k_r01_L5", {}, 200_000, 0.1, 5),
    ("clusters_150k_r05_L2", {"cluster_frac": 0.7, "n_clusters": 6, "cluster_sigma": 3.0}, 150_000, 0.5, 2),
    ("c2style_300k_r8_L0", {"cluster_frac": 0.1, "n_clusters": 5}, 300_000, 8.0, 0),
]


@pytest.mark.parametrize("name,over,n,minr,level", MINRAD, ids=[c[0] for c in MINRAD])
def test_minradius_vs_oracle(ctx, name, over, n, minr, level):
    say("minradius", name)
    x, y, _ = tile_for(over, n).generate()
    both(ctx, x, y, synth.bbox(x, y), owm.QT_MINRADIUS, min_radius=minr, level=level)


def test_lattice_centre_lines(ctx):
    say("lattice")
    for spec in ({"kind": "lattice", "side": 64.0, "step": 0.25},
                 {"kind": "lattice", "side": 50.0, "side_y": 30.0, "step": 0.125,
                  "origin": [512000.0, 4700000.0]}):
        x, y, box = cloud(spec)
        for maxnum, level in ((1, 0), (4, 2), (9, 1), (64, 4)):
            both(ctx, x, y, box, owm.QT_MAXNUMBER, max_number=maxnum, level=level)
        both(ctx, x, y, box, owm.QT_MINRADIUS, min_radius=0.2, level=3)


def test_bbox_from_data_and_device_inputs(ctx):
    import torch
    say("data bbox + device inputs")
    x, y, _ = tile_for({"seed": 8}, 120_000).generate()
    box = synth.bbox(x, y)
    want = O.qtree_stats(x, y, box, O.QT_MAXNUMBER, max_number=16, level=4)
    got, sm = ctx.qtree_stats(x, y, policy="statsmaxnum", max_number=16, level=4)
    assert got == want and tuple(sm["bbox"]) == tuple(box)
    tx = torch.from_numpy(x).cuda()
    ty = torch.from_numpy(y).cuda()
    got, _ = ctx.qtree_stats(tx, ty, policy="statsmaxnum", max_number=16, level=4, bbox=box)
    assert got == want


def test_parsed_cloud_with_header(ctx):
    say("parsed cloud")
    x, y, z = tile_for({"seed": 9}, 50_000).generate()
    box = (float(np.floor(x.min())), float(np.ceil(x.max())),
           float(np.floor(y.min())), float(np.ceil(y.max())))
    text = b"%d\n%.2f\n%.2f\n%.2f\n%.2f\n" % (len(x), *box) + synth.format_xyz(x, y, z)
    c = ctx.parse_xyz(text)
    px, py, _ = c.to_numpy()
    got, sm = ctx.qtree_stats(c, policy="stats", min_radius=0.25, level=5)
    assert tuple(sm["bbox"]) == box
    assert got == O.qtree_stats(px, py, box, O.QT_MINRADIUS, min_radius=0.25, level=5)


def test_repeated_calls_and_sizes(ctx):
    say("repeat")
    x, y, _ = tile_for({"seed": 10}, 200_000).generate()
    box = synth.bbox(x, y)
    a = both(ctx, x, y, box, owm.QT_MAXNUMBER, max_number=32, level=5)
    both(ctx, x[:5000], y[:5000], box, owm.QT_MAXNUMBER, max_number=32, level=5)
    assert both(ctx, x, y, box, owm.QT_MAXNUMBER, max_number=32, level=5) == a


def test_edge_cases(ctx):
    say("edges")
    box = (0.0, 10.0, 0.0, 10.0)
    e = np.zeros(0)
    got, _ = ctx.qtree_stats(e, e, policy=owm.QT_MAXNUMBER, level=2, bbox=box)
    assert got == [(3, 0, 64)]
    one = np.array([3.0])
    both(ctx, one, one, box, owm.QT_MAXNUMBER, max_number=1, level=0)
    p = np.full(9, 5.0)
    assert both(ctx, p, p, box, owm.QT_MAXNUMBER, max_number=9, level=0) == [(1, 0, 3), (1, 9, 1)]
    with pytest.raises(owm.OWMError):  # 9 coincident points, maxNumber 8: never stops
        ctx.qtree_stats(p, p, policy=owm.QT_MAXNUMBER, max_number=8, level=0, bbox=box)
    with pytest.raises(owm.OWMError):
        ctx.qtree_stats(p, p, policy=owm.QT_MAXNUMBER, max_number=0, bbox=box)
    with pytest.raises(owm.OWMError):
        ctx.qtree_stats(p, p, policy=owm.QT_MAXNUMBER, level=11, bbox=box)
    with pytest.raises(owm.OWMError):  # no box and no points
        ctx.qtree_stats(e, e, policy=owm.QT_MAXNUMBER)
    for r in (0.0, -1.0, 1e-30):  # occupied nodes would split past depth 64
        with pytest.raises(owm.OWMError) as err:
            ctx.qtree_stats(p, p, policy=owm.QT_MINRADIUS, min_radius=r, level=0, bbox=box)
        assert err.value.name == "E_ARG" and "minRadius" in str(err.value)
    # the context still works after the errors
    x, y, _ = tile_for({"seed": 12}, 30_000).generate()
    both(ctx, x, y, synth.bbox(x, y), owm.QT_MINRADIUS, min_radius=0.5, level=1)


def deep_cloud():
    """2000 points over a 1 km box and a 12 x 12 lattice 0.1 mm apart: its
    nodes stay split below depth 16 (nodes 1.5 cm wide there)."""
    rng = np.random.default_rng(5)
    g = np.arange(12) * 1e-4
    X, Y = np.meshgrid(500.123 + g, 250.456 + g, indexing="ij")
    x = np.concatenate([rng.uniform(0, 1000, 2000), X.ravel()])
    y = np.concatenate([rng.uniform(0, 1000, 2000), Y.ravel()])
    return x, y, (0.0, 1000.0, 0.0, 1000.0)


def test_leaves_below_the_key_depths(ctx):
    say("deep")
    x, y, box = deep_cloud()
    for maxnum, level in ((4, 0), (1, 3), (20, 5)):
        got = both(ctx, x, y, box, owm.QT_MAXNUMBER, max_number=maxnum, level=level)
        assert max(d for d, _, _ in got) > 16
    got = both(ctx, x, y, box, owm.QT_MINRADIUS, min_radius=1e-5, level=2)
    assert max(d for d, _, _ in got) > 16


def test_non_finite_coordinates(ctx):
    say("non-finite")
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 100, 3000)
    y = rng.uniform(0, 100, 3000)
    x[:5] = np.nan
    y[5:9] = np.inf
    x[9:12] = -np.inf
    y[12] = np.nan
    box = (0.0, 100.0, 0.0, 100.0)
    both(ctx, x, y, box, owm.QT_MAXNUMBER, max_number=8, level=2)
    both(ctx, x, y, box, owm.QT_MINRADIUS, min_radius=0.5, level=2)


def test_full_size_invariants(ctx):
    """C2 (20M points): properties that hold at any size -- every point in one
    leaf, leaves = 1 + 3 x split nodes, counts within maxNumber, no leaf above
    the forced depths -- and the same classes for the points in another order."""
    import torch
    say("c2 invariants")
    x, y, _ = synth.config_c2(0).generate()
    box = synth.bbox(x, y)
    tx, ty = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    got, sm = ctx.qtree_stats(tx, ty, policy=owm.QT_MAXNUMBER, max_number=32, level=5, bbox=box)
    leaves = sum(k for _, _, k in got)
    assert sum(c * k for _, c, k in got) == len(x)
    assert (leaves - 1) % 3 == 0 and leaves == sm["leaves"]
    assert all(c <= 32 for _, c, _ in got) and min(d for d, _, _ in got) >= 6
    perm = torch.randperm(len(x), generator=torch.Generator().manual_seed(1)).cuda()
    again, _ = ctx.qtree_stats(tx[perm].contiguous(), ty[perm].contiguous(),
                               policy=owm.QT_MAXNUMBER, max_number=32, level=5, bbox=box)
    assert again == got
    got, _ = ctx.qtree_stats(tx, ty, policy=owm.QT_MINRADIUS, min_radius=0.1, level=5, bbox=box)
    assert sum(c * k for _, c, k in got) == len(x)
    # minRadius 0.1 on a 2 km box: every occupied node splits down to depth 13
    assert {d for d, c, _ in got if c > 0} == {13}
