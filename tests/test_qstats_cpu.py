"""Quadtree leaf statistics (the reference's owm_stats): the oracle against
the reference's own quadtree code and against an independent formulation.

- tests/golden/ref_qstats.json holds leaf classes written by the reference's
  minRadius quadtree (insertPointF, CPP-TBB-CPU/src/optim1_qtree_func.cpp:250-274)
  through oracle/ref_qstats.cpp (tests/golden/make_qstats_golden.py); with
  level 0 the oracle's parallel_qtree restatement must give the same classes.
- the maxNumber policy (insertPointMaxNum, CPP-TBB-CPU/src/optim3_func.cpp:160-184)
  is sequential insertion with splits and re-insertion in the oracle; here it
  is recomputed level by level with numpy (a node is split iff it is one of
  the depth <= level nodes or holds more than maxNumber points), which must
  agree, together with the closed forms leaves = 1 + 3 * internal nodes and
  sum(points) = n.
"""
import json
import os
import struct
import subprocess
import sys
import tempfile
from collections import Counter

import numpy as np
import pytest

import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import points_digest, tile_for  # noqa: E402
from make_qstats_golden import cloud  # noqa: E402

GOLD = json.load(open(os.path.join(HERE, "golden", "ref_qstats.json")))
REF_BIN = os.path.join(HERE, "..", "oracle", "_ref", "ref_qstats")


def level_sync(x, y, box, policy, max_number=32, min_radius=0.1, level=5, max_depth=64):
    """Leaf classes computed depth by depth: every point is routed from its
    node to a child with the createQuadrants centres; a child is split when
    forced (depth <= level) or by the policy, else it is a leaf."""
    rx, ry = (box[1] - box[0]) / 2.0, (box[3] - box[2]) / 2.0
    r = np.float32(rx if rx >= ry else ry)
    ncx = np.array([box[0] + rx])
    ncy = np.array([box[2] + ry])
    node = np.zeros(len(x), np.int64)
    px, py = np.asarray(x, np.float64), np.asarray(y, np.float64)
    cls = Counter()
    for d in range(1, max_depth + 1):
        off = float(np.float32(r * np.float32(0.5)))  # radius * (+-0.5f), a float
        rd = np.float32(float(r) * 0.5)                 # the child's radius
        q = (px >= ncx[node]).astype(np.int64) * 2 + (py >= ncy[node]).astype(np.int64)
        child = 4 * node + q
        cnt = np.bincount(child, minlength=4 * len(ncx))
        if d <= level:
            split = np.ones(len(cnt), bool)
        elif policy == O.QT_MAXNUMBER:
            split = cnt > max_number
        else:
            split = (cnt >= 1) & (float(rd) * 0.5 > float(np.float32(min_radius)))
        for c, k in Counter(cnt[~split].tolist()).items():
            cls[(d, c)] += k
        if not split.any():
            return sorted((a, b, k) for (a, b), k in cls.items())
        if d == max_depth:
            return None
        parent = np.arange(len(cnt)) // 4
        qq = np.arange(len(cnt)) % 4
        sx = np.where(qq & 2, off, -off)
        sy = np.where(qq & 1, off, -off)
        newid = np.full(len(cnt), -1, np.int64)
        newid[split] = np.arange(split.sum())
        ncx = (ncx[parent] + sx)[split]
        ncy = (ncy[parent] + sy)[split]
        keep = split[child]
        px, py, node = px[keep], py[keep], newid[child[keep]]
        r = rd
    return None


@pytest.mark.parametrize("case", GOLD["cases"], ids=[c["name"] for c in GOLD["cases"]])
def test_oracle_matches_reference_quadtree(case):
    x, y, box = cloud(case["cloud"])
    assert points_digest(x, y, np.zeros_like(x)) == case["points_sha256"], "generator drifted"
    assert list(box) == case["bbox"]
    got = O.qtree_stats(x, y, box, O.QT_MINRADIUS, min_radius=case["min_radius"], level=0)
    assert [list(c) for c in got] == case["ref_classes"]
    # the numpy formulation agrees with the reference too
    assert [list(c) for c in level_sync(x, y, box, O.QT_MINRADIUS,
                                         min_radius=case["min_radius"], level=0)] == case["ref_classes"]


@pytest.mark.skipif(not os.path.exists(REF_BIN), reason="reference quadtree driver not built here")
def test_live_reference_quadtree():
    rng = np.random.default_rng(11)
    x = 700000.0 + np.round(rng.uniform(0, 300, 30000), 2)
    y = 4280000.0 + np.round(rng.uniform(0, 200, 30000), 2)
    box = (700000.0, 700300.0, 4280000.0, 4280200.0)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "pts.bin")
        with open(path, "wb") as f:
            f.write(struct.pack("<Q4df", len(x), *box, 0.25))
            f.write(x.tobytes())
            f.write(y.tobytes())
        out = subprocess.run([REF_BIN, path], capture_output=True, text=True, timeout=300,
                             check=True).stdout
    ref = [tuple(int(v) for v in line.split()) for line in out.splitlines()]
    assert O.qtree_stats(x, y, box, O.QT_MINRADIUS, min_radius=0.25, level=0) == ref


MAXNUM_CASES = [
    ("uniform_30k", {}, 30_000, 32, 5),
    ("uniform_30k_L0", {"seed": 2}, 30_000, 16, 0),
    ("clusters_40k", {"cluster_frac": 0.7, "n_clusters": 6, "cluster_sigma": 4.0}, 40_000, 8, 3),
    ("c2style_60k", {"cluster_frac": 0.1, "n_clusters": 5}, 60_000, 100, 2),
    ("uniform_5k_m1", {"seed": 7}, 5_000, 1, 1),
]


@pytest.mark.parametrize("name,over,n,maxnum,level", MAXNUM_CASES, ids=[c[0] for c in MAXNUM_CASES])
def test_maxnumber_oracle_equals_level_by_level(name, over, n, maxnum, level):
    x, y, _ = tile_for(over, n).generate()
    from fastowm_amd import synth
    box = synth.bbox(x, y)
    got = O.qtree_stats(x, y, box, O.QT_MAXNUMBER, max_number=maxnum, level=level)
    assert got == level_sync(x, y, box, O.QT_MAXNUMBER, max_number=maxnum, level=level)
    leaves = sum(k for _, _, k in got)
    assert sum(c * k for _, c, k in got) == n
    assert all(c <= maxnum for _, c, _ in got)
    assert (leaves - 1) % 3 == 0
    assert min(d for d, _, _ in got) >= level + 1


def test_lattice_boundaries_maxnumber():
    # points on every quadrant centre line go to the upper quadrant (>=)
    x, y, box = cloud({"kind": "lattice", "side": 32.0, "step": 0.5})
    for maxnum, level in ((1, 0), (4, 2), (9, 1)):
        got = O.qtree_stats(x, y, box, O.QT_MAXNUMBER, max_number=maxnum, level=level)
        assert got == level_sync(x, y, box, O.QT_MAXNUMBER, max_number=maxnum, level=level)


def test_edge_cases():
    box = (0.0, 10.0, 0.0, 10.0)
    e = np.zeros(0)
    # no points: the forced nodes' children are empty leaves
    assert O.qtree_stats(e, e, box, O.QT_MAXNUMBER, level=2) == [(3, 0, 64)]
    # more than maxNumber coincident points never stop splitting
    p = np.full(9, 5.0)
    assert O.qtree_stats(p, p, box, O.QT_MAXNUMBER, max_number=8, level=0) is None
    assert level_sync(p, p, box, O.QT_MAXNUMBER, max_number=8, level=0) is None
    got = O.qtree_stats(p, p, box, O.QT_MAXNUMBER, max_number=9, level=0)
    assert got == [(1, 0, 3), (1, 9, 1)]


# ---- the text reports (owm_qtree_stats_write, host only) ------------------

def c_round2d(v):
    import math
    w = v * 100.0
    return (math.floor(w + 0.5) if w >= 0 else math.ceil(w - 0.5)) / 100.0


def expected_reports(path, classes, policy, max_number, min_radius, box, n):
    """save_histogram, save_areas, save_density and save_levels
    (CPP-TBB-CPU/include/stats_func.hpp:120-366) over the leaves the classes
    describe; numbers as an ostream prints them (%g, 6 digits)."""
    rx, ry = (box[1] - box[0]) / 2.0, (box[3] - box[2]) / 2.0
    r0 = np.float32(rx if rx >= ry else ry)

    def radius(d):
        r = r0
        for _ in range(d):
            r = np.float32(float(r) * 0.5)
        return r

    pts, lev, rad, den = Counter(), Counter(), Counter(), Counter()
    for d, c, k in classes:
        r = radius(d)
        pts[c] += k
        lev[d] += k
        rad[float(r)] += k
        den[float(np.float32(c) / (np.float32(4) * r * r))] += k
    g = lambda v: "%g" % v  # noqa: E731
    nleaf = sum(pts.values())
    dens = n / (c_round2d(box[1] - box[0]) * c_round2d(box[3] - box[2]))
    s = [f"HISTOGRAM: {path}", f"Number of leaf nodes: {nleaf}"]
    s += [f"MinRadius: {g(float(np.float32(min_radius)))} m" if policy == O.QT_MINRADIUS
          else f"MaxNumber: {max_number} pts", ""]
    s += ["Observed: ", f"Average number of points/leaf-node: {sum(c * k for c, k in pts.items()) // nleaf}",
          f"__min: {min([999] + list(pts))}", f"__max: {max([0] + list(pts))}", "",
          "Estimation (assuming equal distribution of points): ", f"Average density: {g(dens)}", "",
          "x_i n_i x_i*n_i", ""]
    s += [f"{c} {k} {c * k}" for c, k in sorted(pts.items())] + ["", ""]
    s += ["Radius HISTOGRAM:", f"Number of leaf-nodes: {nleaf}", "", "r_i n_i a_i", ""]
    s += [f"{g(r)} {k} {g(4 * r * r)}" for r, k in sorted(rad.items())] + ["", ""]
    s += ["Density HISTOGRAM:", f"Number of leaf-nodes: {nleaf}", "", "d_i n_i", ""]
    s += [f"{g(v)} {k}" for v, k in sorted(den.items())] + ["", ""]
    s += ["Levels HISTOGRAM:", f"Number of leaf-nodes: {nleaf}", "", "level_i n_i", ""]
    s += [f"{d} {k}" for d, k in sorted(lev.items())] + ["", ""]
    return "\n".join(s) + "\n"


@pytest.mark.parametrize("policy,kw", [(O.QT_MAXNUMBER, {"max_number": 32}),
                                       (O.QT_MINRADIUS, {"min_radius": 0.3})])
def test_reports_text(tmp_path, policy, kw):
    from fastowm_amd import owm, synth
    x, y, _ = tile_for({"cluster_frac": 0.3, "n_clusters": 4}, 40_000).generate()
    box = synth.bbox(x, y)
    classes = O.qtree_stats(x, y, box, policy, level=3, **kw)
    rx, ry = (box[1] - box[0]) / 2.0, (box[3] - box[2]) / 2.0
    summary = {"bbox": box, "max_radius": float(np.float32(rx if rx >= ry else ry)),
               "depth": max(c[0] for c in classes), "leaves": sum(c[2] for c in classes)}
    q = owm.qtree_params(policy, level=3, bbox=box, **kw)
    path = tmp_path / "tbb_histogram.csv"
    owm.write_qtree_stats(path, classes, summary, q, len(x))
    want = expected_reports(str(path), classes, policy, kw.get("max_number", 32),
                            kw.get("min_radius", 0.1), box, len(x))
    assert path.read_text() == want


def test_leaves_below_depth_16():
    rng = np.random.default_rng(5)
    g = np.arange(12) * 1e-4
    X, Y = np.meshgrid(500.123 + g, 250.456 + g, indexing="ij")
    x = np.concatenate([rng.uniform(0, 1000, 2000), X.ravel()])
    y = np.concatenate([rng.uniform(0, 1000, 2000), Y.ravel()])
    box = (0.0, 1000.0, 0.0, 1000.0)
    for maxnum, level in ((4, 0), (1, 3)):
        got = O.qtree_stats(x, y, box, O.QT_MAXNUMBER, max_number=maxnum, level=level)
        assert got == level_sync(x, y, box, O.QT_MAXNUMBER, max_number=maxnum, level=level)
        assert max(d for d, _, _ in got) > 16
    got = O.qtree_stats(x, y, box, O.QT_MINRADIUS, min_radius=1e-5, level=2)
    assert got == level_sync(x, y, box, O.QT_MINRADIUS, min_radius=1e-5, level=2)


def test_non_finite_coordinates():
    # quadrantIdx compares with >=: NaN goes to quadrant 0 at every depth,
    # +inf to the upper and -inf to the lower half
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 100, 3000)
    y = rng.uniform(0, 100, 3000)
    x[:5] = np.nan
    y[5:9] = np.inf
    x[9:12] = -np.inf
    y[12] = np.nan
    box = (0.0, 100.0, 0.0, 100.0)
    for pol, kw in ((O.QT_MAXNUMBER, {"max_number": 8}), (O.QT_MINRADIUS, {"min_radius": 0.5})):
        got = O.qtree_stats(x, y, box, pol, level=2, **kw)
        assert got == level_sync(x, y, box, pol, level=2, **kw)
        assert sum(c * k for _, c, k in got) == len(x)


def test_histograms_match_the_written_reports(tmp_path):
    from fastowm_amd import owm, synth
    x, y, _ = tile_for({"seed": 6}, 30_000).generate()
    box = synth.bbox(x, y)
    classes = O.qtree_stats(x, y, box, O.QT_MINRADIUS, min_radius=0.2, level=4)
    rx, ry = (box[1] - box[0]) / 2.0, (box[3] - box[2]) / 2.0
    summary = {"bbox": box, "max_radius": float(np.float32(rx if rx >= ry else ry)),
               "depth": 0, "leaves": 0}
    h = owm.qtree_histograms(classes, summary)
    path = tmp_path / "h.csv"
    owm.write_qtree_stats(path, classes, summary, owm.qtree_params(O.QT_MINRADIUS, min_radius=0.2,
                                                                   level=4, bbox=box), len(x))
    # each report's rows, after its column header line and a blank line
    text = path.read_text().split("\n")
    tables = {}
    for head, name in (("x_i n_i x_i*n_i", "points"), ("r_i n_i a_i", "radius"),
                       ("d_i n_i", "density"), ("level_i n_i", "levels")):
        i = text.index(head) + 2
        rows = []
        while text[i]:
            rows.append(text[i].split())
            i += 1
        tables[name] = rows
    assert [[str(v), str(k)] for v, k in h["points"]] == [r[:2] for r in tables["points"]]
    assert [["%g" % v, str(k)] for v, k in h["radius"]] == [r[:2] for r in tables["radius"]]
    assert [["%g" % v, str(k)] for v, k in h["density"]] == tables["density"]
    assert [[str(v), str(k)] for v, k in h["levels"]] == tables["levels"]
    assert sum(k for _, k in h["levels"]) == sum(k for _, _, k in classes)
