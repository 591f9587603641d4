"""Writes tests/golden/ref_o1qtree.json: seed lists produced by the reference's own
OpenMP quadtree OWM (CPP-TBB-CPU/application/owm_optim1_qtree.cpp with
src/optim1_qtree_func.cpp), compiled by `make -C oracle ref` into
oracle/_ref/o1qtree with -DDEBUG.

For each case the tile is regenerated from its synthetic spec, written as a
headed .xyz file, and the binary is run as
    o1qtree <file-without-.xyz> Wsize Bsize Overlap threads 1 0.1
in a scratch directory.  LLP ids come from its LLPs.txt (0-based ids); the full
seed list comes from <file>_salida.xyz (LLPs, then stage-3 additions, printed
with %.2f), whose lines are matched back to point ids.

The reference exact stage 1 tests every point against the open window box,
so on tie-free data (continuous z) its seed set is a function of the input and
must equal the oracle's.  Run from the repo root:
    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from fastowm_amd import synth  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "o1qtree")
OUT = os.path.join(ROOT, "tests", "golden", "ref_o1qtree.json")

# name, generator overrides, n, wsize, bsize, overlap
CASES = [
    ("uniform_20k_w12_b20_o08", {}, 20_000, 12, 20, 0.8),
    ("clusters_50k_w12_b20_o08", {"cluster_frac": 0.7, "n_clusters": 6, "cluster_sigma": 8.0},
     50_000, 12, 20, 0.8),
    ("uniform_30k_w13_b21_o08", {}, 30_000, 13, 21, 0.8),
    ("uniform_30k_w7_b11_o05", {"seed": 4}, 30_000, 7, 11, 0.5),
    ("uniform_30k_w10_b20_o09", {"seed": 5}, 30_000, 10, 20, 0.9),
    ("sparse_20k_w8_b15_o00", {"seed": 6, "cluster_frac": 0.5, "n_clusters": 4}, 20_000, 8, 15, 0.0),
    ("c2style_100k_w12_b20_o08", {"cluster_frac": 0.1, "n_clusters": 5}, 100_000, 12, 20, 0.8),
    ("c2style_60k_w16_b30_o075", {"seed": 9, "cluster_frac": 0.3, "n_clusters": 8}, 60_000, 16, 30, 0.75),
]


def tile_for(over, n):
    t = synth.scaled(synth.config_c1(over.get("seed", 0)), n)
    for k, v in over.items():
        setattr(t, k, v)
    return t


def points_digest(x, y, z):
    h = hashlib.sha256()
    for a in (x, y, z):
        h.update(np.ascontiguousarray(a, np.float64).tobytes())
    return h.hexdigest()


def match_ids(x, y, z, rows):
    """Point ids of %.2f-printed (x, y, z) rows (continuous data: unique)."""
    order = np.argsort(x, kind="stable")
    xs = x[order]
    ids = []
    for X, Y, Z in rows:
        lo = np.searchsorted(xs, X - 0.00501)
        hi = np.searchsorted(xs, X + 0.00501)
        cand = order[lo:hi]
        cand = cand[(np.abs(y[cand] - Y) < 0.00501) & (np.abs(z[cand] - Z) < 0.00501)]
        if len(cand) != 1:
            raise RuntimeError(f"ambiguous or missing match for {X} {Y} {Z}: {cand}")
        ids.append(int(cand[0]))
    return ids


def run_case(name, over, n, ws, bs, ov):
    t = tile_for(over, n)
    x, y, z = t.generate()
    box = synth.bbox(x, y)
    with tempfile.TemporaryDirectory() as d:
        stem = os.path.join(d, "tile")
        synth.write_xyz(stem + ".xyz", x, y, z, box)
        r = subprocess.run([REF_BIN, stem, str(ws), str(bs), str(ov), "4", "1", "0.1"], cwd=d,
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        llp = None
        if ov != 0:
            llp = sorted(int(v) for v in np.loadtxt(os.path.join(d, "LLPs.txt"), usecols=0,
                                                     dtype=np.int64, ndmin=1))
        rows = np.loadtxt(stem + "_salida.xyz", ndmin=2)
    seeds = match_ids(x, y, z, rows)
    if len(set(seeds)) != len(seeds):
        raise RuntimeError("duplicate seeds in reference output")
    return {
        "name": name, "generator": {k: v for k, v in over.items()}, "n": n,
        "wsize": ws, "bsize": bs, "overlap": ov, "bbox": list(box),
        "points_sha256": points_digest(x, y, z),
        "ref_llp": llp, "ref_seeds": sorted(seeds),
    }


def main():
    if not os.path.exists(REF_BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = {"source": "oracle/_ref/o1qtree built from CPP-TBB-CPU/application/owm_optim1_qtree.cpp "
                     "+ src/optim1_qtree_func.cpp (-DDEBUG)",
           "cases": [run_case(*c) for c in CASES]}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    for c in out["cases"]:
        print(c["name"], len(c["ref_seeds"]), "seeds", "" if c["ref_llp"] is None else len(c["ref_llp"]))


if __name__ == "__main__":
    main()
