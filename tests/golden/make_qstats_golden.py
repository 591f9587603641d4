"""Writes tests/golden/ref_qstats.json: quadtree leaf classes produced by the
reference's own minRadius quadtree code (CPP-TBB-CPU/src/optim1_qtree_func.cpp,
insertPointF and createQuadrantsF) through oracle/ref_qstats.cpp, which
`make -C oracle ref` compiles into oracle/_ref/ref_qstats.

Each case is a cloud (a synthetic tile, or a lattice whose points lie exactly
on quadrant centre lines) with its box and minRadius; the fixture holds the
distinct (depth, points, leaves) triples of the leaves.  Run from the repo root:
    python tests/golden/make_qstats_golden.py
"""
import json
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from fastowm_amd import synth  # noqa: E402
from make_golden import points_digest, tile_for  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_qstats")
OUT = os.path.join(ROOT, "tests", "golden", "ref_qstats.json")

# name, cloud spec, minRadius
CASES = [
    ("uniform_20k_r01", {"kind": "tile", "over": {}, "n": 20_000}, 0.1),
    ("uniform_50k_r05", {"kind": "tile", "over": {"seed": 3}, "n": 50_000}, 0.5),
    ("clusters_50k_r02", {"kind": "tile", "over": {"cluster_frac": 0.7, "n_clusters": 6,
                                                   "cluster_sigma": 8.0}, "n": 50_000}, 0.2),
    ("c2style_100k_r1", {"kind": "tile", "over": {"cluster_frac": 0.1, "n_clusters": 5},
                         "n": 100_000}, 1.0),
    ("lattice_64_step05_r03", {"kind": "lattice", "side": 64.0, "step": 0.5}, 0.3),
    ("lattice_64_step1_r1", {"kind": "lattice", "side": 64.0, "step": 1.0}, 1.0),
    ("lattice_50x30_step025_r015", {"kind": "lattice", "side": 50.0, "side_y": 30.0,
                                    "step": 0.25, "origin": [512000.0, 4700000.0]}, 0.15),
]


def cloud(spec):
    """(x, y, box) of a case: a synthetic tile with its data box, or a lattice
    of points step apart from the box corner, boundary lines included."""
    if spec["kind"] == "tile":
        x, y, _ = tile_for(spec["over"], spec["n"]).generate()
        return x, y, synth.bbox(x, y)
    sx = spec["side"]
    sy = spec.get("side_y", sx)
    ox, oy = spec.get("origin", [0.0, 0.0])
    gx = ox + np.arange(0.0, sx + spec["step"] / 2, spec["step"])
    gy = oy + np.arange(0.0, sy + spec["step"] / 2, spec["step"])
    X, Y = np.meshgrid(gx, gy, indexing="ij")
    return X.ravel().copy(), Y.ravel().copy(), (ox, ox + sx, oy, oy + sy)


def run_case(name, spec, min_radius):
    x, y, box = cloud(spec)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "pts.bin")
        with open(path, "wb") as f:
            f.write(struct.pack("<Q4df", len(x), *box, min_radius))
            f.write(np.ascontiguousarray(x, np.float64).tobytes())
            f.write(np.ascontiguousarray(y, np.float64).tobytes())
        r = subprocess.run([REF_BIN, path], capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-2000:])
    classes = [[int(v) for v in line.split()] for line in r.stdout.splitlines() if line.strip()]
    return {"name": name, "cloud": spec, "n": len(x), "bbox": list(box), "min_radius": min_radius,
            "points_sha256": points_digest(x, y, np.zeros_like(x)), "ref_classes": classes}


def main():
    if not os.path.exists(REF_BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = {"source": "oracle/_ref/ref_qstats: oracle/ref_qstats.cpp + "
                     "CPP-TBB-CPU/src/optim1_qtree_func.cpp (insertPointF, minRadius policy)",
           "cases": [run_case(*c) for c in CASES]}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    for c in out["cases"]:
        print(c["name"], c["n"], "points", len(c["ref_classes"]), "classes",
              sum(k[2] for k in c["ref_classes"]), "leaves")


if __name__ == "__main__":
    main()
