"""The driver binary (fastowm_amd/bin/owm) end to end on the device.

Checks the reference driver's outputs (CPP-TBB-CPU/application/owm_optim3.cpp):
the seed file -o as its _salida.xyz (LLPs in ascending order, then the
stage-3 additions, "%.2f %.2f %.2f", :242-262) against the seed lists of the
reference's own quadtree OWM (tests/golden/ref_o1qtree.json), the results row
(save_time, src/optim3_func.cpp:797-828) with the flags it records, the GOLD
check's Correctness column, readXYZfile's refusal of header-less files
(src/optim3_func.cpp:681-686), and the statistics mode's reports."""
import json
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O
from fastowm_amd import owm, synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
OWM = os.path.join(ROOT, "fastowm_amd", "bin", "owm")
HEADER = "ExecName InputFile NumThreads MinRadius MaxNumber Level TreeTime OwmTime Correctness"


def say(*a):
    print("[case]", *a, flush=True)


def _fmt(x, y, z, ids):
    return ["%.2f %.2f %.2f" % (x[i], y[i], z[i]) for i in ids]


def _cases():
    gold = json.load(open(os.path.join(HERE, "golden", "ref_o1qtree.json")))
    return {c["name"]: c for c in gold["cases"]}


def _tile(c):
    t = synth.scaled(synth.config_c1(c["generator"].get("seed", 0)), c["n"])
    for k, v in c["generator"].items():
        setattr(t, k, v)
    return t.generate()


def _run(args, cwd):
    r = subprocess.run([OWM] + args, cwd=cwd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    return r


@pytest.mark.parametrize("name", ["clusters_50k_w12_b20_o08", "uniform_30k_w13_b21_o08",
                                  "c2style_60k_w16_b30_o075"])
def test_seed_file_and_results_row(tmp_path, name):
    c = _cases()[name]
    x, y, z = _tile(c)
    stem = str(tmp_path / "tile")
    synth.write_xyz(stem + ".xyz", x, y, z, c["bbox"])
    out, res = str(tmp_path / "seeds.xyz"), str(tmp_path / "times.csv")
    say("cli", name)
    r = _run(["-i", stem, "-W", str(c["wsize"]), "-B", str(c["bsize"]), "-O", str(c["overlap"]),
              "-s", "32", "-l", "2", "-n", "4", "-r", "0.25", "-L", "3", "--results", res,
              "-o", out], tmp_path)
    assert r.returncode == 0
    lines = open(out).read().splitlines()
    llp, seeds = c["ref_llp"], c["ref_seeds"]
    assert len(lines) == len(seeds)
    assert lines[:len(llp)] == _fmt(x, y, z, llp), "LLPs first, ascending"
    adds = sorted(set(seeds) - set(llp))
    assert sorted(lines[len(llp):]) == sorted(_fmt(x, y, z, adds)), "then the stage-3 additions"
    rows = open(res).read().splitlines()
    assert rows[0] == HEADER and len(rows) == 2
    f = rows[1].split(" ")
    assert f[0] == OWM and f[1] == stem + ".xyz"
    assert f[2:6] == ["4", "0.25", "32", "3"]
    assert all(len(v.split(".")[1]) == 6 and float(v) > 0 for v in f[6:8]), "times in %.6f"
    assert f[8] == "-1", "no GOLD file"


def test_gold_correctness_and_second_row(tmp_path):
    c = _cases()["uniform_30k_w7_b11_o05"]
    x, y, z = _tile(c)
    stem = str(tmp_path / "tile")
    synth.write_xyz(stem + ".xyz", x, y, z, c["bbox"])
    # a GOLD file as the reference's are written: the seeds, %.2f
    with open(stem + "_salidaGOLD.xyz", "w") as g:
        g.write("\n".join(_fmt(x, y, z, c["ref_seeds"])) + "\n")
    res = str(tmp_path / "times.csv")
    say("cli gold")
    for _ in range(2):
        r = _run(["-i", stem, "-W", str(c["wsize"]), "-B", str(c["bsize"]), "-O", str(c["overlap"]),
                  "--results", res], tmp_path)
        assert r.returncode == 0
    rows = open(res).read().splitlines()
    assert rows[0] == HEADER and len(rows) == 3  # one header, then a row per run
    for row in rows[1:]:
        f = row.split(" ")
        assert f[2:6] == ["1", "0.1", "32", "5"], "the reference's defaults"
        assert f[8] == "100"


def test_header_less_file_refused_unless_data_bbox(tmp_path):
    c = _cases()["uniform_20k_w12_b20_o08"]
    x, y, z = _tile(c)
    stem = str(tmp_path / "plain")
    np.savetxt(stem + ".xyz", np.stack([x, y, z], axis=1), fmt="%.17g")  # round-trip digits
    say("cli no header")
    r = _run(["-i", stem, "-W", "12", "-B", "20", "-O", "0.8"], tmp_path)
    assert r.returncode != 0 and "No header data!" in r.stdout
    say("cli --data-bbox")
    out = str(tmp_path / "s.xyz")
    r = _run(["-i", stem, "-W", "12", "-B", "20", "-O", "0.8", "--data-bbox", "-o", out], tmp_path)
    assert r.returncode == 0
    want, _, _ = O.run(x, y, z, 12, 20, 0.8, 32)
    assert sorted(open(out).read().splitlines()) == sorted(_fmt(x, y, z, want))


def test_stats_mode_reports(tmp_path):
    c = _cases()["clusters_50k_w12_b20_o08"]
    x, y, z = _tile(c)
    stem = str(tmp_path / "tile")
    synth.write_xyz(stem + ".xyz", x, y, z, c["bbox"])
    for mode, policy, kw in (("maxnum", owm.QT_MAXNUMBER, {"max_number": 16}),
                             ("radius", owm.QT_MINRADIUS, {"min_radius": 0.5})):
        say("cli stats", mode)
        res = str(tmp_path / f"h_{mode}.csv")
        args = ["--stats", mode, "-i", stem, "-L", "4", "--results", res]
        args += ["-s", "16"] if mode == "maxnum" else ["-r", "0.5"]
        assert _run(args, tmp_path).returncode == 0
        with owm.Context(0) as ctx:
            classes, summary = ctx.qtree_stats(x, y, policy=policy, level=4, bbox=c["bbox"], **kw)
        want = str(tmp_path / f"w_{mode}.csv")
        owm.write_qtree_stats(want, classes, summary, summary["params"], len(x))
        # the reports name the file they were written to
        assert open(res).read().replace(res, "<file>") == open(want).read().replace(want, "<file>")
