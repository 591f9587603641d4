"""CPU-side checks of the C ABI: the HIP library loads, exports every entry point
include/owm.h declares, and its host-only parameter derivation equals the
oracle's (no device call is made here)."""
import ctypes
import os
import re

import numpy as np
import pytest

import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "owm.h")
LIB = os.path.join(ROOT, "fastowm_amd", "libfastowm.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(owm_[a-z0-9_]+)\s*\(", text)))


pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libfastowm.so not built")


def test_exports_every_declared_symbol():
    names = declared()
    assert len(names) >= 25
    lib = ctypes.CDLL(LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_is_plain_c():
    raw = open(HEADER).read()
    assert 'extern "C"' in raw
    text = re.sub(r"/\*.*?\*/", "", raw, flags=re.S)
    for bad in ("std::", "torch", "at::", "template", "class "):
        assert bad not in text


def test_status_strings():
    from fastowm_amd import owm
    L = owm.lib()
    assert L.owm_status_str(0) == b"ok"
    assert L.owm_status_str(5) == b"output buffer too small"


@pytest.mark.parametrize("ws,bs,ov,box,n", [
    (12, 20, 0.8, (500000.0, 500500.0, 4700000.0, 4700500.0), 1_000_000),
    (13, 21, 0.8, (512000.123, 512731.4, 4700001.7, 4700200.05), 300_000),
    (7, 11, 0.5, (0.0, 70.7, 0.0, 70.7), 20_000),
    (8, 15, 0.0, (10.0, 1010.0, 20.0, 520.0), 250_000),
    (20, 30, 0.9, (1.0, 2.0, 3.0, 4.0), 3),
])
def test_derive_matches_oracle(ws, bs, ov, box, n):
    from fastowm_amd import owm
    d = owm.derive(ws, bs, ov, n, box)
    p = O.params(ws, bs, ov, box, n)
    assert (d.width, d.height, d.density, d.displace) == (p.width, p.height, p.density, p.displace)
    assert (d.ncols, d.nrows, d.ncolsg, d.nrowsg) == (p.ncols, p.nrows, p.ncolsg, p.nrowsg)
    assert d.min_num_points == p.min_num_points


@pytest.mark.parametrize("ws,ov,box", [
    (12, 1.0, (0.0, 10.0, 0.0, 10.0)),
    (0, 0.8, (0.0, 10.0, 0.0, 10.0)),
    (12, 0.8, (0.0, 0.001, 0.0, 10.0)),
    (12, -0.1, (0.0, 10.0, 0.0, 10.0)),
    (70000, 0.8, (0.0, 10.0, 0.0, 10.0)),
])
def test_derive_rejects_like_oracle(ws, ov, box):
    from fastowm_amd import owm
    with pytest.raises(owm.OWMError) as e:
        owm.derive(ws, 20, ov, 100, box)
    assert e.value.name == "E_ARG"
    assert O.params(ws, 20, ov, box, 100) is None


@pytest.mark.parametrize("ws,bs,ov,n,seed", [(12, 20, 0.8, 20_000, 1), (8, 15, 0.0, 30_000, 2),
                                             (13, 21, 0.5, 5_000, 3), (5, 10, 0.9, 200, 4)])
def test_seed_count_bounds_oracle(ws, bs, ov, n, seed):
    """The sizing bound is min(n, windows + cells) and holds the oracle's seeds."""
    from fastowm_amd import owm, synth
    t = synth.scaled(synth.config_c1(seed), n)
    x, y, z = t.generate()
    box = synth.bbox(x, y)
    p = O.params(ws, bs, ov, box, n)
    bound = owm.query_seed_count(ws, bs, ov, n, box)
    assert bound == min(n, p.ncols * p.nrows + p.ncolsg * p.nrowsg)
    seeds, _, _ = O.run(x, y, z, ws, bs, ov, 32)
    assert len(seeds) <= bound
    assert owm.query_seed_count(ws, bs, ov, n) == n  # no bbox: n


def test_python_api_has_reference_names():
    from fastowm_amd import owm
    for name in ("parallel_qtree", "stage1", "stage2", "stage3", "run_tile", "batch_run"):
        assert hasattr(owm.Context, name)


def _gold_brute(seeds, gold, displace):
    """check_results of the reference, restated literally: every seed against
    every GOLD point (src/optim3_func.cpp:725-794)."""
    r = float(np.float32(np.float32(displace) * 0.5))
    sx, sy, sz = seeds
    gx, gy, gz = gold
    count = 0
    for x, y, z in zip(sx, sy, sz):
        inside = (gx > x - r) & (gy > y - r) & (gx < x + r) & (gy < y + r)
        if np.any(inside & (np.abs(z - gz) < 0.01)):
            count += 1
    return count, count / len(gx) * 100.0


@pytest.mark.parametrize("seed,displace", [(0, 2.4), (1, 1.2), (2, 3.0), (3, 0.6)])
def test_gold_check_equals_the_reference_rule(seed, displace):
    from fastowm_amd import owm
    rng = np.random.default_rng(seed)
    n = 3000
    gx = 5e5 + rng.uniform(0, 200, n)
    gy = 4.7e6 + rng.uniform(0, 200, n)
    gz = rng.uniform(0, 5, n).round(2)
    # seeds: some near GOLD points (inside / on / past the box edge, z within
    # and past the tolerance), some anywhere
    r = float(np.float32(np.float32(displace) * 0.5))
    k = rng.integers(0, n, 1500)
    dx = rng.choice([0.0, 0.3 * r, r * (1 - 1e-12), r, r * (1 + 1e-9), -0.7 * r, -r], 1500)
    dz = rng.choice([0.0, 0.005, 0.00999, 0.01, 0.02, -0.004], 1500)
    sx = np.concatenate([gx[k] + dx, 5e5 + rng.uniform(0, 200, 500)])
    sy = np.concatenate([gy[k] - dx[::-1], 4.7e6 + rng.uniform(0, 200, 500)])
    sz = np.concatenate([gz[k] + dz, rng.uniform(0, 5, 500)])
    want = _gold_brute((sx, sy, sz), (gx, gy, gz), displace)
    got = owm.gold_check((sx, sy, sz), (gx, gy, gz), displace)
    assert got[0] == want[0] and got[1] == pytest.approx(want[1], rel=0, abs=1e-12)
    assert 0 < got[0] < len(sx)


C_CALLER = r"""
#include <stdio.h>
#include <string.h>
#include "owm.h"

int main(int argc, char **argv)
{
    /* the driver's parameter derivation, as a C caller sizes its buffers */
    owm_params p;
    memset(&p, 0, sizeof p);
    p.wsize = 12; p.bsize = 20; p.overlap = 0.8; p.leaf_size = 32;
    double box[4] = {0.0, 500.0, 0.0, 500.0};
    owm_derived d;
    if (owm_derive(&p, 1000000, box, &d) != OWM_OK) return 1;
    printf("%u %u %u %u %u\n", d.min_num_points, d.ncols, d.nrows, d.ncolsg, d.nrowsg);
    /* the statistics tool's reports from a histogram: two leaves of 3 points
       at depth 1, two empty ones */
    owm_qtree_params q;
    memset(&q, 0, sizeof q);
    q.policy = OWM_QT_MAXNUMBER; q.max_number = 4; q.level = 0;
    owm_qtree_summary s;
    memset(&s, 0, sizeof s);
    memcpy(s.bbox, box, sizeof box);
    s.max_radius = 250.0f; s.depth = 1; s.leaves = 4; s.n_bins = 2;
    owm_leaf_bin bins[2] = {{1, 0, 2}, {1, 3, 2}};
    if (owm_qtree_stats_write(argv[1], &q, &s, bins, 2, 6) != OWM_OK) return 2;
    printf("%s\n", owm_status_str(OWM_E_CAPACITY));
    /* argument guards come before any use of the handles: a missing output
       buffer with a nonzero capacity is OWM_E_ARG (the handles here are
       stand-ins that must never be dereferenced) */
    static char fake_ctx[64], fake_tree[64];
    owm_ctx *c = (owm_ctx *)fake_ctx;
    owm_tree *t = (owm_tree *)fake_tree;
    uint64_t cnt = 0;
    if (owm_stage1(c, t, &p, NULL, 16, NULL) != OWM_E_ARG) return 3;
    if (owm_stage2(c, t, &p, NULL, 4, NULL, 4, &cnt) != OWM_E_ARG) return 4;
    if (owm_stage3(c, t, &p, NULL, 0, NULL, 4, &cnt) != OWM_E_ARG) return 5;
    if (owm_tree_run(c, t, &p, NULL, 4, &cnt) != OWM_E_ARG) return 6;
    if (owm_run_tile(c, NULL, NULL, NULL, 10, &p, NULL, 0, &cnt) != OWM_E_ARG) return 7;
    if (owm_stage1(NULL, t, &p, NULL, 0, NULL) != OWM_E_ARG) return 8;
    return 0;
}
"""


def test_plain_c_caller_builds_and_runs(tmp_path):
    """The header compiles as C and a C program links the library and calls
    its host-only entry points."""
    import shutil
    import subprocess
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "caller.c"
    src.write_text(C_CALLER)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(LIB)
    subprocess.run([cc, "-std=c99", "-Wall", "-Werror", str(src), "-I", os.path.join(ROOT, "include"),
                    "-L", libdir, "-lfastowm", "-Wl,-rpath," + libdir, "-o", str(exe)], check=True)
    report = tmp_path / "hist.csv"
    out = subprocess.run([str(exe), str(report)], check=True, capture_output=True, text=True).stdout
    ref = O.params(12, 20, 0.8, (0.0, 500.0, 0.0, 500.0), 1000000)
    assert out.split("\n")[0].split() == [str(v) for v in (ref.min_num_points, ref.ncols, ref.nrows,
                                                            ref.ncolsg, ref.nrowsg)]
    text = report.read_text()
    assert "Number of leaf nodes: 4" in text and "MaxNumber: 4 pts" in text
    assert "\n0 2 0\n3 2 6\n" in text  # x_i n_i x_i*n_i rows
    assert "\n1 4\n" in text           # level 1: four leaves
