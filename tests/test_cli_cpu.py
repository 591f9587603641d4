"""CPU checks of the driver's host-only pieces: the results row of save_time
(CPP-TBB-CPU/src/optim3_func.cpp:797-828) and readXYZfile's table of the
reference's header-less clouds (src/optim3_func.cpp:667-680).

The row is compared byte for byte with what a C++ program writes through the
same stream manipulators save_time uses (defaultfloat for MinRadius and
Correctness, fixed for the two times), so the %g / %.6f formatting of the
library is checked against the C++ library's own number formatting."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fastowm_amd", "libfastowm.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libfastowm.so not built")

HEADER = "ExecName InputFile NumThreads MinRadius MaxNumber Level TreeTime OwmTime Correctness\n"

# writes rows the way save_time streams them (one row per input line:
# exec input threads minRadius maxNumber level tree owm correctness)
STREAM_ROWS = r"""
#include <fstream>
#include <iostream>
#include <string>
int main(int argc, char **argv)
{
    std::fstream out(argv[1], std::ios::out | std::ios::app);
    std::string exec, input;
    int threads, maxNumber, level;
    float minRadius;
    double tree, owm, corr;
    while (std::cin >> exec >> input >> threads >> minRadius >> maxNumber >> level >> tree >> owm >> corr) {
        out << exec << " " << input << " " << threads << " ";
        out << std::defaultfloat << minRadius << " " << maxNumber << " " << level << " ";
        out << std::fixed << tree << " " << owm << " ";
        out << std::defaultfloat << corr << std::endl;
    }
    return 0;
}
"""


def _lib():
    L = ctypes.CDLL(LIB)
    L.owm_save_time.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double]
    L.owm_save_time.restype = ctypes.c_int
    L.owm_xyz_known_cloud.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.c_void_p]
    L.owm_xyz_known_cloud.restype = ctypes.c_int
    return L


def _rows():
    rng = np.random.default_rng(3)
    rows = [("./o3memo", "data/INAER_2011_AlcoyH.xyz", 8, 0.1, 32, 5, 0.123456789, 1.5, 97.123456789),
            ("bin/owm", "a.xyz", 1, 0.25, 64, 3, 0.0, 2e-6, -1.0),
            ("x", "b.xyz", 64, 1e-5, 1, 0, 12.3456785, 1234.5, 100.0),
            ("x", "c.xyz", 2, 3.0, 128, 10, 1e-7, 0.5, 99.99999949)]
    for _ in range(40):
        rows.append(("owm", "t.xyz", int(rng.integers(1, 97)), float(np.float32(rng.uniform(1e-3, 10))),
                     int(rng.integers(1, 512)), int(rng.integers(0, 11)), float(rng.uniform(0, 3)),
                     float(rng.exponential(0.05)), float(rng.uniform(0, 100))))
    return rows


def test_save_time_row_matches_the_stream_layout(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no C++ compiler")
    src = tmp_path / "rows.cpp"
    src.write_text(STREAM_ROWS)
    exe = tmp_path / "rows"
    subprocess.run([cxx, "-std=c++17", str(src), "-o", str(exe)], check=True)
    rows = _rows()
    want = tmp_path / "want.csv"
    want.write_text(HEADER)
    text = "\n".join(" ".join(repr(v) if isinstance(v, float) else str(v) for v in r) for r in rows)
    subprocess.run([str(exe), str(want)], input=text + "\n", text=True, check=True)
    got = tmp_path / "got.csv"
    L = _lib()
    for r in rows:
        st = L.owm_save_time(r[0].encode(), str(got).encode(), r[1].encode(), *r[2:])
        assert st == 0
    assert got.read_text() == want.read_text()
    # a few rows spelled out: the reference's defaults, no GOLD file (-1)
    lines = got.read_text().splitlines()
    assert lines[0] == HEADER.strip()
    assert lines[1] == "./o3memo data/INAER_2011_AlcoyH.xyz 8 0.1 32 5 0.123457 1.500000 97.1235"
    assert lines[2] == "bin/owm a.xyz 1 0.25 64 3 0.000000 0.000002 -1"
    assert lines[3] == "x b.xyz 64 1e-05 1 0 12.345678 1234.500000 100"


def test_save_time_header_only_for_a_new_file(tmp_path):
    L = _lib()
    f = tmp_path / "r.csv"
    f.write_text("")  # exists (empty): save_time tests existence, not size
    assert L.owm_save_time(b"e", str(f).encode(), b"i.xyz", 1, 0.1, 32, 5, 0.1, 0.2, -1.0) == 0
    assert f.read_text() == "e i.xyz 1 0.1 32 5 0.100000 0.200000 -1\n"
    assert L.owm_save_time(b"e", str(tmp_path / "nodir" / "r.csv").encode(), b"i", 1, 0.1, 32, 5,
                           0.0, 0.0, 0.0) == 1  # OWM_E_ARG: cannot be written


def test_known_header_less_clouds():
    L = _lib()
    n = ctypes.c_uint64()
    box = (ctypes.c_double * 4)()
    table = {
        "data/INAER_2011_Alcoy.xyz": (2772832, (715244.96, 716057.75, 4286623.63, 4287447.70)),
        "/x/INAER_2011_Alcoy_Core.xyz": (20380212, (714947.98, 716361.06, 4286501.93, 4288406.23)),
        "BABCOCK_2017_Arzua_3B.xyz": (40706503, (568000.00, 568999.99, 4752320.00, 4753319.99)),
        "V21_group1_densified_point_cloud.xyz": (42384876, (526964.093, 527664.647, 4742610.292, 4743115.738)),
        "d/V19_group1_densified_point_cloud.xyz": (48024480, (526955.908, 527686.445, 4742586.025, 4743124.373)),
        "sample24.xyz": (7492, (513748.12, 513869.97, 5403124.76, 5403197.20)),
    }
    for path, (cnt, bb) in table.items():
        assert L.owm_xyz_known_cloud(path.encode(), ctypes.byref(n), box) == 0, path
        assert n.value == cnt and tuple(box) == bb, path
    for path in ("data/AlcoyH.xyz", "tile.xyz", "INAER_2011_Alcoy.txt"):
        assert L.owm_xyz_known_cloud(path.encode(), ctypes.byref(n), box) == 8  # OWM_E_NOTFOUND
