"""Builds the native libraries in-tree (hipcc for gfx950, gcc for host-only C).

    python -m fastowm_amd.build        # or __graft_entry__.build()

Outputs: fastowm_amd/libfastowm.so (HIP kernels + C ABI), fastowm_amd/libowm_synth.so
(synthetic tile generator), fastowm_amd/bin/owm (CLI).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
ARCH = os.environ.get("OWM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: window edges are computed without fused multiply-add, as
# the reference does; packed fp32 ops are disabled for the device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
             "-Wno-unused-function", "-Wno-unused-variable"]
HIP_SOURCES = ["sort.hip", "tree.hip", "stage1.hip", "stages.hip", "xyz.hip", "owm_host.cpp",
               "owm_capi.cpp", "owm_serve.cpp", "owm_tune.cpp", "owm_xyz.cpp", "owm_gold.cpp",
               "qstats.hip", "owm_qstats.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _compile(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    _run(cmd)
    return out


def build(verbose: bool = False) -> None:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.join(HERE, "bin"), exist_ok=True)
    jobs = min(len(HIP_SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES))
    lib = os.path.join(HERE, "libfastowm.so")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", lib + ".tmp"] + objs +
         ["-lpthread"])
    os.replace(lib + ".tmp", lib)
    synth = os.path.join(HERE, "libowm_synth.so")
    _run(["gcc", "-O2", "-fPIC", "-fopenmp", "-shared", "-o", synth + ".tmp",
          os.path.join(CSRC, "synth.c"), "-lm"])
    os.replace(synth + ".tmp", synth)
    cli_src = os.path.join(CSRC, "owm_cli.cpp")
    if os.path.exists(cli_src):
        cli = os.path.join(HERE, "bin", "owm")
        _run(["g++", "-O2", "-std=c++17", "-o", cli + ".tmp", cli_src, "-I", os.path.join(HERE, "..", "include"),
              "-L", HERE, "-lfastowm", "-Wl,-rpath,$ORIGIN/..", "-lpthread"])
        os.replace(cli + ".tmp", cli)
    if verbose:
        print("built", lib, synth)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
