"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Importable only from tests/, __graft_entry__.smoke() and the cpu_baseline leg
of bench.py, as the checker or the timed CPU baseline; the product package
never imports it.  Semantics: see owm_oracle.h.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NONE = 0xFFFFFFFF
_LIB = None


class Params(ctypes.Structure):
    _fields_ = [
        ("wsize", ctypes.c_uint32), ("bsize", ctypes.c_uint32), ("overlap", ctypes.c_double),
        ("leaf_size", ctypes.c_uint32), ("level", ctypes.c_int32),
        ("minx", ctypes.c_double), ("maxx", ctypes.c_double),
        ("miny", ctypes.c_double), ("maxy", ctypes.c_double),
        ("width", ctypes.c_double), ("height", ctypes.c_double),
        ("density", ctypes.c_double), ("displace", ctypes.c_double),
        ("min_num_points", ctypes.c_uint32),
        ("ncols", ctypes.c_uint32), ("nrows", ctypes.c_uint32),
        ("ncolsg", ctypes.c_uint32), ("nrowsg", ctypes.c_uint32),
        ("h", ctypes.c_int32), ("hb", ctypes.c_int32),
        ("initx", ctypes.c_double), ("inity", ctypes.c_double),
        ("initxg", ctypes.c_double), ("inityg", ctypes.c_double),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        up = ctypes.POINTER(ctypes.c_uint32)
        pp = ctypes.POINTER(Params)
        L.oracle_derive.argtypes = [pp, ctypes.c_uint64]
        L.oracle_derive.restype = ctypes.c_int
        L.oracle_tree_build.argtypes = [dp, dp, dp, ctypes.c_uint64, pp]
        L.oracle_tree_build.restype = ctypes.c_void_p
        L.oracle_tree_free.argtypes = [ctypes.c_void_p]
        L.oracle_tree_nodes.argtypes = [ctypes.c_void_p]
        L.oracle_tree_nodes.restype = ctypes.c_uint64
        L.oracle_tree_check.argtypes = [ctypes.c_void_p]
        L.oracle_tree_check.restype = ctypes.c_int
        L.oracle_stage1.argtypes = [ctypes.c_void_p, pp, up]
        L.oracle_stage1_cols.argtypes = [ctypes.c_void_p, pp, ctypes.c_uint32, ctypes.c_uint32, up]
        L.oracle_stage2.argtypes = [pp, up, ctypes.c_uint32, up]
        L.oracle_stage2.restype = ctypes.c_uint32
        L.oracle_stage3.argtypes = [ctypes.c_void_p, pp, up, ctypes.c_uint32, up]
        L.oracle_stage3.restype = ctypes.c_uint32
        L.oracle_run.argtypes = [dp, dp, dp, ctypes.c_uint64, pp, up, ctypes.c_uint64,
                                 ctypes.POINTER(ctypes.c_uint64), dp]
        L.oracle_run.restype = ctypes.c_int
        L.oracle_stage1_brute.argtypes = [dp, dp, dp, ctypes.c_uint64, pp, up]
        L.oracle_stage3_brute.argtypes = [dp, dp, dp, ctypes.c_uint64, pp, up, ctypes.c_uint32, up]
        L.oracle_stage3_brute.restype = ctypes.c_uint32
        L.oracle_parse_xyz.argtypes = [ctypes.c_char_p, ctypes.c_uint64, dp, dp, dp, ctypes.c_uint64]
        L.oracle_parse_xyz.restype = ctypes.c_int64
        L.oracle_qtree_stats.argtypes = [dp, dp, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_uint32, ctypes.c_float, ctypes.c_int,
                                         ctypes.c_int, up, up,
                                         ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64]
        L.oracle_qtree_stats.restype = ctypes.c_int64
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_get_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _up(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def params(wsize, bsize, overlap, box, n, leaf_size=32, level=5):
    """Derived parameters (application/owm_optim3.cpp:92-132); None if rejected."""
    p = Params()
    p.wsize, p.bsize, p.overlap, p.leaf_size, p.level = wsize, bsize, overlap, leaf_size, level
    p.minx, p.maxx, p.miny, p.maxy = box
    if lib().oracle_derive(ctypes.byref(p), n) != 0:
        return None
    return p


def set_threads(n):
    lib().oracle_set_threads(int(n))


def threads():
    return int(lib().oracle_get_threads())


class Tree:
    def __init__(self, x, y, z, p):
        self.x, self.y, self.z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
        self.p = p
        self.h = lib().oracle_tree_build(_dp(self.x), _dp(self.y), _dp(self.z), len(self.x),
                                         ctypes.byref(p))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_tree_free(self.h)
            self.h = None

    def nodes(self):
        return int(lib().oracle_tree_nodes(self.h))

    def check(self):
        return int(lib().oracle_tree_check(self.h))

    def stage1(self):
        out = np.empty(self.p.ncols * self.p.nrows, np.uint32)
        lib().oracle_stage1(self.h, ctypes.byref(self.p), _up(out))
        return out

    def stage1_cols(self, col0, col1, out=None):
        """stage1 of the windows in columns [col0, col1) (every row), at the
        indices stage1 uses; other entries of out are left as they are."""
        if out is None:
            out = np.full(self.p.ncols * self.p.nrows, NONE, np.uint32)
        lib().oracle_stage1_cols(self.h, ctypes.byref(self.p), col0, col1, _up(out))
        return out

    def stage3(self, llp):
        llp = np.ascontiguousarray(llp, np.uint32)
        out = np.empty(max(self.p.ncolsg * self.p.nrowsg, 1), np.uint32)
        k = lib().oracle_stage3(self.h, ctypes.byref(self.p), _up(llp), len(llp), _up(out))
        return out[:k].copy()


def stage2(p, min_ids):
    m = np.array(min_ids, np.uint32, copy=True)
    out = np.empty(max(len(m), 1), np.uint32)
    k = lib().oracle_stage2(ctypes.byref(p), _up(m), len(m), _up(out))
    return out[:k].copy()


def stage1_brute(x, y, z, p):
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    out = np.empty(p.ncols * p.nrows, np.uint32)
    lib().oracle_stage1_brute(_dp(x), _dp(y), _dp(z), len(x), ctypes.byref(p), _up(out))
    return out


def stage3_brute(x, y, z, p, llp):
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    llp = np.ascontiguousarray(llp, np.uint32)
    out = np.empty(max(p.ncolsg * p.nrowsg, 1), np.uint32)
    k = lib().oracle_stage3_brute(_dp(x), _dp(y), _dp(z), len(x), ctypes.byref(p), _up(llp),
                                  len(llp), _up(out))
    return out[:k].copy()


def run(x, y, z, wsize, bsize, overlap, leaf_size=32, box=None, level=5):
    """Whole pipeline; returns (seeds sorted ascending, times {tree,s1,s2,s3}, params)."""
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    if box is None:
        box = (float(np.nanmin(x)), float(np.nanmax(x)), float(np.nanmin(y)), float(np.nanmax(y)))
    p = Params()
    p.wsize, p.bsize, p.overlap, p.leaf_size, p.level = wsize, bsize, overlap, leaf_size, level
    p.minx, p.maxx, p.miny, p.maxy = box
    cap = len(x) + 16
    seeds = np.empty(cap, np.uint32)
    ns = ctypes.c_uint64(0)
    t = np.zeros(4, np.float64)
    rc = lib().oracle_run(_dp(x), _dp(y), _dp(z), len(x), ctypes.byref(p), _up(seeds), cap,
                          ctypes.byref(ns), _dp(t))
    if rc != 0:
        raise ValueError(f"oracle rejected the parameters (rc={rc})")
    return seeds[:ns.value].copy(), dict(zip(("tree", "s1", "s2", "s3"), t.tolist())), p


def seeds_from_stages(x, y, z, wsize, bsize, overlap, leaf_size=32, box=None, brute=False):
    """Pipeline assembled from the stage functions (or the brute-force ones)."""
    if box is None:
        box = (float(np.nanmin(x)), float(np.nanmax(x)), float(np.nanmin(y)), float(np.nanmax(y)))
    p = params(wsize, bsize, overlap, box, len(x), leaf_size)
    if p is None:
        raise ValueError("oracle rejected the parameters")
    if brute:
        mins = stage1_brute(x, y, z, p)
    else:
        tree = Tree(x, y, z, p)
        mins = tree.stage1()
    llp = stage2(p, mins)
    adds = stage3_brute(x, y, z, p, llp) if brute else tree.stage3(llp)
    return np.sort(np.concatenate([llp, adds]).astype(np.uint32)), mins, llp, adds, p


def parse_xyz(text: bytes, cap: int):
    """The reference reader's point loop over a text (strtod per number)."""
    x, y, z = (np.empty(cap, np.float64) for _ in range(3))
    n = lib().oracle_parse_xyz(text, len(text), _dp(x), _dp(y), _dp(z), cap)
    if n < 0:
        raise ValueError("a line holds fewer than three numbers")
    return x[:n], y[:n], z[:n]


QT_MAXNUMBER, QT_MINRADIUS = 0, 1


def qtree_stats(x, y, box, policy=QT_MAXNUMBER, max_number=32, min_radius=0.1, level=5,
                max_depth=64):
    """Leaf classes of the reference's parallel_qtree as owm_stats walks it:
    a list of (depth, points, leaves), ascending; None when a split would go
    below max_depth."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    cap = 4096
    while True:
        lv = np.empty(cap, np.uint32)
        ct = np.empty(cap, np.uint32)
        nl = np.empty(cap, np.uint64)
        r = lib().oracle_qtree_stats(_dp(x), _dp(y), len(x), box[0], box[1], box[2], box[3],
                                     policy, max_number, min_radius, level, max_depth,
                                     _up(lv), _up(ct),
                                     nl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), cap)
        if r == -2:
            return None
        if r < 0:
            raise ValueError("oracle_qtree_stats: bad arguments")
        if r <= cap:
            return [(int(lv[i]), int(ct[i]), int(nl[i])) for i in range(r)]
        cap = int(r)
