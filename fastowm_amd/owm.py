"""Python host interface of the MI355X OWM ground-seed pipeline (ctypes over
libfastowm.so, the C ABI of include/owm.h).

The names mirror the reference's C++ interface for this path
(CPP-TBB-CPU/include/optim3_func.hpp:85-93): ``parallel_qtree`` builds the
spatial index (there: quadtree + storeMinAndNumPoints; here: a device-resident
LBVH), ``stage1`` / ``stage2`` / ``stage3`` run the three OWM stages on it, and
``run_tile`` is the reference driver's timed loop (owm_optim3.cpp:143-230).
Point ids are 0-based indices into the caller's arrays (the reference's 1-based
id minus one).  Errors raise OWMError carrying the library's status and message;
nothing falls back to a CPU path: without the HIP library every call raises.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import sys
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfastowm.so")
NONE = 0xFFFFFFFF
F_HAVE_BBOX = 1
B_GRAPH = 1

STATUS = {0: "OK", 1: "E_ARG", 2: "E_NOMEM", 3: "E_BUDGET", 4: "E_DEVICE", 5: "E_CAPACITY",
          6: "E_CANCELLED", 7: "E_NOTREADY", 8: "E_NOTFOUND", 9: "E_INTERNAL"}


class OWMError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status
        self.name = STATUS.get(status, str(status))


class Params(ctypes.Structure):
    _fields_ = [("wsize", ctypes.c_uint32), ("bsize", ctypes.c_uint32),
                ("overlap", ctypes.c_double), ("leaf_size", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("bbox", ctypes.c_double * 4)]


class Derived(ctypes.Structure):
    _fields_ = [("width", ctypes.c_double), ("height", ctypes.c_double),
                ("density", ctypes.c_double), ("displace", ctypes.c_double),
                ("min_num_points", ctypes.c_uint32), ("ncols", ctypes.c_uint32),
                ("nrows", ctypes.c_uint32), ("ncolsg", ctypes.c_uint32),
                ("nrowsg", ctypes.c_uint32), ("bbox", ctypes.c_double * 4)]


class PhaseStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("ms", ctypes.c_double), ("calls", ctypes.c_uint64),
                ("bytes", ctypes.c_uint64), ("items", ctypes.c_uint64)]


class TileDesc(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint64), ("n", ctypes.c_uint64), ("params", Params)]


class CloudDesc(ctypes.Structure):
    _fields_ = [("x", ctypes.c_void_p), ("y", ctypes.c_void_p), ("z", ctypes.c_void_p),
                ("n", ctypes.c_uint64), ("bbox", ctypes.c_double * 4), ("flags", ctypes.c_uint32),
                ("host_lines", ctypes.c_uint64)]


class TuneEntry(ctypes.Structure):
    _fields_ = [("max_n", ctypes.c_uint64), ("leaf_size", ctypes.c_uint32),
                ("block_x", ctypes.c_uint32), ("block_y", ctypes.c_uint32),
                ("block_k", ctypes.c_uint32), ("summaries", ctypes.c_uint32),
                ("sort_bits", ctypes.c_uint32),
                ("ms", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


TUNE_CLASSES = 6
# leaf histogram settings of the statistics-driven tuner (owm_autotune_stats)
TUNE_STATS_MAXNUMBER, TUNE_STATS_LEVEL = 256, 0
TREE_LEAF = 0x80000000


class TreeView(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("n_leaves", ctypes.c_uint64),
                ("n_internal", ctypes.c_uint64), ("leaf_size", ctypes.c_uint32),
                ("key_bits", ctypes.c_uint32), ("root", ctypes.c_uint32),
                ("qmin", ctypes.c_double * 2), ("qscale", ctypes.c_double)] + \
               [(f, ctypes.c_void_p) for f in ("ids", "xs", "ys", "zs", "leaf_key", "leaf_start",
                                               "leaf_points", "leaf_count", "leaf_bbox", "leaf_min",
                                               "node_child", "node_count", "node_bbox", "node_min")]

QT_MAXNUMBER, QT_MINRADIUS = 0, 1


class QtreeParams(ctypes.Structure):
    _fields_ = [("policy", ctypes.c_uint32), ("max_number", ctypes.c_uint32),
                ("min_radius", ctypes.c_float), ("level", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("bbox", ctypes.c_double * 4)]


class LeafBin(ctypes.Structure):
    _fields_ = [("level", ctypes.c_uint32), ("points", ctypes.c_uint32),
                ("leaves", ctypes.c_uint64)]


class QtreeSummary(ctypes.Structure):
    _fields_ = [("bbox", ctypes.c_double * 4), ("max_radius", ctypes.c_float),
                ("depth", ctypes.c_uint32), ("leaves", ctypes.c_uint64),
                ("n_bins", ctypes.c_uint64)]


class TuneSample(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("weight", ctypes.c_double), ("params", Params),
                ("bins", ctypes.POINTER(LeafBin)), ("n_bins", ctypes.c_uint64),
                ("summary", QtreeSummary)]


def qtree_params(policy=QT_MAXNUMBER, max_number=32, min_radius=0.1, level=5, bbox=None):
    """owm_qtree_params; policy is QT_MAXNUMBER / QT_MINRADIUS or the
    reference binary's name, "statsmaxnum" / "stats"."""
    if isinstance(policy, str):
        policy = {"statsmaxnum": QT_MAXNUMBER, "maxnumber": QT_MAXNUMBER,
                  "stats": QT_MINRADIUS, "minradius": QT_MINRADIUS}[policy.lower()]
    q = QtreeParams(policy, max_number, min_radius, level, 0)
    if bbox is not None:
        q.flags = F_HAVE_BBOX
        for k in range(4):
            q.bbox[k] = bbox[k]
    return q


def qtree_histograms(classes, summary):
    """The statistics tool's four histograms from the leaf classes, as its
    makeHistogram calls build them (CPP-TBB-CPU/include/stats_func.hpp:58-118,
    application/owm_stats.cpp:98-133): (value, leaves) pairs in ascending
    value order for points per leaf, leaf radius (the root's float radius
    halved per depth), leaf density (points / (4 r r) in float arithmetic, as
    getDensity) and leaf depth."""
    r0 = np.float32(summary["max_radius"])
    pts, rad, den, lev = {}, {}, {}, {}
    for d, c, k in classes:
        r = r0
        for _ in range(d):
            r = np.float32(float(r) * 0.5)
        dens = float(np.float32(c) / (np.float32(4) * r * r))
        for h, v in ((pts, c), (rad, float(r)), (den, dens), (lev, d)):
            h[v] = h.get(v, 0) + k
    return {name: sorted(h.items()) for name, h in
            (("points", pts), ("radius", rad), ("density", den), ("levels", lev))}


def write_qtree_stats(path, classes, summary, params, n_points):
    """Appends the statistics tool's four reports (owm_qtree_stats_write);
    classes and summary as Context.qtree_stats returns them."""
    arr = (LeafBin * max(len(classes), 1))(*[LeafBin(*c) for c in classes])
    sm = QtreeSummary()
    for k in range(4):
        sm.bbox[k] = summary["bbox"][k]
    sm.max_radius = summary["max_radius"]
    sm.depth, sm.leaves, sm.n_bins = summary["depth"], summary["leaves"], len(classes)
    st = lib().owm_qtree_stats_write(str(path).encode(), ctypes.byref(params), ctypes.byref(sm),
                                     arr, len(classes), n_points)
    if st != 0:
        raise OWMError(st, f"owm_qtree_stats_write({path}) failed")


_lib = None
# Contexts not closed yet.  At interpreter exit they are closed by an atexit
# hook, which runs before module teardown and before the HIP runtime is torn
# down; finalizers that run later (sys.is_finalizing()) make no library call.
_live_contexts = weakref.WeakSet()


def _close_live_contexts():
    for c in list(_live_contexts):
        try:
            c.close()
        except Exception:
            pass


def lib():
    """The loaded HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        # One HIP runtime per process: when torch is installed its bundled
        # libamdhip64.so.7 must be the one loaded first, so that this library's
        # dependency of the same soname binds to it instead of a second copy.
        try:
            import torch  # noqa: F401
        except Exception:
            pass
        L = ctypes.CDLL(LIB_PATH)
        # registered after torch's own exit handlers, so it runs before them
        atexit.register(_close_live_contexts)
        vp, u64, u32, dp = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p
        P = ctypes.POINTER
        sig = {
            "owm_ctx_create": ([ctypes.c_int, u64, P(vp)], ctypes.c_int),
            "owm_ctx_destroy": ([vp], None),
            "owm_last_error": ([vp], ctypes.c_char_p),
            "owm_status_str": ([ctypes.c_int], ctypes.c_char_p),
            "owm_ctx_device_bytes": ([vp], u64),
            "owm_host_alloc": ([u64, P(vp)], ctypes.c_int),
            "owm_host_free": ([vp], None),
            "owm_derive": ([P(Params), u64, dp, P(Derived)], ctypes.c_int),
            "owm_query_seed_count": ([P(Params), u64, dp, P(u64)], ctypes.c_int),
            "owm_run_tile": ([vp, dp, dp, dp, u64, P(Params), dp, u64, P(u64)], ctypes.c_int),
            "owm_run_tile_device": ([vp, dp, dp, dp, u64, P(Params), dp, u64, P(u64)], ctypes.c_int),
            "owm_run_tile_device_noalloc": ([vp, dp, dp, dp, u64, P(Params), dp, u64, P(u64)],
                                            ctypes.c_int),
            "owm_xyz_parse": ([vp, ctypes.c_char_p, u64, P(CloudDesc)], ctypes.c_int),
            "owm_gold_check": ([dp, dp, dp, u64, dp, dp, dp, u64, ctypes.c_double, P(u64),
                                P(ctypes.c_double)], ctypes.c_int),
            "owm_cloud_read": ([vp, P(CloudDesc), dp, dp, dp], ctypes.c_int),
            "owm_batch_run": ([vp, dp, dp, dp, P(TileDesc), u32, u32, dp, u64, dp], ctypes.c_int),
            "owm_batch_run_device": ([vp, dp, dp, dp, P(TileDesc), u32, u32, dp, u64, dp],
                                     ctypes.c_int),
            "owm_tree_build": ([vp, dp, dp, dp, u64, ctypes.c_int, u32, P(vp)], ctypes.c_int),
            "owm_tree_free": ([vp], None),
            "owm_tree_points": ([vp], u64),
            "owm_tree_info": ([vp, P(u64), P(u64), P(u64), dp], ctypes.c_int),
            "owm_tree_read": ([vp, P(TreeView)], ctypes.c_int),
            "owm_stage1": ([vp, vp, P(Params), dp, u64, P(Derived)], ctypes.c_int),
            "owm_stage2": ([vp, vp, P(Params), dp, u64, dp, u64, P(u64)], ctypes.c_int),
            "owm_stage3": ([vp, vp, P(Params), dp, u64, dp, u64, P(u64)], ctypes.c_int),
            "owm_tree_run": ([vp, vp, P(Params), dp, u64, P(u64)], ctypes.c_int),
            "owm_serve_start": ([vp, u32, u32], ctypes.c_int),
            "owm_serve_submit": ([vp, dp, dp, dp, u64, P(Params), u64, P(u64)], ctypes.c_int),
            "owm_serve_submit_device": ([vp, dp, dp, dp, u64, P(Params), u64, P(u64)], ctypes.c_int),
            "owm_serve_poll": ([vp, u64, dp, u64, P(u64), P(ctypes.c_double)], ctypes.c_int),
            "owm_serve_wait": ([vp, u64, dp, u64, P(u64), P(ctypes.c_double)], ctypes.c_int),
            "owm_serve_cancel": ([vp, u64], ctypes.c_int),
            "owm_serve_evict": ([vp, u64], ctypes.c_int),
            "owm_serve_stop": ([vp], ctypes.c_int),
            "owm_profile_enable": ([vp, ctypes.c_int], ctypes.c_int),
            "owm_profile_phases": ([vp, u32], ctypes.c_int),
            "owm_profile_read": ([vp, P(PhaseStat), u32, P(u32)], ctypes.c_int),
            "owm_autotune": ([vp, dp, dp, dp, ctypes.c_int, P(TileDesc), u32, P(TuneEntry), u32,
                              P(u32)], ctypes.c_int),
            "owm_autotune_stats": ([vp, dp, dp, dp, ctypes.c_int, P(TileDesc), u32, P(TuneEntry),
                                    u32, P(u32)], ctypes.c_int),
            "owm_tune_model": ([P(TuneSample), u32, P(TuneEntry), u32, P(u32)], ctypes.c_int),
            "owm_tune_get": ([vp, P(TuneEntry), u32, P(u32)], ctypes.c_int),
            "owm_tune_set": ([vp, P(TuneEntry), u32], ctypes.c_int),
            "owm_qtree_stats": ([vp, dp, dp, u64, ctypes.c_int, P(QtreeParams), P(LeafBin),
                                 u64, P(QtreeSummary)], ctypes.c_int),
            "owm_qtree_stats_write": ([ctypes.c_char_p, P(QtreeParams), P(QtreeSummary),
                                       P(LeafBin), u64, u64], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def make_params(wsize=12, bsize=20, overlap=0.8, leaf_size=0, bbox=None) -> Params:
    p = Params()
    p.wsize, p.bsize, p.overlap, p.leaf_size = int(wsize), int(bsize), float(overlap), int(leaf_size)
    p.flags = 0
    if bbox is not None:
        p.flags |= F_HAVE_BBOX
        for k in range(4):
            p.bbox[k] = float(bbox[k])
    return p


def gold_check(seeds_xyz, gold_xyz, displace):
    """check_results of the reference: (correct seeds, percent of the GOLD
    points); seeds_xyz and gold_xyz are (x, y, z) array triples."""
    sx, sy, sz = (np.ascontiguousarray(a, np.float64) for a in seeds_xyz)
    gx, gy, gz = (np.ascontiguousarray(a, np.float64) for a in gold_xyz)
    n = ctypes.c_uint64()
    pct = ctypes.c_double()
    st = lib().owm_gold_check(sx.ctypes.data, sy.ctypes.data, sz.ctypes.data, len(sx),
                              gx.ctypes.data, gy.ctypes.data, gz.ctypes.data, len(gx),
                              float(displace), ctypes.byref(n), ctypes.byref(pct))
    if st:
        raise OWMError(st, "gold_check")
    return n.value, pct.value


def query_seed_count(wsize, bsize, overlap, n, bbox=None) -> int:
    """Upper bound on the seeds of an n-point tile (seed-buffer sizing)."""
    p = make_params(wsize, bsize, overlap, 0, None)
    out = ctypes.c_uint64()
    bb = (ctypes.c_double * 4)(*bbox) if bbox is not None else None
    st = lib().owm_query_seed_count(ctypes.byref(p), n, ctypes.cast(bb, ctypes.c_void_p) if bb else None,
                                    ctypes.byref(out))
    if st != 0:
        raise OWMError(st, "owm_query_seed_count")
    return out.value


def derive(wsize, bsize, overlap, n, bbox) -> Derived:
    """Derived parameters (owm_optim3.cpp:92-132) without device work."""
    d = Derived()
    p = make_params(wsize, bsize, overlap, 0, bbox)
    st = lib().owm_derive(ctypes.byref(p), int(n), None, ctypes.byref(d))
    if st:
        raise OWMError(st, "parameters rejected")
    return d


def _host(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data


def _is_device(a):
    return hasattr(a, "is_cuda") and a.is_cuda


def _dev_ptr(t):
    import torch
    if t.dtype != torch.float64 or not t.is_contiguous():
        raise TypeError("device arrays must be contiguous float64 torch tensors")
    # the library runs on its own stream: make the producer's work visible first
    torch.cuda.current_stream(t.device).synchronize()
    return t.data_ptr()


def tune_model(samples):
    """The statistics model alone (owm_tune_model, host code): samples are
    (n, weight, dict(wsize=, overlap=, bsize=, bbox=), bins, summary) with
    bins/summary as Context.qtree_stats returns them (maxNumber policy,
    TUNE_STATS_MAXNUMBER, level TUNE_STATS_LEVEL).  Returns the table."""
    arr = (TuneSample * max(len(samples), 1))()
    keep = []
    for i, (n, weight, prm, bins, summary) in enumerate(samples):
        b = (LeafBin * max(len(bins), 1))()
        for k, (lv, pts, leaves) in enumerate(bins):
            b[k].level, b[k].points, b[k].leaves = lv, pts, leaves
        keep.append(b)
        arr[i].n, arr[i].weight = n, weight
        arr[i].params = make_params(prm.get("wsize", 12), prm.get("bsize", 20),
                                    prm.get("overlap", 0.8), 0, prm.get("bbox"))
        arr[i].bins = ctypes.cast(b, ctypes.POINTER(LeafBin))
        arr[i].n_bins = len(bins)
        arr[i].summary.bbox[:] = list(summary["bbox"])
        arr[i].summary.max_radius = summary["max_radius"]
        arr[i].summary.n_bins = len(bins)
    out = (TuneEntry * TUNE_CLASSES)()
    n = ctypes.c_uint32()
    st = lib().owm_tune_model(arr, len(samples), out, TUNE_CLASSES, ctypes.byref(n))
    if st != 0:
        raise OWMError(st, "owm_tune_model rejected the samples")
    return [out[i].as_dict() for i in range(n.value)]


class Tree:
    """A device-resident spatial index of one cloud (parallel_qtree +
    storeMinAndNumPoints of the reference)."""

    def __init__(self, ctx: "Context", handle, n: int):
        self.ctx, self.h, self.n = ctx, handle, n

    def info(self):
        nl, ni, rc = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        box = (ctypes.c_double * 4)()
        self.ctx._check(lib().owm_tree_info(self.h, ctypes.byref(nl), ctypes.byref(ni),
                                            ctypes.byref(rc), box))
        return {"leaves": nl.value, "internal": ni.value, "root_count": rc.value,
                "root_bbox": tuple(box)}

    def read(self):
        """The tree's structure (owm_tree_read): scalars and every array as
        numpy (bboxes as (k, 4) minx, maxx, miny, maxy; node_child (k, 2))."""
        v = TreeView()
        self.ctx._check(lib().owm_tree_read(self.h, ctypes.byref(v)))
        n, nl, nn = v.n, v.n_leaves, v.n_internal
        shapes = {"ids": (n, np.uint32), "xs": (n, np.float64), "ys": (n, np.float64),
                  "zs": (n, np.float64), "leaf_key": (nl, np.uint32),
                  "leaf_start": (nl, np.uint32), "leaf_points": (nl, np.uint32),
                  "leaf_count": (nl, np.uint32), "leaf_bbox": (4 * nl, np.float64),
                  "leaf_min": (nl, np.uint32), "node_child": (2 * nn, np.uint32),
                  "node_count": (nn, np.uint32), "node_bbox": (4 * nn, np.float64),
                  "node_min": (nn, np.uint32)}
        out = {k: np.empty(max(c, 1), t) for k, (c, t) in shapes.items()}
        for k, a in out.items():
            setattr(v, k, a.ctypes.data)
        self.ctx._check(lib().owm_tree_read(self.h, ctypes.byref(v)))
        res = {k: out[k][:c] for k, (c, _) in shapes.items()}
        res["leaf_bbox"] = res["leaf_bbox"].reshape(-1, 4)
        res["node_bbox"] = res["node_bbox"].reshape(-1, 4)
        res["node_child"] = res["node_child"].reshape(-1, 2)
        res.update(n=n, n_leaves=nl, n_internal=nn, leaf_size=v.leaf_size, key_bits=v.key_bits,
                   root=v.root, qmin=(v.qmin[0], v.qmin[1]), qscale=v.qscale)
        return res

    def free(self):
        if self.h and getattr(self.ctx, "h", None):
            lib().owm_tree_free(self.h)
        self.h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # bound at definition: module globals may already be gone at exit
        if _finalizing():
            return  # the exit hook closed the context and its trees already
        try:
            self.free()
        except Exception:
            pass


class Cloud:
    """A point cloud parsed on the device (owm_xyz_parse).  Its arrays live
    in the context's reader buffers: the next parse on the context, or
    closing it, retires the cloud (using it then raises)."""

    def __init__(self, ctx: "Context", desc, generation: int):
        self.ctx, self.desc, self.generation = ctx, desc, generation
        self.n = desc.n
        self.bbox = tuple(desc.bbox) if desc.flags & F_HAVE_BBOX else None
        self.host_lines = desc.host_lines

    def check(self):
        if not getattr(self.ctx, "h", None) or self.ctx._clouds != self.generation:
            raise OWMError(1, "this cloud was retired by a later parse or by closing its context")

    def to_numpy(self):
        self.check()
        x, y, z = (np.empty(self.n, np.float64) for _ in range(3))
        self.ctx._check(lib().owm_cloud_read(self.ctx.h, ctypes.byref(self.desc), x.ctypes.data,
                                             y.ctypes.data, z.ctypes.data))
        return x, y, z


class BatchDesc:
    """The owm_tile array of a batch (tiles as for Context.batch_run), built
    once so that repeated runs of the same batch skip rebuilding it."""

    def __init__(self, tiles):
        self.descs = (TileDesc * len(tiles))()
        off = 0
        for i, (n, kw) in enumerate(tiles):
            self.descs[i].offset = off
            self.descs[i].n = n
            self.descs[i].params = make_params(**kw)
            off += n
        self.points = off

    def __len__(self):
        return len(self.descs)


def batch_desc(tiles):
    return BatchDesc(tiles)


class Context:
    """One HIP device, one stream, reusable device arenas."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = ctypes.c_void_p()
        st = lib().owm_ctx_create(device, hbm_budget, ctypes.byref(h))
        if st:
            raise OWMError(st, f"cannot create a context on device {device}")
        self.h = h
        _live_contexts.add(self)
        self._trees = weakref.WeakSet()
        self._seed_buf = np.empty(0, np.uint32)
        self._seed_mem = None
        self._clouds = 0  # parses so far (a Cloud is valid for the latest only)

    def _seeds_out(self, cap):
        """Reusable page-locked host buffer the library copies seeds into
        (results are copied out of it): a fresh array per call would cost the
        page faults of an n-sized allocation every time, and pageable memory
        a staged copy.  It only grows."""
        if len(self._seed_buf) < cap:
            self._free_seed_buf()
            p = ctypes.c_void_p()
            if lib().owm_host_alloc(cap * 4, ctypes.byref(p)) == 0 and p.value:
                self._seed_mem = p
                buf = (ctypes.c_uint32 * cap).from_address(p.value)
                self._seed_buf = np.frombuffer(buf, np.uint32)
            else:
                self._seed_buf = np.empty(cap, np.uint32)
        return self._seed_buf

    def _free_seed_buf(self):
        self._seed_buf = np.empty(0, np.uint32)
        if self._seed_mem is not None:
            lib().owm_host_free(self._seed_mem)
            self._seed_mem = None

    def close(self):
        if getattr(self, "h", None):
            for t in list(self._trees):
                t.free()
            h, self.h = self.h, None
            lib().owm_ctx_destroy(h)
            self._free_seed_buf()
            _live_contexts.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self, _finalizing=sys.is_finalizing):
        # bound at definition: module globals may already be gone at exit
        if _finalizing():
            return  # closed by the exit hook; no HIP call during interpreter teardown
        try:
            self.close()
        except Exception:
            pass

    def last_error(self) -> str:
        return lib().owm_last_error(self.h).decode()

    def device_bytes(self) -> int:
        return int(lib().owm_ctx_device_bytes(self.h))

    def _check(self, st):
        if st:
            raise OWMError(st, self.last_error())

    def profile(self, enable: bool = True, phases=None):
        """Start (zeroing) or stop per-phase HIP-event timing; `phases` (names
        as in profile_read) limits the events to those phases."""
        mask = 0
        if enable and phases:
            names = list(self.profile_read())
            for p in phases:
                mask |= 1 << names.index(p)
        self._check(lib().owm_profile_phases(self.h, mask))
        self._check(lib().owm_profile_enable(self.h, 1 if enable else 0))

    def profile_read(self):
        """{phase: {ms, calls, bytes, items}} summed since profile(True)."""
        arr = (PhaseStat * 32)()
        n = ctypes.c_uint32()
        self._check(lib().owm_profile_read(self.h, arr, 32, ctypes.byref(n)))
        return {arr[i].name.decode(): {"ms": arr[i].ms, "calls": arr[i].calls,
                                       "bytes": arr[i].bytes, "items": arr[i].items}
                for i in range(n.value)}

    # -- reader ------------------------------------------------------------
    def parse_xyz(self, data):
        """readXYZfile on the device: `data` is the file's bytes (or a path).
        Returns a Cloud whose arrays stay in device memory until the next
        parse on this context."""
        if isinstance(data, str):
            with open(data, "rb") as f:
                data = f.read()
        c = CloudDesc()
        self._clouds += 1  # earlier clouds' memory is reused from here on
        self._check(lib().owm_xyz_parse(self.h, data, len(data), ctypes.byref(c)))
        return Cloud(self, c, self._clouds)

    def run_cloud(self, cloud, wsize=12, bsize=20, overlap=0.8, leaf_size=0, bbox=None, cap=None):
        """run_tile on a parsed cloud; the file header's bbox is used when it
        had one and no bbox is given (as the reference driver does)."""
        cloud.check()
        if bbox is None and cloud.bbox is not None:
            bbox = cloud.bbox
        p = make_params(wsize, bsize, overlap, leaf_size, bbox)
        cap = cloud.n + 1 if cap is None else cap
        out = self._seeds_out(max(cap, 1))
        ns = ctypes.c_uint64()
        d = cloud.desc
        self._check(lib().owm_run_tile_device(self.h, d.x, d.y, d.z, d.n, ctypes.byref(p),
                                              out.ctypes.data, cap, ctypes.byref(ns)))
        return out[:ns.value].copy()

    def qtree_stats(self, x, y=None, policy=QT_MAXNUMBER, max_number=32, min_radius=0.1,
                    level=5, bbox=None):
        """Leaf statistics of the reference's quadtree (owm_stats): a list of
        (level, points, leaves), ascending, and a summary dict.  x/y are numpy
        arrays or float64 torch tensors on the GPU, or x is a parsed Cloud
        (whose header box is used when it had one and bbox is None)."""
        if isinstance(x, Cloud):
            x.check()
            if bbox is None and x.bbox is not None:
                bbox = x.bbox
            d = x.desc
            dev, px, py, n = 1, d.x, d.y, d.n
        elif _is_device(x):
            dev, px, py, n = 1, _dev_ptr(x), _dev_ptr(y), len(x)
        else:
            (xa, px), (ya, py) = _host(x), _host(y)
            dev, n = 0, len(xa)
        q = qtree_params(policy, max_number, min_radius, level, bbox)
        sm = QtreeSummary()
        cap = 4096
        while True:
            arr = (LeafBin * cap)()
            st = lib().owm_qtree_stats(self.h, px, py, n, dev, ctypes.byref(q), arr, cap,
                                       ctypes.byref(sm))
            if st == 5 and sm.n_bins > cap:  # OWM_E_CAPACITY
                cap = int(sm.n_bins)
                continue
            self._check(st)
            break
        classes = [(arr[k].level, arr[k].points, arr[k].leaves) for k in range(sm.n_bins)]
        summary = {"bbox": tuple(sm.bbox), "max_radius": sm.max_radius, "depth": sm.depth,
                   "leaves": sm.leaves, "params": q, "n": n}
        return classes, summary

    # -- whole pipeline ----------------------------------------------------
    def run_tile(self, x, y, z, wsize=12, bsize=20, overlap=0.8, leaf_size=0, bbox=None,
                 cap=None):
        """Seeds (sorted 0-based ids) of one tile; x/y/z are numpy arrays or
        contiguous float64 torch tensors already on the GPU."""
        p = make_params(wsize, bsize, overlap, leaf_size, bbox)
        n = len(x)
        cap = n + 1 if cap is None else cap
        out = self._seeds_out(max(cap, 1))
        ns = ctypes.c_uint64()
        if _is_device(x):
            st = lib().owm_run_tile_device(self.h, _dev_ptr(x), _dev_ptr(y), _dev_ptr(z), n,
                                           ctypes.byref(p), out.ctypes.data, cap, ctypes.byref(ns))
        else:
            (xa, xp), (ya, yp), (za, zp) = _host(x), _host(y), _host(z)
            st = lib().owm_run_tile(self.h, xp, yp, zp, n, ctypes.byref(p), out.ctypes.data, cap,
                                    ctypes.byref(ns))
        self._check(st)
        return out[:ns.value].copy()

    @staticmethod
    def _descs(tiles):
        if isinstance(tiles, BatchDesc):
            return tiles.descs, tiles.points
        d = BatchDesc(tiles)
        return d.descs, d.points

    def batch_run(self, x, y, z, tiles, graph=False, copy=True, split=True):
        """tiles: list of (n, dict(wsize=, bsize=, overlap=, leaf_size=, bbox=)) in
        cloud order, or a BatchDesc of such a list (built once, reused).
        Returns a list of seed arrays (tile-local ids).  graph=True captures the
        device chain once per batch layout and replays it.  copy=False returns
        views of the context's page-locked seed buffer instead of copies: they
        are valid until the next call on this context.  split=False returns
        the C ABI's form instead of the list: all seeds in one array and the
        n_tiles + 1 offsets of each tile's range in it (no per-tile objects)."""
        descs, off = self._descs(tiles)
        n_tiles = len(descs)
        cap = off + n_tiles + 1
        out = self._seeds_out(cap)
        so = np.zeros(n_tiles + 1, np.uint64)
        flags = B_GRAPH if graph else 0
        if _is_device(x):
            st = lib().owm_batch_run_device(self.h, _dev_ptr(x), _dev_ptr(y), _dev_ptr(z), descs,
                                            n_tiles, flags, out.ctypes.data, cap, so.ctypes.data)
        else:
            (xa, xp), (ya, yp), (za, zp) = _host(x), _host(y), _host(z)
            st = lib().owm_batch_run(self.h, xp, yp, zp, descs, n_tiles, flags,
                                     out.ctypes.data, cap, so.ctypes.data)
        self._check(st)
        seeds = out[:int(so[-1])]
        if copy:
            seeds = seeds.copy()  # one copy out of the reusable buffer
        if not split:
            return seeds, so
        return np.split(seeds, so[1:-1].astype(np.intp))  # a view per tile

    # -- autotuner ----------------------------------------------------------
    def autotune(self, x, y, z, tiles, method="trials"):
        """Tunes the table for this batch (same form as batch_run) and returns
        it.  method "trials" times leaf sizes, window blocks and sort widths
        and keeps the winners (owm_autotune); "stats" lets the cost model pick
        them from leaf histograms of sampled tiles (owm_autotune_stats)."""
        fn = {"trials": lib().owm_autotune, "stats": lib().owm_autotune_stats}[method]
        descs, _ = self._descs(tiles)
        out = (TuneEntry * TUNE_CLASSES)()
        n = ctypes.c_uint32()
        if _is_device(x):
            st = fn(self.h, _dev_ptr(x), _dev_ptr(y), _dev_ptr(z), 1, descs, len(descs), out,
                    TUNE_CLASSES, ctypes.byref(n))
        else:
            (xa, xp), (ya, yp), (za, zp) = _host(x), _host(y), _host(z)
            st = fn(self.h, xp, yp, zp, 0, descs, len(descs), out, TUNE_CLASSES, ctypes.byref(n))
        self._check(st)
        return [out[i].as_dict() for i in range(n.value)]

    def tune_table(self):
        out = (TuneEntry * TUNE_CLASSES)()
        n = ctypes.c_uint32()
        self._check(lib().owm_tune_get(self.h, out, TUNE_CLASSES, ctypes.byref(n)))
        return [out[i].as_dict() for i in range(n.value)]

    def set_tune_table(self, entries):
        arr = (TuneEntry * len(entries))()
        for i, e in enumerate(entries):
            for f, _ in TuneEntry._fields_:
                setattr(arr[i], f, e[f])
        self._check(lib().owm_tune_set(self.h, arr, len(entries)))

    # -- stage-level interface (reference names) ----------------------------
    def parallel_qtree(self, x, y, z, leaf_size=32) -> Tree:
        n = len(x)
        h = ctypes.c_void_p()
        if _is_device(x):
            st = lib().owm_tree_build(self.h, _dev_ptr(x), _dev_ptr(y), _dev_ptr(z), n, 1,
                                      leaf_size, ctypes.byref(h))
        else:
            (xa, xp), (ya, yp), (za, zp) = _host(x), _host(y), _host(z)
            st = lib().owm_tree_build(self.h, xp, yp, zp, n, 0, leaf_size, ctypes.byref(h))
        self._check(st)
        tree = Tree(self, h, n)
        self._trees.add(tree)
        return tree

    def stage1(self, tree: Tree, wsize, overlap, bsize=20, bbox=None):
        """Per-window minima, min_ids[jj*ncols + ii] (NONE = invalid window)."""
        p = make_params(wsize, bsize, overlap, 0, bbox)
        d = Derived()
        cap = 1 << 20
        while True:
            out = np.empty(cap, np.uint32)
            st = lib().owm_stage1(self.h, tree.h, ctypes.byref(p), out.ctypes.data, cap,
                                  ctypes.byref(d))
            if st == 5 and d.ncols:
                cap = d.ncols * d.nrows
                continue
            self._check(st)
            return out[:d.ncols * d.nrows].copy(), d

    def stage2(self, tree: Tree, min_ids, wsize, overlap, bsize=20, bbox=None):
        p = make_params(wsize, bsize, overlap, 0, bbox)
        m = np.ascontiguousarray(min_ids, np.uint32)
        out = np.empty(max(len(m), 1), np.uint32)
        n = ctypes.c_uint64()
        self._check(lib().owm_stage2(self.h, tree.h, ctypes.byref(p), m.ctypes.data, len(m),
                                     out.ctypes.data, len(out), ctypes.byref(n)))
        return out[:n.value].copy()

    def stage3(self, tree: Tree, llp, wsize, overlap, bsize=20, bbox=None):
        p = make_params(wsize, bsize, overlap, 0, bbox)
        l_ = np.ascontiguousarray(llp, np.uint32)
        cap = 1 << 16
        while True:
            out = np.empty(cap, np.uint32)
            n = ctypes.c_uint64()
            st = lib().owm_stage3(self.h, tree.h, ctypes.byref(p), l_.ctypes.data, len(l_),
                                  out.ctypes.data, cap, ctypes.byref(n))
            if st == 5:
                cap = int(n.value)
                continue
            self._check(st)
            return out[:n.value].copy()

    def tree_run(self, tree: Tree, wsize, bsize, overlap, bbox=None):
        p = make_params(wsize, bsize, overlap, 0, bbox)
        cap = tree.n + 1
        out = np.empty(cap, np.uint32)
        n = ctypes.c_uint64()
        self._check(lib().owm_tree_run(self.h, tree.h, ctypes.byref(p), out.ctypes.data, cap,
                                       ctypes.byref(n)))
        return out[:n.value].copy()


class Server:
    """Continuous-batching serving loop on a context (owm_serve_*).

    Requests are admitted as slots and device memory free up; groups of
    different slots overlap on the device.  ``key`` != 0 keeps the tile's tree
    resident so that re-queries with other parameters skip the tree build.
    Input arrays are referenced until their request completes.
    """

    def __init__(self, ctx: Context, max_inflight: int = 2, max_batch: int = 64):
        self.ctx = ctx
        ctx._check(lib().owm_serve_start(ctx.h, max_inflight, max_batch))
        self._keep = {}

    def submit(self, x, y, z, wsize=12, bsize=20, overlap=0.8, leaf_size=0, bbox=None, key=0):
        """x/y/z: numpy arrays (host memory) or contiguous float64 torch
        tensors on the GPU (owm_serve_submit_device: no PCIe transfer)."""
        p = make_params(wsize, bsize, overlap, leaf_size, bbox)
        rid = ctypes.c_uint64()
        if _is_device(x):
            self.ctx._check(lib().owm_serve_submit_device(
                self.ctx.h, _dev_ptr(x), _dev_ptr(y), _dev_ptr(z), len(x), ctypes.byref(p),
                int(key), ctypes.byref(rid)))
            self._keep[rid.value] = (x, y, z)
            return rid.value
        (xa, xp), (ya, yp), (za, zp) = _host(x), _host(y), _host(z)
        self.ctx._check(lib().owm_serve_submit(self.ctx.h, xp, yp, zp, len(xa), ctypes.byref(p),
                                               int(key), ctypes.byref(rid)))
        self._keep[rid.value] = (xa, ya, za)
        return rid.value

    def _result(self, rid, wait):
        n = len(self._keep[rid][0]) if rid in self._keep else 1 << 20
        out = np.empty(max(n + 1, 1), np.uint32)
        ns, lat = ctypes.c_uint64(), ctypes.c_double()
        f = lib().owm_serve_wait if wait else lib().owm_serve_poll
        st = f(self.ctx.h, rid, out.ctypes.data, len(out), ctypes.byref(ns), ctypes.byref(lat))
        if st == 7:  # not ready
            return None
        self._keep.pop(rid, None)
        self.ctx._check(st)
        return out[:ns.value].copy(), lat.value

    def poll(self, rid):
        """(seeds, latency_s) when done, None while running."""
        return self._result(rid, False)

    def wait(self, rid):
        return self._result(rid, True)

    def cancel(self, rid):
        self.ctx._check(lib().owm_serve_cancel(self.ctx.h, rid))

    def evict(self, key):
        self.ctx._check(lib().owm_serve_evict(self.ctx.h, int(key)))

    def stop(self):
        if self.ctx.h:
            lib().owm_serve_stop(self.ctx.h)
        self._keep.clear()
