"""The device tree itself, read back whole (owm_tree_read) and checked by
brute force, in the spirit of the reference's checkNumPoints
(CPP-TBB-CPU/src/optim3_func.cpp:270-283):

- Morton keys: each point's sort key is the top key_bits of the bit
  interleave of its quantised x (even bits) and y (odd bits) -- a closed form
  computed here bit by bit (basic/morton.h:6-35 is the reference's 16-bit form),
  quantised against the square over a sample of the cloud (the frame the bbox
  pass takes its key counts with); de-interleaving a key gives back the top
  bits of the quantised coordinates;
- tree order: the tree's point order is numpy's stable argsort of those keys
  (the reference sorts (key, index) pairs with cub::DeviceRadixSort,
  ordered_builder.inl:460-468, whose output equals std::stable_sort), a
  permutation, and the records moved with it are the input's;
- leaves and nodes: every leaf's count, tight bbox and lowest (z, index) point
  equal a brute-force reduction over its points; every internal node's equal
  the merge of its two children's; the root holds every point; every leaf and
  node is reached exactly once and each node spans a contiguous leaf range.

Planted mistakes in a copy of the read-back tree (host arrays) must fail the
checks, so the checks are not vacuous."""
import numpy as np
import pytest

from fastowm_amd import owm, synth

pytestmark = pytest.mark.gpu
DBL_MAX = np.finfo(np.float64).max


def say(*a):
    print("[case]", *a, flush=True)


def quantise(v, vmin, scale):
    q = (v - vmin) * scale  # IEEE subtract, then multiply
    out = np.zeros(len(v), np.uint64)
    pos = q > 0  # False for NaN
    big = q >= 4294967295.0
    mid = pos & ~big
    out[mid] = np.floor(q[mid]).astype(np.uint64)
    out[big] = 0xFFFFFFFF
    return out


def interleave(qx, qy):
    m = np.zeros(len(qx), np.uint64)
    one = np.uint64(1)
    for b in range(32):
        m |= ((qx >> np.uint64(b)) & one) << np.uint64(2 * b)
        m |= ((qy >> np.uint64(b)) & one) << np.uint64(2 * b + 1)
    return m


def deinterleave(m, bits):
    """Even and odd bits of the low `bits` bits of m."""
    ex = np.zeros(len(m), np.uint64)
    ey = np.zeros(len(m), np.uint64)
    one = np.uint64(1)
    for b in range(bits // 2):
        ex |= ((m >> np.uint64(2 * b)) & one) << np.uint64(b)
        ey |= ((m >> np.uint64(2 * b + 1)) & one) << np.uint64(b)
    return ex, ey


def key_frame(x, y):
    """The keys' quantisation frame: the square over the box of 128 runs of 16
    consecutive points starting at r*n//128 (every point when n <= 2048), NaN
    skipped per axis (no finite sample: origin 0), its side mapped just under
    2^32 (zero extent: side 1)."""
    n = len(x)
    idx = np.unique(np.concatenate([np.arange(r * n // 128, min(r * n // 128 + 16, n))
                                    for r in range(128)]))
    lo, hi = [], []
    for a in (x[idx], y[idx]):
        f = a[~np.isnan(a)]
        lo.append(float(f.min()) if len(f) else 0.0)
        hi.append(float(f.max()) if len(f) else 0.0)
    side = max(hi[0] - lo[0], hi[1] - lo[1])
    if not (side > 0 and np.isfinite(side)):
        side = 1.0
    return (lo[0], lo[1]), 4294967296.0 / (side * (1.0 + 1e-12))


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def check_tree(t, x, y, z):
    n, kb, L = len(x), t["key_bits"], t["leaf_size"]
    ids = t["ids"].astype(np.int64)
    assert t["n"] == n
    # a permutation, and the records that moved with it are the input's
    assert np.array_equal(np.sort(ids), np.arange(n)), "ids are not a permutation"
    for a, b in ((t["xs"], x), (t["ys"], y), (t["zs"], z)):
        assert np.array_equal(_bits(a), _bits(b[ids])), "records do not follow the ids"
    # quantisation frame: the square over the sampled box, exactly
    qmin, qscale = key_frame(x, y)
    assert t["qmin"] == qmin and t["qscale"] == qscale, "quantisation frame"
    assert np.nanmin(x) <= qmin[0] and np.nanmin(y) <= qmin[1]
    # keys by the closed form; order = stable argsort
    qx, qy = quantise(x, t["qmin"][0], t["qscale"]), quantise(y, t["qmin"][1], t["qscale"])
    keys = interleave(qx, qy) >> np.uint64(64 - kb)
    ex, ey = deinterleave(keys, kb)
    assert np.array_equal(ex, qx >> np.uint64(32 - kb // 2)), "decode(encode) x"
    assert np.array_equal(ey, qy >> np.uint64(32 - kb // 2)), "decode(encode) y"
    assert np.array_equal(ids, np.argsort(keys, kind="stable")), "tree order != stable argsort"
    # leaves: fixed runs of L tree positions
    nl = (n + L - 1) // L
    assert t["n_leaves"] == nl and t["n_internal"] == max(nl - 1, 0)
    starts = np.arange(nl, dtype=np.int64) * L
    assert np.array_equal(t["leaf_start"], starts)
    assert np.array_equal(t["leaf_points"], np.minimum(L, n - starts))
    assert np.array_equal(t["leaf_key"], keys[ids[starts]].astype(np.uint32))
    xs, ys, zs = x[ids], y[ids], z[ids]
    ok = ~np.isnan(xs) & ~np.isnan(ys)
    assert np.array_equal(t["leaf_count"], np.add.reduceat(ok.astype(np.int64), starts))
    inf = np.inf
    want_box = np.stack([np.minimum.reduceat(np.where(ok, xs, inf), starts),
                         np.maximum.reduceat(np.where(ok, xs, -inf), starts),
                         np.minimum.reduceat(np.where(ok, ys, inf), starts),
                         np.maximum.reduceat(np.where(ok, ys, -inf), starts)], axis=1)
    assert np.array_equal(t["leaf_bbox"], want_box)
    elig = ok & (zs < DBL_MAX)
    leaf_of = np.arange(n) // L
    zz = np.where(zs == 0.0, 0.0, zs)
    e = np.nonzero(elig)[0]
    order = e[np.lexsort((ids[e], zz[e], leaf_of[e]))]
    first = np.unique(leaf_of[order], return_index=True)
    want_min = np.full(nl, 0xFFFFFFFF, np.uint32)
    want_min[first[0]] = ids[order[first[1]]]
    assert np.array_equal(t["leaf_min"], want_min), "leaf minima"
    # internal nodes: merge of the children, each reached once, contiguous ranges
    zk = np.where(z == 0.0, 0.0, z)

    def better(a, b):
        if a == 0xFFFFFFFF:
            return b
        if b == 0xFFFFFFFF:
            return a
        return a if (zk[a], a) <= (zk[b], b) else b

    nn = t["n_internal"]
    child, ncnt, nbox, nmin = t["node_child"], t["node_count"], t["node_bbox"], t["node_min"]
    seen_leaf = np.zeros(nl, bool)
    seen_node = np.zeros(nn, bool)
    agg = {}
    stack = [(t["root"], False)]
    while stack:
        ref, done = stack.pop()
        if ref & owm.TREE_LEAF:
            j = ref & ~owm.TREE_LEAF
            assert not seen_leaf[j], "leaf reached twice"
            seen_leaf[j] = True
            agg[ref] = (int(t["leaf_count"][j]), t["leaf_bbox"][j], int(t["leaf_min"][j]), j, j)
            continue
        if not done:
            assert not seen_node[ref], "node reached twice"
            seen_node[ref] = True
            stack.append((ref, True))
            stack += [(int(c), False) for c in child[ref]]
            continue
        (ca, ba, ma, la, ha), (cb, bb, mb, lb, hb) = agg[int(child[ref][0])], agg[int(child[ref][1])]
        assert ha + 1 == lb, "children do not tile a contiguous leaf range"
        box = np.array([min(ba[0], bb[0]), max(ba[1], bb[1]), min(ba[2], bb[2]), max(ba[3], bb[3])])
        assert ncnt[ref] == ca + cb, f"node {ref} count"
        assert np.array_equal(nbox[ref], box), f"node {ref} bbox"
        assert nmin[ref] == better(ma, mb), f"node {ref} minimum"
        agg[ref] = (ca + cb, box, int(nmin[ref]), la, hb)
    assert seen_leaf.all() and seen_node.all(), "unreachable leaves or nodes"
    rc, _, _, lo, hi = agg[t["root"]]
    assert rc == int(ok.sum()) and (lo, hi) == (0, nl - 1)


def _cloud(kind, n, seed):
    if kind == "c2":
        return synth.scaled(synth.config_c2(seed), n).generate()
    if kind == "roofs":
        return synth.scaled(synth.config_c3(seed), n).generate()
    t = synth.scaled(synth.config_c1(seed), n)
    if kind == "clusters":
        t.cluster_frac, t.n_clusters, t.cluster_sigma = 0.6, 5, 6.0
    x, y, z = t.generate()
    if kind == "odd":
        x[10], y[20], z[30], z[40], z[50], z[51] = np.nan, np.nan, np.nan, np.inf, 0.0, -0.0
        x[100:700], y[100:700] = x[0], y[0]  # coincident points over many leaves
        z[200:300] = -np.inf
    return x, y, z


CASES = [("clusters", 200_000, 32, 32), ("c2", 2_000_000, 64, 16), ("c2", 300_000, 7, 24),
         ("odd", 50_000, 4, 32), ("roofs", 500_000, 128, 16), ("clusters", 1, 16, 32),
         ("c2", 1000, 1000, 24)]


@pytest.fixture(scope="module")
def ctx():
    c = owm.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("kind,n,leaf,bits", CASES, ids=[f"{c[0]}_{c[1]}_l{c[2]}_k{c[3]}" for c in CASES])
def test_tree_structure(ctx, kind, n, leaf, bits):
    x, y, z = _cloud(kind, n, seed=n % 97 + leaf)
    saved = ctx.tune_table()
    try:
        entries = [dict(e, sort_bits=bits) for e in saved]
        ctx.set_tune_table(entries)
        say("tree", kind, n, "leaf", leaf, "bits", bits)
        tree = ctx.parallel_qtree(x, y, z, leaf_size=leaf)
        t = tree.read()
        tree.free()
    finally:
        ctx.set_tune_table(saved)
    assert t["key_bits"] == bits and t["leaf_size"] == leaf
    check_tree(t, x, y, z)


def test_planted_mistakes_fail_the_checks(ctx):
    x, y, z = _cloud("clusters", 20_000, seed=3)
    say("planted")
    tree = ctx.parallel_qtree(x, y, z, leaf_size=16)
    t = tree.read()
    tree.free()
    check_tree(t, x, y, z)

    def broken(mutate):
        u = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in t.items()}
        mutate(u)
        with pytest.raises(AssertionError):
            check_tree(u, x, y, z)

    def swap_ids(u):
        a, b = 5, 5000
        for k in ("ids", "xs", "ys", "zs"):
            u[k][[a, b]] = u[k][[b, a]]
    broken(swap_ids)                                                    # order, not records
    broken(lambda u: u["xs"].__setitem__(7, u["xs"][7] + 1e-9))         # a moved record
    broken(lambda u: u.__setitem__("qscale", u["qscale"] * (1 - 1e-6)))  # quantisation
    broken(lambda u: u["leaf_key"].__setitem__(3, u["leaf_key"][3] ^ 1))
    broken(lambda u: u["leaf_min"].__setitem__(9, u["ids"][0]))          # a point of leaf 0
    broken(lambda u: u["node_count"].__setitem__(0, u["node_count"][0] + 1))
    broken(lambda u: u["node_bbox"].__setitem__((4, 1), u["node_bbox"][4, 1] + 1.0))
    broken(lambda u: u["node_min"].__setitem__(2, 0xFFFFFFFF))
    broken(lambda u: u["node_child"].__setitem__((1, 0), u["node_child"][2, 1]))
