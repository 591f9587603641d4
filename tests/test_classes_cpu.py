"""The arithmetic behind stamped window classes, restated in IEEE double on the
CPU and checked against brute force.

A tree built in the same chain as its query stores, for every point, its
global window classes G(x) = #{e < x} + #{e <= x} over all left and right
window edges of its tile (one per axis); stage 1 then bins by them.  The sort
computes G by arithmetic (owm_gaxis_fast in fastowm_amd/csrc/owm_dev.h): one
estimate f = (x - L_0) / displace for the left edges, f - c for the right
edges, and the counts clamp(floor + 1, 0, n) -- claimed exact whenever both
estimates lie farther than tol = 2^-46 M / displace from an integer, with M
bounding |init| and every edge.  Points where the claim is not made are read
by their coordinates instead, so the only requirement is that every claim is
right.  The edges are the device's (window centre init + i*displace, then -+
(float)(Wsize/2), each step rounded, no fused multiply-add); the parameters
come from the oracle's derivation."""
import numpy as np
import pytest

import pyoracle as O


def edges(init, disp, hd, n):
    i = np.arange(n, dtype=np.float64)
    c = init + i * disp  # numpy: two roundings, no fused multiply-add
    return c - hd, c + hd


def brute(x, L, R):
    E = np.sort(np.concatenate([L, R]))
    return np.searchsorted(E, x, side="left") + np.searchsorted(E, x, side="right")


def fast(x, L, R, inv, m):
    """The fast path: (classes, claimed) per point."""
    e0 = L[0]
    c = (R[0] - L[0]) * inv
    tol = np.ldexp(m, -46) * inv
    with np.errstate(invalid="ignore", over="ignore"):
        fl = (x - e0) * inv
        fr = fl - c
        il, ir = np.floor(fl), np.floor(fr)
        dl, dr = fl - il, fr - ir
        n = float(len(L))
        # fmax(NaN, 0) is 0 on the device, as nan_to_num makes it here
        kl = np.nan_to_num(np.clip(il + 1, 0, n)).astype(np.int64)
        kr = np.nan_to_num(np.clip(ir + 1, 0, n)).astype(np.int64)
        g = 2 * (kl + kr)
        ok = (np.minimum(dl, dr) > tol) & (np.maximum(dl, dr) < 1.0 - tol)
    return g, ok


def nudged(v, rng, k):
    """v moved by 0..k ulps either way."""
    out = v.copy()
    steps = rng.integers(-k, k + 1, len(v))
    for s in range(1, k + 1):
        for sign in (1, -1):
            m = steps == sign * s
            w = out[m]
            for _ in range(s):
                w = np.nextafter(w, sign * np.inf)
            out[m] = w
    return out


CASES = [
    # Wsize, Overlap, bbox origin, side (m), points
    (12, 0.8, 0.0, 500.0),
    (12, 0.8, 612_345.0, 2000.0),
    (8, 0.5, 4_512_345.25, 800.0),
    (16, 0.7, -300_000.0, 1500.0),
    (10, 0.9, 0.0, 200.0),
    (5, 0.6, 1_000_000.0, 300.0),
    (2, 0.95, 400_000.0, 60.0),
    (20, 0.99, 5_000_000.0, 400.0),
    (7, 0.75, -1.0e7, 900.0),
]


@pytest.mark.parametrize("ws,ov,origin,side", CASES)
def test_fast_classes_equal_brute_force_where_claimed(ws, ov, origin, side):
    rng = np.random.default_rng(int(ws * 1000 + ov * 100))
    box = (origin, origin + side, origin - side / 3, origin + side * 2 / 3)
    p = O.params(ws, 20, ov, box, 100_000, 32)
    assert p is not None
    hd = float(np.float32(p.h))
    inv = 1.0 / p.displace
    for init, n, lo, hi in ((p.initx, p.ncols, box[0], box[1]), (p.inity, p.nrows, box[2], box[3])):
        L, R = edges(init, p.displace, hd, n)
        m = max(abs(init), abs(L[0]), abs(R[-1])) + hd
        # uniform points over the tile and beyond its window grid
        x = rng.uniform(lo - 2 * ws, hi + 2 * ws, 200_000)
        # points on and a few ulps either side of every edge
        e = np.concatenate([L, R])
        x = np.concatenate([x, nudged(np.repeat(e, 6), rng, 3), e])
        # far out, non-finite
        x = np.concatenate([x, [lo - 1e6, hi + 1e6, -1e300, 1e300, np.inf, -np.inf, np.nan, 0.0, -0.0]])
        g, ok = fast(x, L, R, inv, m)
        want = brute(x, L, R)
        assert np.array_equal(g[ok], want[ok]), np.flatnonzero(ok & (g != want))[:5]
        # NaN is never claimed; the claims cover nearly all uniform points
        assert not ok[np.isnan(x)].any()
        assert ok[:200_000].mean() > 0.999
        # the classes fit 16 bits for grids up to 16383 windows
        assert want.max() <= 4 * n


def test_points_on_edges_are_left_to_coordinates():
    """An edge value itself (odd class) or one ulp from it is never claimed."""
    p = O.params(10, 20, 0.9, (0.0, 200.0, 0.0, 200.0), 40_000, 16)
    hd = float(np.float32(p.h))
    L, R = edges(p.initx, p.displace, hd, p.ncols)
    m = max(abs(p.initx), abs(L[0]), abs(R[-1])) + hd
    x = np.concatenate([L, R, np.nextafter(L, np.inf), np.nextafter(R, -np.inf)])
    g, ok = fast(x, L, R, 1.0 / p.displace, m)
    assert not ok.any()
    # and the brute-force classes of edge values are odd unless two edges meet
    b = brute(np.concatenate([L, R]), L, R)
    coincide = np.isin(np.concatenate([L, R]), np.intersect1d(L, R))
    assert (b[~coincide] % 2 == 1).all()


def fast_columns(x, L, R, inv, mm, m):
    """owm_paxis_fast: (column + 1, claimed) when the window side is m steps."""
    e0 = L[0]
    c = (R[0] - L[0]) * inv
    tol = np.ldexp(mm, -46) * inv
    n = len(L)
    with np.errstate(invalid="ignore", over="ignore"):
        fl = (x - e0) * inv
        fr = fl - c
        il, ir = np.floor(fl), np.floor(fr)
        dl, dr = fl - il, fr - ir
        # fmax(NaN, -1) is -1 on the device
        v = np.nan_to_num(np.clip(il, -1, n + m - 1), nan=-1.0).astype(np.int64) + 1
        ok = (np.minimum(dl, dr) > tol) & (np.maximum(dl, dr) < 1.0 - tol) & (ir == il - m)
    return v, ok


@pytest.mark.parametrize("ws,ov,origin,side", CASES)
def test_window_columns_equal_brute_force_where_claimed(ws, ov, origin, side):
    """Where the column stamp is claimed, the windows c - m + 1 .. c of its
    column c are exactly the windows whose open interval (L_i, R_i) holds x;
    with a window side of a whole number of steps nearly every point is
    claimed, and no point on an edge or NaN ever is."""
    rng = np.random.default_rng(int(ws * 7 + ov * 1000))
    box = (origin, origin + side, origin - side / 3, origin + side * 2 / 3)
    p = O.params(ws, 20, ov, box, 100_000, 32)
    hd = float(np.float32(p.h))
    inv = 1.0 / p.displace
    for init, n, lo, hi in ((p.initx, p.ncols, box[0], box[1]), (p.inity, p.nrows, box[2], box[3])):
        L, R = edges(init, p.displace, hd, n)
        mm = max(abs(init), abs(L[0]), abs(R[-1])) + hd
        m = int(round((R[0] - L[0]) * inv))
        x = rng.uniform(lo - 2 * ws, hi + 2 * ws, 200_000)
        e = np.concatenate([L, R])
        x = np.concatenate([x, nudged(np.repeat(e, 6), rng, 3), e,
                            [lo - 1e6, hi + 1e6, -1e300, 1e300, np.inf, -np.inf, np.nan]])
        v, ok = fast_columns(x, L, R, inv, mm, m)
        col = v - 1
        # windows holding x by brute force: #{R <= x} .. #{L < x} - 1
        w0 = np.searchsorted(R, x, side="right")
        w1 = np.searchsorted(L, x, side="left")
        c0 = np.clip(col - m + 1, 0, n)
        c1 = np.clip(col + 1, 0, n)
        empty = w0 >= w1
        same = np.where(empty, c0 >= c1, (c0 == w0) & (c1 == w1))
        assert same[ok].all(), np.flatnonzero(ok & ~same)[:5]
        assert not ok[np.isnan(x)].any()
        assert not ok[np.isin(x, e)].any()
        assert (v[ok] >= 0).all() and (v[ok] <= n + m).all()
        if abs((R[0] - L[0]) * inv - m) < 1e-6:
            assert ok[:200_000].mean() > 0.999
