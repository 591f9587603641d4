"""Fits the cost model of the statistics-driven tuner (owm_tune_model in
fastowm_amd/csrc/owm_tune.cpp) to trial times written by collect_trials.py.

Model, per point of a tile, in ns, for leaf size L, window block b = (bx, by,
k, summaries) and sort key width S:

    c_leaf / L + c_amp * amp0 + c_coarse * ampc + c_pass * S / 8
      + c_blk[b] + c_read * rf + c_quant * q + c_under * u

over the tile's leaf histogram (statsmaxnum leaves: density rho = points /
side^2 of each leaf, weighted by its points):
  - amp0: mean over points of (1 + s0/ex)(1 + s0/ey) - 1, with s0 = sqrt(L/rho)
    the side of an L-point leaf and ex, ey = bx*D + W - D, by*D + W - D the
    extent of a window block (W = Wsize, D = Displace): how many window blocks
    beyond the first read a point (leaves that straddle block edges);
  - ampc: the same with s = max(s0, tile side / 2^(S/2)) minus amp0: what a
    key too short to order points inside one Morton cell adds (a leaf then
    spans the whole cell);
  - rf: the points a tile's window blocks read over its points, from its window
    grid: blocks of bx x by windows (fewer at the grid's right and top edges)
    each read an extent of ni*D + W - D by nj*D + W - D;
  - with `slots` = workgroups the chip holds at once for the block shape (256
    compute units x 4, 4, 4, 3, 2, 1 workgroups of the shapes below) and
    `blocks` = window blocks in the launch: q = ceil(blocks / slots) * slots /
    blocks - 1, the idle share of the last round of workgroups, and
    u = max(0, slots / blocks - 1), the idle share of a launch that does not
    fill the chip.
The first term counts leaves (leaf records, Karras nodes, aggregate levels),
the fourth the sort passes.  The fit is least squares on the times of each set
about that set's mean (a free intercept per set), so only differences between
configurations of one set inform it.

    python -m fastowm_amd.tune.fit_model fastowm_amd/tune/trials_mi355x.json
"""
from __future__ import annotations

import json
import math
import sys

import numpy as np

LEAVES = (16, 32, 64, 128)
BLOCKS = ((8, 8, 4, 0), (16, 8, 4, 0), (8, 8, 2, 0), (16, 8, 4, 1), (16, 16, 4, 0), (32, 16, 4, 0))
SORTS = (32, 24, 16, 8)
CU = 256
SLOTS = dict(zip(BLOCKS, (4, 4, 4, 3, 2, 1)))  # workgroups per compute unit
N_COEF = 4 + len(BLOCKS) + 3


def density(sample):
    """(rho, weight) of the sample's leaves with points."""
    r0 = float(sample["max_radius"])
    rho, w = [], []
    for level, pts, leaves in sample["bins"]:
        if pts == 0:
            continue
        side = 2.0 * r0 / 2.0 ** level
        rho.append(pts / (side * side))
        w.append(float(pts) * leaves)
    return np.array(rho), np.array(w)


def features(sample, wsize, leaf, block, sort_bits, blocks_in_launch, rho=None, w=None):
    if rho is None:
        rho, w = density(sample)
    bx, by, _, _ = block
    D = float(sample["displace"])
    W = float(wsize)
    ex, ey = bx * D + W - D, by * D + W - D
    bb = sample["bbox"]
    cell = max(bb[1] - bb[0], bb[3] - bb[2]) / 2.0 ** (sort_bits / 2)
    s0 = np.sqrt(leaf / rho)
    s = np.maximum(s0, cell)
    sw = w.sum()
    amp0 = float((w * ((1 + s0 / ex) * (1 + s0 / ey) - 1)).sum() / sw)
    ampc = float((w * ((1 + s / ex) * (1 + s / ey) - 1)).sum() / sw) - amp0
    nc, nr = sample["ncols"], sample["nrows"]
    nbx, nby = math.ceil(nc / bx), math.ceil(nr / by)
    rf = (nc * D + nbx * (W - D)) * (nr * D + nby * (W - D)) / ((nc * D + W - D) * (nr * D + W - D))
    slots = CU * SLOTS[tuple(block)]
    q = math.ceil(blocks_in_launch / slots) * slots / blocks_in_launch - 1.0
    u = max(0.0, slots / blocks_in_launch - 1.0)
    f = [1.0 / leaf, amp0, ampc, sort_bits / 8.0]
    f += [1.0 if b == block else 0.0 for b in BLOCKS]
    f += [rf, q, u]
    return np.array(f)


def blocks(sample, tiles, block):
    return tiles * math.ceil(sample["ncols"] / block[0]) * math.ceil(sample["nrows"] / block[1])


def config_key(leaf, block, sort_bits):
    return "%d,%d,%d,%d,%d,%d" % ((leaf,) + tuple(block) + (sort_bits,))


def set_matrix(rec):
    """Feature rows and times (ns per point) of one set, in config order."""
    rho, w = density(rec["sample"])
    keys, F, t = [], [], []
    points = rec["n"] * rec["tiles"]
    for leaf in LEAVES:
        for block in BLOCKS:
            for sb in SORTS:
                k = config_key(leaf, block, sb)
                keys.append((leaf, block, sb))
                F.append(features(rec["sample"], rec["wsize"], leaf, block, sb,
                                  blocks(rec["sample"], rec["tiles"], block), rho, w))
                t.append(rec["times"][k] / points * 1e6)
    return keys, np.array(F), np.array(t)


def fit(records):
    X, Y = [], []
    for rec in records:
        _, F, t = set_matrix(rec)
        X.append(F - F.mean(0))
        Y.append(t - t.mean())
    coef, *_ = np.linalg.lstsq(np.vstack(X), np.concatenate(Y), rcond=None)
    return coef


def pick(coef, rec):
    keys, F, _ = set_matrix(rec)
    return keys[int(np.argmin(F @ coef))]


def regret(coef, rec):
    leaf, block, sb = pick(coef, rec)
    t = rec["times"]
    return t[config_key(leaf, block, sb)] / min(t.values()) - 1.0


def cpp_block(coef):
    names = ["LEAF", "AMP", "COARSE", "PASS"]
    lines = ["constexpr double M_%s = %.6g;" % (n, c) for n, c in zip(names, coef[:4])]
    blk = ", ".join("%.6g" % coef[4 + j] for j in range(len(BLOCKS)))
    lines.append("constexpr double M_BLK[%d] = {%s};" % (len(BLOCKS), blk))
    lines.append("constexpr double M_READ = %.6g;" % coef[-3])
    lines.append("constexpr double M_QUANT = %.6g;" % coef[-2])
    lines.append("constexpr double M_UNDER = %.6g;" % coef[-1])
    return "\n".join(lines)


def main(argv):
    recs = json.load(open(argv[1]))
    coef = fit(recs)
    rg = np.array([regret(coef, r) for r in recs])
    loo = np.array([regret(fit(recs[:i] + recs[i + 1:]), recs[i]) for i in range(len(recs))])
    dflt = np.array([r["times"][config_key(32, (8, 8, 4, 0), 32)] / min(r["times"].values()) - 1
                     for r in recs])
    print("sets: %d" % len(recs))
    print("regret of the model's pick: mean %.4f max %.4f (leave-one-set-out: mean %.4f max %.4f)"
          % (rg.mean(), rg.max(), loo.mean(), loo.max()))
    print("regret of leaf 32, block 8x8x4, 32-bit keys: mean %.4f max %.4f" % (dflt.mean(), dflt.max()))
    print(cpp_block(coef))


if __name__ == "__main__":
    main(sys.argv)
