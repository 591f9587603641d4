// Device-side data layout and helpers shared by the OWM kernels and the host
// orchestration.  Layout in HBM (all tiles of a batch concatenated):
//   points    x, y, z            fp64 SoA, caller order, tile t = [pt_off, pt_off+n)
//   sorted    xs, ys, zs, ids    fp64 SoA + u32 original index, Morton order
//   leaves    box (minx,miny,maxx,maxy) f64x4, cnt, start, npts, zkey (u64), zid,
//             parent, key0 (Morton key of the first point)
//   internal  box f64x4, cnt, zkey, zid, child (u32x2, bit 31 = leaf), parent, flag
//   windows   u32 minimum index per window (0xFFFFFFFF = invalid), row-major
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OWM_NONE 0xFFFFFFFFu
#define OWM_LEAF_BIT 0x80000000u
#define OWM_ZKEY_NONE 0xFFFFFFFFFFFFFFFFull

// Per-tile description (device table).  Tree fields depend only on the cloud
// and the leaf size; query fields on the OWM parameters.
struct TileDev {
    // points / tree
    uint32_t pt_off, n;
    uint32_t leaf_off, nleaf;
    uint32_t node_off, leaf_size;
    uint32_t root;      // child reference of the root (OWM_LEAF_BIT | leaf) or internal
    uint32_t grp_off;   // first leaf group (64/leaf_size leaves per wave) of the tile
    uint32_t pad_t;
    double qminx, qminy, qscale;  // quantisation: q = (x - qminx) * qscale in [0, 2^32)
    // query (stages 1-3)
    uint32_t ncols, nrows, ncolsg, nrowsg;
    uint32_t win_off, cell_off, blk_off, nbx;
    uint32_t min_num_points, stage2_need;
    int32_t h, hb;
    uint32_t bsize;
    uint32_t lcs;  // stage 1 reads leaf cell summaries (else every leaf point by point)
    double displace, initx, inity, initxg, inityg;
    double inv_disp;  // 1 / displace (estimates only; exact compares follow)
};

// Offsets of every per-tile range, one array per kind, for binary searches.
struct BatchDev {
    uint32_t ntiles;
    uint32_t np, nleaf, nnode, nwin, ncell, nblk, ngrp;
    uint32_t key_bits;     // Morton bits in the sort key (below the tile bits)
    const TileDev *tiles;
    const uint32_t *pt_off;    // [ntiles+1]
    const uint32_t *leaf_off;  // [ntiles+1]
    const uint32_t *node_off;  // [ntiles+1]
    const uint32_t *win_off;   // [ntiles+1]
    const uint32_t *cell_off;  // [ntiles+1]
    const uint32_t *blk_off;   // [ntiles+1]
    const uint32_t *grp_off;   // [ntiles+1]
    const uint32_t *sort_units;  // [ntiles+1] first sort tile of each tile's segment
};

// Tree arrays of a batch.
struct TreeDev {
    const double *xs, *ys, *zs;
    const uint32_t *ids;  // sorted position -> original (batch-global) index
    const uint32_t *pos;  // original index -> sorted position (resident trees)
    const double *X, *Y;  // original cloud when still available (else pos is used)
    const double4 *lf_box;
    const uint32_t *lf_cnt, *lf_start, *lf_npts, *lf_zid;
    const unsigned long long *lf_zkey;
    const double4 *nd_box;
    const uint32_t *nd_cnt, *nd_zid;
    const unsigned long long *nd_zkey;
    const uint2 *nd_child;
};

// Leaf cell summaries of stage 1: per leaf up to LCS_SLOTS aggregates of runs
// of its points that share a global window-edge class in x and y.
#define LCS_SLOTS 16u
#define LCS_RAW 0xFFFFFFFFu  // the leaf has more classes than slots
struct LeafCells {
    uint32_t *ne;            // [nleaf] summaries of the leaf, or LCS_RAW
    unsigned long long *kc;  // [nleaf * LCS_SLOTS] class key (gy << 16 | gx) | count << 32
    uint32_t *id;            // lowest (z, index) of the summary
    unsigned long long *zk;
};

#ifdef __HIPCC__

__device__ __forceinline__ uint32_t owm_find_tile(const uint32_t *off, uint32_t ntiles, uint32_t v)
{
    // largest t with off[t] <= v (empty tiles are skipped since off[t+1] > v)
    uint32_t lo = 0, hi = ntiles;  // answer in [lo, hi)
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= v)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Order-preserving 64-bit key of an eligible z (z < DBL_MAX): equal doubles
// give equal keys (-0.0 is folded onto +0.0), so (key, index) order is the
// reference's "lowest z, then lowest index" order.
__device__ __forceinline__ unsigned long long owm_zkey(double z)
{
    z = (z == 0.0) ? 0.0 : z;
    unsigned long long b = (unsigned long long)__double_as_longlong(z);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ bool owm_eligible(double z) { return z < 1.7976931348623157e308; }

__device__ __forceinline__ bool owm_less(unsigned long long ka, uint32_t ia, unsigned long long kb,
                                         uint32_t ib)
{
    return ka < kb || (ka == kb && ia < ib);
}

// Window edges of stage 1 (src/optim3_func.cpp:541-561 + makeBox): centre
// initX + ii*Displace, half side (float)(Wsize/2); no fused multiply-add.
__device__ __forceinline__ double owm_wcenter(double init, double disp, uint32_t i)
{
    return __dadd_rn(init, __dmul_rn((double)(int)i, disp));
}

// Stage-3 cell centre: initX + ii*Bsize with the integer product.
__device__ __forceinline__ double owm_gcenter(double init, uint32_t bsize, uint32_t i)
{
    return __dadd_rn(init, (double)(int)(i * bsize));
}

__device__ __forceinline__ uint32_t owm_lane() { return __lane_id(); }

#endif
