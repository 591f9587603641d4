// Device-side data layout and helpers shared by the OWM kernels and the host
// orchestration.  Layout in HBM (all tiles of a batch concatenated):
//   points    x, y, z            fp64 SoA, caller order, tile t = [pt_off, pt_off+n)
//   sorted    xy (double2) + PointZ {z, id, key} (16 B each), Morton order
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

// One point of the sorted cloud: coordinates, batch index and sort key (the
// top key_bits of its Morton code within its tile).  Stored as two 16-byte
// streams, (x, y) and PointZ (z, id, key): every access is one 16-byte load
// or store per lane, coalesced when lanes take consecutive points, and the
// tree sort's digit runs are written at about the speed of a contiguous copy
// (four SoA streams of 8- and 4-byte fields written the same way are about a
// third slower; lanes reading whole 32-byte records use half of each
// fetched span per load).
struct PointRec {
    double x, y, z;
    uint32_t id, key;
};
struct __align__(16) PointZ {
    double z;
    uint32_t id, key;
};
struct PointsDev {
    const double2 *xy;
    const PointZ *zi;
};
struct PointsOut {
    double2 *xy;
    PointZ *zi;
    __host__ __device__ operator PointsDev() const { return PointsDev{xy, zi}; }
};

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
    PointsDev pts;  // the cloud in Morton order; pts.zi[i].id = batch index
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

// (v - vmin) * scale truncated to 32 bits, 0 when not positive (or NaN).
__device__ __forceinline__ uint32_t quantise(double v, double vmin, double scale)
{
    double q = __dmul_rn(__dsub_rn(v, vmin), scale);
    if (!(q > 0.0))
        return 0u;
    if (q >= 4294967295.0)
        return 0xFFFFFFFFu;
    return (uint32_t)q;
}

// A point's Morton code interleaves its quantised x (even bits) and y (odd
// bits), 32 bits per axis over the tile's square bbox; the sort key is its top
// key_bits bits.

// The low 16 bits of v spread to the even bit positions of 32.
__device__ __forceinline__ uint32_t spread16(uint32_t v)
{
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// Sort key of a point: the top bits of its Morton code (shift = 64 - key
// bits, key bits even and at most 32), computed as the interleave of the top
// key_bits / 2 bits of each quantised coordinate (the same bits, with 32-bit
// arithmetic).  The Morton pass (histograms) and the first sort pass (which
// stores it in the point's record) both compute it with this function.
__device__ __forceinline__ uint32_t sort_key_q(double qminx, double qminy, double qscale, double x,
                                               double y, uint32_t shift)
{
    const uint32_t drop = 32u - (64u - shift) / 2u;  // low bits of each axis left out
    const uint32_t hx = (uint32_t)((uint64_t)quantise(x, qminx, qscale) >> drop);
    const uint32_t hy = (uint32_t)((uint64_t)quantise(y, qminy, qscale) >> drop);
    return spread16(hx) | (spread16(hy) << 1);
}

__device__ __forceinline__ uint32_t sort_key(const TileDev &T, double x, double y, uint32_t shift)
{
    return sort_key_q(T.qminx, T.qminy, T.qscale, x, y, shift);
}

// The quantisation frame of a tile's sort keys: its origin and scale.
struct KeyFrame {
    double qminx, qminy, qscale;
};

// A whole point in two 16-byte loads.
__device__ __forceinline__ PointRec owm_load_pt(const PointsDev &P, uint32_t i)
{
    const double2 xy = P.xy[i];
    const uint4 b = *reinterpret_cast<const uint4 *>(P.zi + i);
    PointRec r;
    r.x = xy.x;
    r.y = xy.y;
    r.z = __hiloint2double((int)b.y, (int)b.x);
    r.id = b.z;
    r.key = b.w;
    return r;
}

#endif
