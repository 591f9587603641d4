// Device-side data layout and helpers shared by the OWM kernels and the host
// orchestration.  Layout in HBM (all tiles of a batch concatenated):
//   points    x, y, z            fp64 SoA, caller order, tile t = [pt_off, pt_off+n)
//   sorted    xy (double2) + PointZ {z, id, aux} (16 B each), Morton order
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

// One point of the sorted cloud: coordinates, batch index and one more word
// (aux).  Between sort passes aux is the point's sort key (the top key_bits
// of its Morton code within its tile); in a tree's final order it is the
// key, or, when the tile's gcls is set, the point's global window-edge
// classes for the query planned with the tree (owm_gstamp).  Stored as two
// 16-byte streams, (x, y) and PointZ (z, id, aux): every access is one 16-byte load
// or store per lane, coalesced when lanes take consecutive points, and the
// tree sort's digit runs are written at about the speed of a contiguous copy
// (four SoA streams of 8- and 4-byte fields written the same way are about a
// third slower; lanes reading whole 32-byte records use half of each
// fetched span per load).
struct PointRec {
    double x, y, z;
    uint32_t id, aux;
};
struct __align__(16) PointZ {
    double z;
    uint32_t id, aux;
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
    // 1: the final sort pass writes each point's global window-edge classes
    // (for this tile's query fields) into aux and stage 1 reads them instead
    // of x and y; only set when the tree is built in the same call as the
    // query and the classes fit 16 bits per axis
    uint32_t gcls;
    // m when the window side is m window steps (every right edge R_i within
    // rounding of the left edge L_{i+m}) and stage 1 runs on window columns:
    // the records then carry each point's column per axis instead of its
    // classes (owm_cstamp); 0 otherwise
    uint32_t pairs;
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
    // global window classes, per axis (x, y): the first left edge and the
    // right edges' offset in steps ((R_0 - L_0) / displace, rounded); the
    // tolerance 2^-46 M / displace for the larger axis bound
    // M = max(|init|, |first left edge|, |last right edge|) + half side
    double ge0[2], gc[2], gtol;
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

// Global window-edge classes.  Over the multiset of all left edges L_i and
// right edges R_i (i < n) of a tile's windows along one axis,
// G(x) = #{e < x} + #{e <= x}.  For every edge e of the multiset,
// e < x <=> G(e) < G(x) and e <= x <=> G(e) <= G(x), so a window block's
// class of x (the same count over the block's own edges, a subset) and
// "strictly inside the block" are functions of G(x) alone.  G(NaN) = 0.
__device__ __forceinline__ double owm_edge_at(double init, double disp, double hd, int i, bool right)
{
    const double c = owm_wcenter(init, disp, (uint32_t)i);
    return right ? __dadd_rn(c, hd) : __dsub_rn(c, hd);
}

// 2 #{edges of one side < x} + [x is such an edge]; the edges of one side
// are strictly increasing in the window index.  The estimate from the step
// is exact but for rounding near an edge, so two exact edge compares confirm
// it and a walk fixes it otherwise.
__device__ __forceinline__ uint32_t owm_gside(double x, double init, double disp, double inv,
                                              double hd, bool right, uint32_t n)
{
    const double f = (x - (right ? init + hd : init - hd)) * inv;
    int k = f >= 0.0 ? (f < (double)n ? (int)f + 1 : (int)n) : 0;
    const double elo = owm_edge_at(init, disp, hd, k > 0 ? k - 1 : 0, right);
    double ehi = owm_edge_at(init, disp, hd, k < (int)n ? k : (int)n - 1, right);
    if (!((k == 0 || elo < x) && (k == (int)n || !(ehi < x)))) {
        while (k > 0 && !(owm_edge_at(init, disp, hd, k - 1, right) < x))
            --k;
        while (k < (int)n && owm_edge_at(init, disp, hd, k, right) < x)
            ++k;
        ehi = owm_edge_at(init, disp, hd, k < (int)n ? k : (int)n - 1, right);
    }
    return 2u * (uint32_t)k + ((k < (int)n && ehi == x) ? 1u : 0u);
}

__device__ __forceinline__ uint32_t owm_gclass(double x, double init, double disp, double inv,
                                               double hd, uint32_t n)
{
    return owm_gside(x, init, disp, inv, hd, false, n) + owm_gside(x, init, disp, inv, hd, true, n);
}

// One axis of owm_gclass by arithmetic, when that is exact.  The edges of one
// side are E_i = fl(fl(init + fl(i*disp)) +- hd), so with M bounding |init|
// and every edge, E_i = E_0 + i*disp + e_i with |e_i| <= 2^-51 M, and for
// |x| <= M, fl(fl(x - E_0) * inv) is within 3 * 2^-52 M / disp of
// (x - E_0) / disp; the right side's estimate f_L - c (c = (R_0 - L_0) * inv,
// rounded) adds at most 2^-50 M / disp.  When an estimate is farther than
// tol = 2^-46 M / disp from every integer (four times the sum of the bounds),
// x lies strictly between E_floor(f) and E_floor(f)+1 (or beyond the first or
// last edge), so that side contributes 2 clamp(floor(f) + 1, 0, n).  For
// |x| > M, x lies beyond every edge, at least half a window step beyond the
// outermost (M exceeds it by the half side), so the estimate clamps to the
// exact 0 or n.  Returns false where the estimate is too close to an integer
// (and for NaN).  e0 = L_0, c and tol are the tile's constants.
__device__ __forceinline__ bool owm_gaxis_fast(double x, double e0, double c, double tol, double inv,
                                               uint32_t n, uint32_t &g)
{
    const double fl = (x - e0) * inv, fr = fl - c;
    const double il = floor(fl), ir = floor(fr);
    const double dl = fl - il, dr = fr - ir;
    const double nn = (double)n;
    g = 2u * ((uint32_t)fmin(fmax(il + 1.0, 0.0), nn) + (uint32_t)fmin(fmax(ir + 1.0, 0.0), nn));
    return fmin(dl, dr) > tol && fmax(dl, dr) < 1.0 - tol;
}

// The window column of x when a tile's window side is m steps: the left edges
// cut the axis into columns [L_c, L_{c+1}), and when #{R <= x} = #{L <= x} - m
// (every point but those within rounding of a right edge and the left edge it
// pairs with) the windows holding x are exactly c - m + 1 .. c.  With both
// estimates of owm_gaxis_fast exact, clamp(floor(fl) + 1, 0, n) = #{L <= x}
// and clamp(floor(fr) + 1, 0, n) = #{R <= x}, so floor(fr) = floor(fl) - m
// gives the pairing and floor(fl) the column; the column is clamped to
// [-1, n + m - 1] (a column beyond either end lies in no window, as does
// every column past those) and stored + 1.  Returns false where the
// estimates are not exact, the pairing fails or x is NaN.
__device__ __forceinline__ bool owm_paxis_fast(double x, double e0, double c, double tol, double inv,
                                               uint32_t n, uint32_t m, uint32_t &v)
{
    const double fl = (x - e0) * inv, fr = fl - c;
    const double il = floor(fl), ir = floor(fr);
    const double dl = fl - il, dr = fr - ir;
    v = (uint32_t)(fmin(fmax(il, -1.0), (double)(n + m - 1)) + 1.0);
    return fmin(dl, dr) > tol && fmax(dl, dr) < 1.0 - tol && ir == il - (double)m;
}

// Most windows per axis for which a tile's classes are stamped (G <= 4n).
#define OWM_GCLS_MAX_WINDOWS 16383u

// (gy << 16 | gx) of a point for a tile's query fields (at most
// OWM_GCLS_MAX_WINDOWS windows per axis, so that G <= 4n <= 0xFFFC fits 16
// bits), or OWM_GCLS_NONE where the arithmetic of owm_gaxis_fast is not exact
// (the reader then classifies the point by its coordinates).  The tile's g*
// fields hold the constants per axis (set with its query fields).
#define OWM_GCLS_NONE 0xFFFFFFFFu
template <class TD>
__device__ __forceinline__ uint32_t owm_gstamp(const TD &T, double x, double y)
{
    uint32_t gx, gy;
    const bool okx = owm_gaxis_fast(x, T.ge0[0], T.gc[0], T.gtol, T.inv_disp, T.ncols, gx);
    const bool oky = owm_gaxis_fast(y, T.ge0[1], T.gc[1], T.gtol, T.inv_disp, T.nrows, gy);
    return okx && oky ? gy << 16 | gx : OWM_GCLS_NONE;
}

// The same with a tile's window columns + 1 (pairs = m; owm_paxis_fast, at
// most n + m <= 16390 per axis), or OWM_GCLS_NONE.
template <class TD>
__device__ __forceinline__ uint32_t owm_cstamp(const TD &T, double x, double y)
{
    uint32_t cx, cy;
    const bool okx = owm_paxis_fast(x, T.ge0[0], T.gc[0], T.gtol, T.inv_disp, T.ncols, T.pairs, cx);
    const bool oky = owm_paxis_fast(y, T.ge0[1], T.gc[1], T.gtol, T.inv_disp, T.nrows, T.pairs, cy);
    return okx && oky ? cy << 16 | cx : OWM_GCLS_NONE;
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
    r.aux = b.w;
    return r;
}

#endif
