// Spatial index of the OWM pipeline on gfx950: bbox reduction, 2D Morton
// keys (the record sort in sort.hip then leaves the cloud in Morton order),
// per-leaf aggregates, Karras LBVH over the leaves and an atomic-flag
// bottom-up pass that propagates each node's tight bbox, point count and
// lowest (z, index).
//
// Reference counterparts: ordered/cuda/ordered_builder.inl (get_morton_code
// :33-58, order_points :60-74, init_leafs_size :78-126, init_nodes5 :129-205,
// create_aabb_size :212-254) and the CPU quadtree + storeMinAndNumPoints of
// CPP-TBB-CPU/src/optim3_func.cpp:218-307.  Differences by design: 64-bit
// Morton codes (32 bits per axis) over the square bbox, whose top bits form the
// sort key; Karras splits over the leaves' Morton keys (spatial splits) rather
// than over leaf indices; tight point bboxes (not cell geometry), so a node
// whose bbox lies strictly inside a window can be taken whole, exactly; the
// (z, index) minimum replaces the reference's 1e-10 tolerance rules.
#include "owm_kernels.h"

namespace owm {

constexpr int TB = 256;
constexpr uint32_t CHUNK = 8192;

__device__ __forceinline__ unsigned long long dkey(double v)
{
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double wave_min_d(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ double wave_max_d(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ uint32_t wave_sum_u(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// lexicographic (key, id) minimum over the wave
__device__ __forceinline__ void wave_min_kid(unsigned long long &k, uint32_t &id)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long k2 = __shfl_xor(k, o, 64);
        uint32_t i2 = __shfl_xor(id, o, 64);
        if (owm_less(k2, i2, k, id)) {
            k = k2;
            id = i2;
        }
    }
}

// The bbox reduction spreads its atomics over BBOX_REPL copies of every
// tile's four words (copy = block index mod BBOX_REPL) so that thousands of
// blocks do not serialise on four addresses; k_bbox_fold folds the copies.
constexpr uint32_t BBOX_REPL = 16;

__global__ void k_bbox_init(unsigned long long *bk, uint32_t nwords)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += gridDim.x * blockDim.x)
        bk[i] = (i & 2) ? 0ull : ~0ull;  // words: min x, min y, max x, max y
}

__device__ __forceinline__ void bbox_flush(unsigned long long *bk, uint32_t t, double mnx, double mny,
                                           double mxx, double mxy)
{
    if (mnx <= mxx) {
        atomicMin(&bk[4 * t + 0], dkey(mnx));
        atomicMax(&bk[4 * t + 2], dkey(mxx));
    }
    if (mny <= mxy) {
        atomicMin(&bk[4 * t + 1], dkey(mny));
        atomicMax(&bk[4 * t + 3], dkey(mxy));
    }
}

// Block-wide reduction of the four accumulators, flushed by thread 0.
__device__ __forceinline__ void bbox_block_flush(unsigned long long *bk, uint32_t t, double mnx,
                                                 double mny, double mxx, double mxy,
                                                 double (*red)[TB / 64])
{
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    mnx = wave_min_d(mnx);
    mny = wave_min_d(mny);
    mxx = wave_max_d(mxx);
    mxy = wave_max_d(mxy);
    if (lane == 0) {
        red[0][w] = mnx;
        red[1][w] = mny;
        red[2][w] = mxx;
        red[3][w] = mxy;
    }
    __syncthreads();
    if (tid == 0) {
        for (int k = 1; k < TB / 64; ++k) {
            mnx = fmin(mnx, red[0][k]);
            mny = fmin(mny, red[1][k]);
            mxx = fmax(mxx, red[2][k]);
            mxy = fmax(mxy, red[3][k]);
        }
        bbox_flush(bk, t, mnx, mny, mxx, mxy);
    }
    __syncthreads();
}

// Exact fp64 min/max of x and y per tile (NaNs ignored).  A block keeps its
// accumulators across consecutive chunks of one tile and flushes once per tile.
__global__ __launch_bounds__(TB) void k_bbox(BatchDev B, const double *__restrict__ X,
                                              const double *__restrict__ Y, unsigned long long *bk_all)
{
    __shared__ double red[4][TB / 64];
    unsigned long long *bk = bk_all + (size_t)(blockIdx.x % BBOX_REPL) * 4 * B.ntiles;
    const uint32_t nch = (B.np + CHUNK - 1) / CHUNK;
    const uint32_t tid = threadIdx.x;
    uint32_t cur = OWM_NONE;  // tile of the accumulators (uniform over the block)
    double mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        uint32_t c0 = c * CHUNK, c1 = min(c0 + CHUNK, B.np);
        uint32_t t0 = owm_find_tile(B.pt_off, B.ntiles, c0);
        uint32_t t1 = owm_find_tile(B.pt_off, B.ntiles, c1 - 1);
        if (t0 != cur && cur != OWM_NONE) {
            bbox_block_flush(bk, cur, mnx, mny, mxx, mxy, red);
            mnx = mny = INFINITY;
            mxx = mxy = -INFINITY;
            cur = OWM_NONE;
        }
        if (t0 != t1) {
            // a chunk across tile boundaries: per-thread flushes (rare)
            double ax = INFINITY, ay = INFINITY, bx = -INFINITY, by = -INFINITY;
            uint32_t t = t0, next = B.pt_off[t0 + 1];
            for (uint32_t i = c0 + tid; i < c1; i += TB) {
                while (i >= next) {
                    bbox_flush(bk, t, ax, ay, bx, by);
                    ax = ay = INFINITY;
                    bx = by = -INFINITY;
                    ++t;
                    next = B.pt_off[t + 1];
                }
                double x = X[i], y = Y[i];
                ax = fmin(ax, x);
                bx = fmax(bx, x);
                ay = fmin(ay, y);
                by = fmax(by, y);
            }
            bbox_flush(bk, t, ax, ay, bx, by);
            continue;
        }
        cur = t0;
        // eight independent loads per array in flight per thread
        for (uint32_t base = c0 + tid; base < c1; base += 8 * TB) {
            double xv[8], yv[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                uint32_t i = base + k * TB;
                xv[k] = i < c1 ? X[i] : NAN;
                yv[k] = i < c1 ? Y[i] : NAN;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                mnx = fmin(mnx, xv[k]);
                mxx = fmax(mxx, xv[k]);
                mny = fmin(mny, yv[k]);
                mxy = fmax(mxy, yv[k]);
            }
        }
    }
    if (cur != OWM_NONE)
        bbox_block_flush(bk, cur, mnx, mny, mxx, mxy, red);
}

// Fold the copies of every tile's words into copy 0.
__global__ void k_bbox_fold(unsigned long long *bk, uint32_t ntiles)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 4 * ntiles; i += gridDim.x * blockDim.x) {
        unsigned long long v = bk[i];
        const bool is_max = (i & 2) != 0;
        for (uint32_t r = 1; r < BBOX_REPL; ++r) {
            unsigned long long u = bk[(size_t)r * 4 * ntiles + i];
            v = is_max ? (u > v ? u : v) : (u < v ? u : v);
        }
        bk[i] = v;
    }
}

__device__ __forceinline__ unsigned long long spread2(uint32_t v)
{
    unsigned long long x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

__device__ __forceinline__ uint32_t quantise(double v, double vmin, double scale)
{
    double q = __dmul_rn(__dsub_rn(v, vmin), scale);
    if (!(q > 0.0))
        return 0u;
    if (q >= 4294967295.0)
        return 0xFFFFFFFFu;
    return (uint32_t)q;
}

// 64-bit 2D Morton code: 32 bits per axis over the tile's square bbox,
// x in the even bits, y in the odd bits.
__device__ __forceinline__ unsigned long long morton64(const TileDev &T, double x, double y)
{
    uint32_t qx = quantise(x, T.qminx, T.qscale);
    uint32_t qy = quantise(y, T.qminy, T.qscale);
    return spread2(qx) | (spread2(qy) << 1);
}

// Sort key of a point: the top bits of its Morton code (shift = 64 - key bits).
// The tree pass recomputes a leaf's first key from its first point with this
// same function, so the last sort pass need not write keys.
__device__ __forceinline__ uint32_t sort_key(const TileDev &T, double x, double y, uint32_t shift)
{
    return (uint32_t)(morton64(T, x, y) >> shift);
}

// Sort keys (the top key_bits of each point's Morton code within its tile)
// plus each tile's digit histograms of every radix pass.
__global__ __launch_bounds__(TB) void k_morton(BatchDev B, const double *__restrict__ X,
                                                const double *__restrict__ Y,
                                                uint32_t *__restrict__ keys, uint32_t *hist, int passes)
{
    // digit counts of the tile `cur` (uniform over the block), flushed to the
    // tile's histogram when the block moves to another tile
    __shared__ uint32_t h[4][256];
    for (int i = threadIdx.x; i < 4 * 256; i += TB)
        (&h[0][0])[i] = 0;
    uint32_t cur = OWM_NONE;
    auto flush = [&]() {
        __syncthreads();
        uint32_t *g = hist + (size_t)cur * SORT_SEG_HIST;
        for (int i = threadIdx.x; i < passes * 256; i += TB) {
            const uint32_t c = (&h[0][0])[i];
            if (c) {
                atomicAdd(&g[i], c);
                (&h[0][0])[i] = 0;
            }
        }
        __syncthreads();
        cur = OWM_NONE;
    };
    __syncthreads();
    const uint32_t nch = (B.np + CHUNK - 1) / CHUNK;
    const uint32_t shift = 64 - B.key_bits;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const uint32_t c0 = c * CHUNK, c1 = min(c0 + CHUNK, B.np);
        uint32_t t = owm_find_tile(B.pt_off, B.ntiles, c0);
        uint32_t next = B.pt_off[t + 1];
        const bool single = c1 <= next;
        if (cur != OWM_NONE && (!single || t != cur))
            flush();
        if (single) {
            // one tile: eight independent point loads in flight per thread
            cur = t;
            const TileDev &T = B.tiles[t];
            for (uint32_t base = c0 + threadIdx.x; base < c1; base += 8 * TB) {
                double xv[8], yv[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint32_t i = base + j * TB;
                    xv[j] = i < c1 ? X[i] : 0.0;
                    yv[j] = i < c1 ? Y[i] : 0.0;
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint32_t i = base + j * TB;
                    if (i < c1) {
                        const uint32_t k = sort_key(T, xv[j], yv[j], shift);
                        keys[i] = k;
                        for (int p = 0; p < passes; ++p)
                            atomicAdd(&h[p][(k >> (8 * p)) & 255u], 1u);
                    }
                }
            }
            continue;
        }
        // a chunk across tiles: counts go straight to each tile's histogram
        for (uint32_t i = c0 + threadIdx.x; i < c1; i += TB) {
            while (i >= next)
                next = B.pt_off[++t + 1];
            const TileDev &T = B.tiles[t];
            const double x = X[i], y = Y[i];
            const uint32_t k = sort_key(T, x, y, shift);
            keys[i] = k;
            uint32_t *g = hist + (size_t)t * SORT_SEG_HIST;
            for (int p = 0; p < passes; ++p)
                atomicAdd(&g[p * 256 + ((k >> (8 * p)) & 255u)], 1u);
        }
    }
    if (cur != OWM_NONE)
        flush();
}

// Segmented reductions over contiguous lane segments of length ls (a segment
// never spans more than one leaf): after the loop the first lane of each
// segment holds the segment's result.
template <typename F>
__device__ __forceinline__ void seg_reduce(uint32_t off, uint32_t ls, F &&step)
{
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1)
        step(o, off + o < ls);
}

// Exchange with the lane whose index differs in bit log2(STEP), or (for
// STEP 4 and 8) with the mirrored lane of the 8- or 16-lane row: after the
// steps 1..STEP every lane of an aligned segment of 2*STEP lanes has combined
// all of them, which is all a commutative reduction needs.  Steps up to 8
// are DPP moves on the VALU, 16 a swizzle, 32 a permute.
template <int STEP>
__device__ __forceinline__ uint32_t xchg_u(uint32_t v)
{
    if constexpr (STEP == 1)
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);  // quad [1,0,3,2]
    else if constexpr (STEP == 2)
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);  // quad [2,3,0,1]
    else if constexpr (STEP == 4)
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false);  // row half mirror
    else if constexpr (STEP == 8)
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, false);  // row mirror
    else if constexpr (STEP == 16)
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F);  // xor 16 within 32
    else
        return (uint32_t)__shfl_xor((int)v, 32, 64);
}

template <int STEP>
__device__ __forceinline__ unsigned long long xchg_u64(unsigned long long v)
{
    uint32_t lo = xchg_u<STEP>((uint32_t)v), hi = xchg_u<STEP>((uint32_t)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

template <int STEP>
__device__ __forceinline__ double xchg_d(double v)
{
    return __longlong_as_double((long long)xchg_u64<STEP>((unsigned long long)__double_as_longlong(v)));
}

struct LeafAgg {
    double mnx, mny, mxx, mxy;
    uint32_t cnt, bid;
    unsigned long long bk;
};

template <int STEP>
__device__ __forceinline__ void leaf_agg_step(LeafAgg &a)
{
    const double b0 = xchg_d<STEP>(a.mnx), b1 = xchg_d<STEP>(a.mny);
    const double b2 = xchg_d<STEP>(a.mxx), b3 = xchg_d<STEP>(a.mxy);
    const uint32_t c = xchg_u<STEP>(a.cnt), i2 = xchg_u<STEP>(a.bid);
    const unsigned long long k2 = xchg_u64<STEP>(a.bk);
    a.mnx = fmin(a.mnx, b0);
    a.mny = fmin(a.mny, b1);
    a.mxx = fmax(a.mxx, b2);
    a.mxy = fmax(a.mxy, b3);
    a.cnt += c;
    if (owm_less(k2, i2, a.bk, a.bid)) {
        a.bk = k2;
        a.bid = i2;
    }
}

// Reduce over aligned segments of ls lanes (ls a power of two <= 64); every
// lane of a segment ends with its result.  Must run with the whole wave active.
__device__ __forceinline__ void leaf_agg_reduce(LeafAgg &a, uint32_t ls)
{
    if (ls > 1)
        leaf_agg_step<1>(a);
    if (ls > 2)
        leaf_agg_step<2>(a);
    if (ls > 4)
        leaf_agg_step<4>(a);
    if (ls > 8)
        leaf_agg_step<8>(a);
    if (ls > 16)
        leaf_agg_step<16>(a);
    if (ls > 32)
        leaf_agg_step<32>(a);
}

__device__ __forceinline__ void leaf_agg_add(LeafAgg &a, double x, double y, double z, uint32_t v)
{
    if (!isnan(x) && !isnan(y)) {
        ++a.cnt;
        a.mnx = fmin(a.mnx, x);
        a.mxx = fmax(a.mxx, x);
        a.mny = fmin(a.mny, y);
        a.mxy = fmax(a.mxy, y);
        if (owm_eligible(z)) {
            unsigned long long zk = owm_zkey(z);
            if (owm_less(zk, v, a.bk, a.bid)) {
                a.bk = zk;
                a.bid = v;
            }
        }
    }
}

__device__ __forceinline__ LeafAgg leaf_agg_empty()
{
    return LeafAgg{INFINITY, INFINITY, -INFINITY, -INFINITY, 0u, OWM_NONE, OWM_ZKEY_NONE};
}

__device__ __forceinline__ void leaf_store(const TileDev &TD, TreeOut &T, uint32_t key0, uint32_t jl,
                                           uint32_t start, uint32_t npts, const LeafAgg &a)
{
    const uint32_t leaf = TD.leaf_off + jl;
    T.lf_box[leaf] = make_double4(a.mnx, a.mny, a.mxx, a.mxy);
    T.lf_cnt[leaf] = a.cnt;
    T.lf_start[leaf] = start;
    T.lf_npts[leaf] = npts;
    T.lf_zkey[leaf] = a.bk;
    T.lf_zid[leaf] = a.bid;
    T.lf_key0[leaf] = key0;
    T.lf_parent[leaf] = OWM_NONE;
}

constexpr int LEAF_G = 8;  // leaf groups a wave reduces at once

template <typename V>
__device__ __forceinline__ V shfl_any(V v, int src)
{
    if constexpr (sizeof(V) == 8) {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)u, src, 64);
        const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(u >> 32), src, 64);
        return __builtin_bit_cast(V, ((unsigned long long)hi << 32) | lo);
    } else {
        return __builtin_bit_cast(V, __shfl(__builtin_bit_cast(int, v), src, 64));
    }
}

// Leaf sizes 16, 32, 64 and 128: the wave reads its LEAF_G groups in rounds
// of 512 points, a segment of 8 lanes taking 64 consecutive points (8 per
// lane, each load a 64-byte run per segment).  Each lane first folds its own
// points of every leaf, then three butterflies inside the segment (four when a
// leaf spans two segments) finish the leaves.  Lane k of a segment reads its
// points of the segment's k-th leaf rotated by k, so that it holds that
// leaf's first point and stores the leaf.
template <int LL>
__device__ __forceinline__ void leaf_rounds(const TileDev &TD, TreeOut &T, uint32_t g0,
                                            uint32_t lane, uint32_t shift)
{
    constexpr int NK = LL <= 64 ? 64 / LL : 1;  // leaves of a lane's points
    constexpr int QL = 8 / NK;                  // a lane's points per leaf
    constexpr int GPTS = LL <= 64 ? 64 : LL;    // points per group
    const uint32_t seg = lane >> 3, sub = lane & 7;
#pragma unroll 1
    for (int r = 0; r < GPTS / 64; ++r) {
        const uint32_t e0 = (g0 - TD.grp_off) * GPTS + r * 512 + seg * 64;
        uint32_t v[8];
        double x[8], y[8], z[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t k = q / QL;
            const uint32_t e = e0 + q * 8 + ((sub - k) & 7);
            // loads clamped to the tile's last point (no branch per load)
            const uint32_t p = TD.pt_off + min(e, TD.n - 1);
            v[q] = T.ids[p];
            x[q] = T.xs[p];
            y[q] = T.ys[p];
            z[q] = T.zs[p];
        }
        LeafAgg a[NK];
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            a[k] = leaf_agg_empty();
#pragma unroll
            for (int q = k * QL; q < (k + 1) * QL; ++q)
                if (e0 + q * 8 + ((sub - k) & 7) < TD.n)
                    leaf_agg_add(a[k], x[q], y[q], z[q], v[q]);
            leaf_agg_step<1>(a[k]);
            leaf_agg_step<2>(a[k]);
            leaf_agg_step<4>(a[k]);
            if (LL > 64)
                leaf_agg_step<8>(a[k]);
        }
        LeafAgg mine = a[0];
        double fx = x[0], fy = y[0];
#pragma unroll
        for (int k = 1; k < NK; ++k)
            if (sub == (uint32_t)k) {
                mine = a[k];
                fx = x[k * QL];
                fy = y[k * QL];
            }
        const uint32_t jl = e0 / LL + sub;
        const bool store = LL <= 64 ? sub < (uint32_t)NK : (sub == 0 && (seg & 1) == 0);
        if (store && jl < TD.nleaf)
            leaf_store(TD, T, sort_key(TD, fx, fy, shift), jl, TD.pt_off + jl * LL,
                       min((uint32_t)LL, TD.n - jl * LL), mine);
    }
}

// Per-leaf aggregates over the points in Morton order.  A wave owns groups of
// 64/leaf_size consecutive leaves of one tile (one leaf when leaf_size > 64,
// its points strided over the lanes): every lane reads one point per leaf
// group and segmented butterflies reduce each leaf's tight bbox, count
// (non-NaN x and y) and lowest (z, index).  With a power-of-two leaf size of at
// most 64 a wave takes LEAF_G groups at once, so that all their loads are in
// flight together, and the leaves' records are then written by consecutive
// lanes (one store per field for all of them).  The group index is uniform over
// the wave, so the tile lookup and the tile's fields are scalar loads.
__global__ __launch_bounds__(TB) void k_leaf_aggs(BatchDev B, TreeOut T)
{
    const uint32_t shift = 64 - B.key_bits;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * (TB / 64);
    for (uint32_t g0 = (blockIdx.x * (TB / 64) + wv) * LEAF_G; g0 < B.ngrp; g0 += nw * LEAF_G) {
        const uint32_t t = owm_find_tile(B.grp_off, B.ntiles, g0);
        const TileDev &TD = B.tiles[t];
        const uint32_t L = TD.leaf_size;
        if (g0 + LEAF_G <= B.grp_off[t + 1] && (L == 16 || L == 32 || L == 64 || L == 128)) {
            if (L == 16)
                leaf_rounds<16>(TD, T, g0, lane, shift);
            else if (L == 32)
                leaf_rounds<32>(TD, T, g0, lane, shift);
            else if (L == 64)
                leaf_rounds<64>(TD, T, g0, lane, shift);
            else
                leaf_rounds<128>(TD, T, g0, lane, shift);
            continue;
        }
        if (L <= 64 && (L & (L - 1)) == 0 && g0 + LEAF_G <= B.grp_off[t + 1]) {
            const uint32_t nl_grp = 64 / L, sub = lane / L, off = lane & (L - 1);
            const uint32_t jl0 = (g0 - TD.grp_off) * nl_grp;  // first leaf of the groups
            uint32_t v[LEAF_G];
            double x[LEAF_G], y[LEAF_G], z[LEAF_G];
            bool have[LEAF_G];
#pragma unroll
            for (int q = 0; q < LEAF_G; ++q) {
                const uint32_t jl = jl0 + q * nl_grp + sub, e = jl * L + off;
                have[q] = jl < TD.nleaf && e < TD.n;
                // loads clamped to the tile's last point (no branch per load)
                const uint32_t p = TD.pt_off + min(e, TD.n - 1);
                v[q] = T.ids[p];
                x[q] = T.xs[p];
                y[q] = T.ys[p];
                z[q] = T.zs[p];
            }
            // lane m ends with the aggregate of leaf jl0 + m when the groups hold
            // at most 64 leaves (leaf_size >= 4); smaller leaves are stored
            // group by group from each leaf's first lane
            const bool gather = LEAF_G * nl_grp <= 64;
            LeafAgg mine = leaf_agg_empty();
            double mine_x = 0.0, mine_y = 0.0;  // the leaf's first point
#pragma unroll
            for (int q = 0; q < LEAF_G; ++q) {
                LeafAgg a = leaf_agg_empty();
                if (have[q])
                    leaf_agg_add(a, x[q], y[q], z[q], v[q]);
                leaf_agg_reduce(a, L);
                if (!gather) {
                    const uint32_t jl = jl0 + q * nl_grp + sub;
                    const uint32_t k0 = off == 0 && have[q] ? sort_key(TD, x[q], y[q], shift) : 0u;
                    if (off == 0 && jl < TD.nleaf)
                        leaf_store(TD, T, k0, jl, TD.pt_off + jl * L, min(L, TD.n - jl * L), a);
                    continue;
                }
                const bool take = lane / nl_grp == (uint32_t)q;
                if (nl_grp > 1) {
                    const int src = (int)((lane & (nl_grp - 1)) * L);
                    a.mnx = shfl_any(a.mnx, src);
                    a.mny = shfl_any(a.mny, src);
                    a.mxx = shfl_any(a.mxx, src);
                    a.mxy = shfl_any(a.mxy, src);
                    a.cnt = shfl_any(a.cnt, src);
                    a.bid = shfl_any(a.bid, src);
                    a.bk = shfl_any(a.bk, src);
                }
                const int first = nl_grp > 1 ? (int)((lane & (nl_grp - 1)) * L) : 0;
                const double fx = shfl_any(x[q], first), fy = shfl_any(y[q], first);
                if (take) {
                    mine = a;
                    mine_x = fx;
                    mine_y = fy;
                }
            }
            // every leaf's first key from its first point, once per wave
            const uint32_t mine_key = sort_key(TD, mine_x, mine_y, shift);
            const uint32_t jl = jl0 + lane;
            if (gather && lane < LEAF_G * nl_grp && jl < TD.nleaf)
                leaf_store(TD, T, mine_key, jl, TD.pt_off + jl * L, min(L, TD.n - jl * L), mine);
            continue;
        }
        // any leaf size, one group at a time
        for (uint32_t g = g0; g < min(g0 + LEAF_G, B.ngrp); ++g) {
            const TileDev &TG = B.tiles[owm_find_tile(B.grp_off, B.ntiles, g)];
            const uint32_t LG = TG.leaf_size;
            const uint32_t nl_grp = LG <= 64 ? 64 / LG : 1;
            const uint32_t ls = LG <= 64 ? LG : 64;
            const uint32_t sub = lane / ls, off = lane - sub * ls;
            const uint32_t jl = (g - TG.grp_off) * nl_grp + sub;
            const bool active = sub < nl_grp && jl < TG.nleaf;
            const uint32_t start = TG.pt_off + jl * LG;
            const uint32_t npts = active ? min(LG, TG.n - jl * LG) : 0u;
            LeafAgg a = leaf_agg_empty();
            uint32_t k0 = 0;
            for (uint32_t k = off; k < npts; k += ls) {
                const uint32_t pp = start + k;
                const double px = T.xs[pp], py = T.ys[pp];
                if (k == 0)
                    k0 = sort_key(TG, px, py, shift);
                leaf_agg_add(a, px, py, T.zs[pp], T.ids[pp]);
            }
            seg_reduce(off, ls, [&](uint32_t o, bool take) {
                uint32_t c2 = __shfl_down(a.cnt, o, 64);
                double w0 = __shfl_down(a.mnx, o, 64), w1 = __shfl_down(a.mny, o, 64);
                double w2 = __shfl_down(a.mxx, o, 64), w3 = __shfl_down(a.mxy, o, 64);
                unsigned long long k2 = __shfl_down(a.bk, o, 64);
                uint32_t i2 = __shfl_down(a.bid, o, 64);
                if (take) {
                    a.cnt += c2;
                    a.mnx = fmin(a.mnx, w0);
                    a.mny = fmin(a.mny, w1);
                    a.mxx = fmax(a.mxx, w2);
                    a.mxy = fmax(a.mxy, w3);
                    if (owm_less(k2, i2, a.bk, a.bid)) {
                        a.bk = k2;
                        a.bid = i2;
                    }
                }
            });
            if (active && off == 0)
                leaf_store(TG, T, k0, jl, start, npts, a);
        }
    }
}

__device__ __forceinline__ int kdelta(const uint32_t *key0, uint32_t nl, int a, int b)
{
    if (b < 0 || b >= (int)nl)
        return -1;
    uint32_t ka = key0[a], kb = key0[b];
    if (ka == kb)
        return 32 + __clz((uint32_t)a ^ (uint32_t)b);
    return __clz(ka ^ kb);
}

// Karras (2012) binary radix tree over the leaves of each tile, keys = Morton
// key of each leaf's first point, ties broken by leaf index.
__global__ __launch_bounds__(TB) void k_karras(BatchDev B, TreeOut T)
{
    for (uint32_t g = blockIdx.x * TB + threadIdx.x; g < B.nnode; g += gridDim.x * TB) {
        uint32_t t = owm_find_tile(B.node_off, B.ntiles, g);
        const TileDev &TD = B.tiles[t];
        const uint32_t nl = TD.nleaf;
        const uint32_t *key0 = T.lf_key0 + TD.leaf_off;
        const int i = (int)(g - TD.node_off);
        int d = (kdelta(key0, nl, i, i + 1) - kdelta(key0, nl, i, i - 1)) > 0 ? 1 : -1;
        int dmin = kdelta(key0, nl, i, i - d);
        int lmax = 2;
        while (kdelta(key0, nl, i, i + lmax * d) > dmin)
            lmax <<= 1;
        int l = 0;
        for (int s = lmax >> 1; s >= 1; s >>= 1)
            if (kdelta(key0, nl, i, i + (l + s) * d) > dmin)
                l += s;
        int j = i + l * d;
        int dnode = kdelta(key0, nl, i, j);
        int s = 0;
        for (int div = 2;; div <<= 1) {
            int step = (l + div - 1) / div;
            if (kdelta(key0, nl, i, i + (s + step) * d) > dnode)
                s += step;
            if (step <= 1)
                break;
        }
        int gamma = i + s * d + (d < 0 ? -1 : 0);
        int lo = min(i, j), hi = max(i, j);
        uint32_t left, right;
        if (lo == gamma) {
            left = OWM_LEAF_BIT | (TD.leaf_off + gamma);
            T.lf_parent[TD.leaf_off + gamma] = g;
        } else {
            left = TD.node_off + gamma;
            T.nd_parent[TD.node_off + gamma] = g;
        }
        if (hi == gamma + 1) {
            right = OWM_LEAF_BIT | (TD.leaf_off + gamma + 1);
            T.lf_parent[TD.leaf_off + gamma + 1] = g;
        } else {
            right = TD.node_off + gamma + 1;
            T.nd_parent[TD.node_off + gamma + 1] = g;
        }
        T.nd_child[g] = make_uint2(left, right);
        T.nd_range[g] = make_uint2((uint32_t)lo, (uint32_t)hi);
        if (i == 0)
            T.nd_parent[g] = OWM_NONE;
    }
}

struct Agg {
    double4 box;
    uint32_t cnt, zid;
    unsigned long long zk;
};

// Leaf aggregates come from an earlier kernel and are read with plain loads.
// Node aggregates are written during the climb by another wave of the same
// workgroup (chunk climb, SCOPE = workgroup) or by another workgroup (global
// climb, SCOPE = agent), so they are read with loads coherent at that scope.
template <int SCOPE>
__device__ __forceinline__ double ld_d(const double *p)
{
    unsigned long long v = __hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED, SCOPE);
    return __longlong_as_double((long long)v);
}

template <int SCOPE>
__device__ __forceinline__ Agg load_agg(const TreeOut &T, uint32_t ref)
{
    Agg a;
    if (ref & OWM_LEAF_BIT) {
        const uint32_t l = ref & ~OWM_LEAF_BIT;
        a.box = T.lf_box[l];
        a.cnt = T.lf_cnt[l];
        a.zk = T.lf_zkey[l];
        a.zid = T.lf_zid[l];
    } else {
        const double *bp = (const double *)(T.nd_box + ref);
        a.cnt = __hip_atomic_load(T.nd_cnt + ref, __ATOMIC_RELAXED, SCOPE);
        a.zk = __hip_atomic_load(T.nd_zkey + ref, __ATOMIC_RELAXED, SCOPE);
        a.zid = __hip_atomic_load(T.nd_zid + ref, __ATOMIC_RELAXED, SCOPE);
        a.box = make_double4(ld_d<SCOPE>(bp), ld_d<SCOPE>(bp + 1), ld_d<SCOPE>(bp + 2),
                             ld_d<SCOPE>(bp + 3));
    }
    return a;
}

__device__ __forceinline__ Agg merge_agg(const Agg &a, const Agg &b)
{
    Agg r;
    r.box = make_double4(fmin(a.box.x, b.box.x), fmin(a.box.y, b.box.y), fmax(a.box.z, b.box.z),
                         fmax(a.box.w, b.box.w));
    r.cnt = a.cnt + b.cnt;
    if (owm_less(b.zk, b.zid, a.zk, a.zid)) {
        r.zk = b.zk;
        r.zid = b.zid;
    } else {
        r.zk = a.zk;
        r.zid = a.zid;
    }
    return r;
}

__device__ __forceinline__ void store_agg(const TreeOut &T, uint32_t p, const Agg &a)
{
    T.nd_box[p] = a.box;
    T.nd_cnt[p] = a.cnt;
    T.nd_zkey[p] = a.zk;
    T.nd_zid[p] = a.zid;
}

constexpr uint32_t BU_CHUNK = 2048;  // leaves per workgroup in the chunk climb
constexpr int BU_TB = 1024;           // threads per chunk-climb workgroup
// A chunk exports at most the children of the crossing paths at its two
// boundaries, each no longer than the depth limit.
constexpr uint32_t BU_MAX_DEPTH = 128;
constexpr uint32_t BU_EXP_PER_CHUNK = 2 * (BU_MAX_DEPTH + 2);

// Chunk climb: a workgroup owns leaves [A, E] of the batch; a node whose
// leaf range lies inside [A, E] has both children completed by this
// workgroup, so its arrival counter lives in LDS and needs no device-wide
// fence.  A completed node whose parent's range leaves the chunk is exported
// for the global climb.
__global__ __launch_bounds__(BU_TB) void k_bottomup_local(BatchDev B, TreeOut T, uint32_t *exp_cnt,
                                                        uint32_t *exported, uint32_t *err)
{
    __shared__ uint32_t sFlag[BU_CHUNK];
    __shared__ uint32_t sNexp;
    const uint32_t nch = (B.nleaf + BU_CHUNK - 1) / BU_CHUNK;
    for (uint32_t ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        const uint32_t A = ch * BU_CHUNK, E = min(A + BU_CHUNK, B.nleaf) - 1;
        uint32_t *slots = exported + (size_t)ch * BU_EXP_PER_CHUNK;
        for (uint32_t k = threadIdx.x; k < BU_CHUNK; k += BU_TB)
            sFlag[k] = 0;
        if (threadIdx.x == 0)
            sNexp = 0;
        __syncthreads();
        for (uint32_t l = A + threadIdx.x; l <= E; l += BU_TB) {
            const TileDev &TD = B.tiles[owm_find_tile(B.leaf_off, B.ntiles, l)];
            uint32_t c = OWM_LEAF_BIT | l;
            Agg mine = load_agg<__HIP_MEMORY_SCOPE_WORKGROUP>(T, c);
            uint32_t p = T.lf_parent[l];
            for (uint32_t depth = 0; p != OWM_NONE; ++depth) {
                if (depth > BU_MAX_DEPTH) {
                    atomicOr(err, 4u);
                    break;
                }
                // everything about p in one round trip
                const uint2 rg = T.nd_range[p];
                const uint2 chd = T.nd_child[p];
                const uint32_t pp = T.nd_parent[p];
                const uint32_t glo = TD.leaf_off + rg.x, ghi = TD.leaf_off + rg.y;
                if (glo < A || ghi > E) {
                    const uint32_t k = atomicAdd(&sNexp, 1u);
                    if (k < BU_EXP_PER_CHUNK)
                        slots[k] = c;
                    else
                        atomicOr(err, 8u);
                    break;
                }
                const uint32_t slot = TD.leaf_off + (p - TD.node_off) - A;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                const uint32_t old = atomicAdd(&sFlag[slot], 1u);
                if (old == 0)
                    break;
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                const Agg other = load_agg<__HIP_MEMORY_SCOPE_WORKGROUP>(T, chd.x == c ? chd.y : chd.x);
                mine = chd.x == c ? merge_agg(mine, other) : merge_agg(other, mine);
                store_agg(T, p, mine);
                c = p;
                p = pp;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0)
            exp_cnt[ch] = min(sNexp, BU_EXP_PER_CHUNK);
        __syncthreads();
    }
}

__device__ __forceinline__ bool in_tile(const TileDev &TD, uint32_t ref)
{
    if (ref & OWM_LEAF_BIT) {
        const uint32_t l = ref & ~OWM_LEAF_BIT;
        return l >= TD.leaf_off && l < TD.leaf_off + TD.nleaf;
    }
    return ref >= TD.node_off && ref + 1 < TD.node_off + TD.nleaf;
}

// Global climb from the exported nodes (those whose parent's leaf range
// crosses a chunk).  Climbs of one tile can meet, climbs of different tiles
// never do, so one workgroup owns each tile: all writes and arrivals of a
// tile's climb come from one workgroup and workgroup-scope release/acquire
// suffice (no device-wide cache write-back per level).  Aggregates of the
// chunk climb were written by an earlier kernel and are visible.
__global__ __launch_bounds__(BU_TB) void k_bottomup_global(BatchDev B, TreeOut T,
                                                            const uint32_t *exp_cnt,
                                                            const uint32_t *exported, uint32_t *err)
{
    __shared__ uint32_t sOff[BU_TB + 1];
    __shared__ uint32_t sWave[BU_TB / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (uint32_t t = blockIdx.x; t < B.ntiles; t += gridDim.x) {
        const TileDev &TD = B.tiles[t];
        if (TD.nleaf < 2)
            continue;
        const uint32_t c0 = TD.leaf_off / BU_CHUNK, c1 = (TD.leaf_off + TD.nleaf - 1) / BU_CHUNK;
        if (c0 == c1)
            continue;  // the chunk climb finished the whole tree
        for (uint32_t cb = c0; cb <= c1; cb += BU_TB) {
            const uint32_t m = min((uint32_t)BU_TB, c1 - cb + 1);
            // exclusive scan of the export counts of chunks cb .. cb+m-1
            const uint32_t v = tid < m ? exp_cnt[cb + tid] : 0u;
            uint32_t inc = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                uint32_t u = __shfl_up(inc, o, 64);
                if (lane >= (uint32_t)o)
                    inc += u;
            }
            if (lane == 63)
                sWave[w] = inc;
            __syncthreads();
            uint32_t add = 0, total = 0;
            for (uint32_t k = 0; k < BU_TB / 64; ++k) {
                if (k < w)
                    add += sWave[k];
                total += sWave[k];
            }
            sOff[tid] = add + inc - v;
            if (tid == 0)
                sOff[BU_TB] = total;
            __syncthreads();
            for (uint32_t i = tid; i < total; i += BU_TB) {
                uint32_t lo = 0, hi = m;  // last chunk with sOff <= i
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (sOff[mid] <= i)
                        lo = mid;
                    else
                        hi = mid;
                }
                uint32_t c = exported[(size_t)(cb + lo) * BU_EXP_PER_CHUNK + (i - sOff[lo])];
                if (!in_tile(TD, c))
                    continue;  // a chunk shared with a neighbouring tile
                Agg mine = load_agg<__HIP_MEMORY_SCOPE_WORKGROUP>(T, c);
                uint32_t p = (c & OWM_LEAF_BIT) ? T.lf_parent[c & ~OWM_LEAF_BIT] : T.nd_parent[c];
                for (uint32_t depth = 0; p != OWM_NONE; ++depth) {
                    if (depth > BU_MAX_DEPTH) {
                        atomicOr(err, 4u);
                        break;
                    }
                    const uint2 chd = T.nd_child[p];
                    const uint32_t pp = T.nd_parent[p];
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    const uint32_t old = __hip_atomic_fetch_add(&T.nd_flag[p], 1u, __ATOMIC_RELAXED,
                                                                __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (old == 0)
                        break;
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                    const Agg other =
                        load_agg<__HIP_MEMORY_SCOPE_WORKGROUP>(T, chd.x == c ? chd.y : chd.x);
                    mine = chd.x == c ? merge_agg(mine, other) : merge_agg(other, mine);
                    store_agg(T, p, mine);
                    c = p;
                    p = pp;
                }
            }
            __syncthreads();  // sOff is rewritten for the next piece
        }
    }
}

static uint32_t grid_for(uint64_t work, uint32_t per_block, uint32_t cap)
{
    uint64_t g = (work + per_block - 1) / per_block;
    if (g < 1)
        g = 1;
    if (g > cap)
        g = cap;
    return (uint32_t)g;
}

uint64_t bbox_key_words(uint32_t ntiles) { return (uint64_t)BBOX_REPL * 4 * ntiles; }

void launch_bbox_init(hipStream_t s, unsigned long long *bk, uint32_t ntiles)
{
    uint64_t nw = bbox_key_words(ntiles);
    hipLaunchKernelGGL(k_bbox_init, dim3(grid_for(nw, TB, 256)), dim3(TB), 0, s, bk, (uint32_t)nw);
}

void launch_bbox(hipStream_t s, const BatchDev &B, const double *x, const double *y,
                 unsigned long long *bk)
{
    // about four chunks per block: fewer flushes, still every CU busy
    hipLaunchKernelGGL(k_bbox, dim3(grid_for(B.np, 4 * CHUNK, 2048)), dim3(TB), 0, s, B, x, y, bk);
    hipLaunchKernelGGL(k_bbox_fold, dim3(grid_for(4ull * B.ntiles, TB, 64)), dim3(TB), 0, s, bk,
                       B.ntiles);
}

void launch_morton(hipStream_t s, const BatchDev &B, const double *x, const double *y,
                   uint32_t *keys, uint32_t *hist, int passes)
{
    hipLaunchKernelGGL(k_morton, dim3(grid_for(B.np, CHUNK, 2048)), dim3(TB), 0, s, B, x, y, keys,
                       hist, passes);
}

void launch_leaf_aggs(hipStream_t s, const BatchDev &B, const TreeOut &T)
{
    hipLaunchKernelGGL(k_leaf_aggs, dim3(grid_for(B.ngrp, LEAF_G * TB / 64, 16384)), dim3(TB), 0, s,
                       B, T);
}

void launch_karras(hipStream_t s, const BatchDev &B, const TreeOut &T)
{
    if (B.nnode == 0)
        return;
    hipLaunchKernelGGL(k_karras, dim3(grid_for(B.nnode, TB, 8192)), dim3(TB), 0, s, B, T);
}

uint64_t bottomup_work_words(uint64_t nleaf, uint32_t ntiles)
{
    (void)ntiles;
    const uint64_t chunks = (nleaf + BU_CHUNK - 1) / BU_CHUNK;
    return chunks * (1 + BU_EXP_PER_CHUNK) + 64;
}

void launch_bottomup(hipStream_t s, const BatchDev &B, const TreeOut &T, uint32_t *work,
                     uint64_t work_words, uint32_t *err)
{
    (void)work_words;
    if (B.nnode == 0)
        return;
    launch_fill_u32(s, T.nd_flag, B.nnode, 0u);
    const uint32_t nch = (B.nleaf + BU_CHUNK - 1) / BU_CHUNK;
    uint32_t *exp_cnt = work, *exported = work + nch;
    hipLaunchKernelGGL(k_bottomup_local, dim3(nch < 1 ? 1 : (nch > 16384 ? 16384 : nch)), dim3(BU_TB),
                       0, s, B, T, exp_cnt, exported, err);
    const uint32_t g = B.ntiles < 1 ? 1 : (B.ntiles > 4096 ? 4096 : B.ntiles);
    hipLaunchKernelGGL(k_bottomup_global, dim3(g), dim3(BU_TB), 0, s, B, T, (const uint32_t *)exp_cnt,
                       (const uint32_t *)exported, err);
}

}  // namespace owm
