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

// ------------------------------------------------- bbox + key histograms --
//
// When the tiles' bboxes are not given, one pass over x and y computes both
// the exact bbox (for the window grid) and the digit histograms of the sort
// keys.  The keys then cannot be quantised over the exact bbox, which is known
// only after the pass: they use a frame taken from a sample of each tile
// (FRAME_RUNS runs of FRAME_RUN consecutive points spread over the tile;
// every point of a tile of up to FRAME_RUNS x FRAME_RUN points).  A point
// outside the sampled box clamps to the frame's edge cell (quantise), which
// costs a little leaf tightness there and nothing else: the seeds do not
// depend on the sort order.
constexpr uint32_t FRAME_RUNS = 128, FRAME_RUN = 16;

__global__ __launch_bounds__(FRAME_RUNS) void k_frame_sample(const uint32_t *__restrict__ pt_off,
                                                             uint32_t ntiles,
                                                             const double *__restrict__ X,
                                                             const double *__restrict__ Y,
                                                             KeyFrame *__restrict__ frames)
{
    __shared__ double red[4][FRAME_RUNS / 64];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t off = pt_off[t], n = pt_off[t + 1] - off;
        const uint32_t i0 = off + (uint32_t)((uint64_t)threadIdx.x * n / FRAME_RUNS);
        const uint32_t i1 = min(i0 + FRAME_RUN, off + n);
        double mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
        for (uint32_t i = i0; i < i1; ++i) {
            const double x = X[i], y = Y[i];
            mnx = fmin(mnx, x);
            mxx = fmax(mxx, x);
            mny = fmin(mny, y);
            mxy = fmax(mxy, y);
        }
        const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
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
        if (threadIdx.x == 0) {
            for (uint32_t k = 1; k < FRAME_RUNS / 64; ++k) {
                mnx = fmin(mnx, red[0][k]);
                mny = fmin(mny, red[1][k]);
                mxx = fmax(mxx, red[2][k]);
                mxy = fmax(mxy, red[3][k]);
            }
            // the square frame of the sampled box (an empty or NaN sample: any
            // frame sorts correctly)
            if (!(mnx <= mxx))
                mnx = mxx = 0.0;
            if (!(mny <= mxy))
                mny = mxy = 0.0;
            double side = fmax(mxx - mnx, mxy - mny);
            if (!(side > 0.0) || !isfinite(side))
                side = 1.0;
            frames[t] = KeyFrame{mnx, mny, 4294967296.0 / (side * (1.0 + 1e-12))};
        }
        __syncthreads();
    }
}

// k_bbox and k_morton in one pass: exact per-tile bbox (replicated words,
// folded by k_bbox_fold) and per-tile digit histograms of the keys under the
// sampled frames.
__global__ __launch_bounds__(TB) void k_bbox_keys(BatchDev B, const double *__restrict__ X,
                                                   const double *__restrict__ Y,
                                                   const KeyFrame *__restrict__ frames,
                                                   unsigned long long *bk_all, uint32_t *hist,
                                                   int passes, uint32_t key_bits)
{
    __shared__ uint32_t h[4][256];
    __shared__ double red[4][TB / 64];
    unsigned long long *bk = bk_all + (size_t)(blockIdx.x % BBOX_REPL) * 4 * B.ntiles;
    for (int i = threadIdx.x; i < 4 * 256; i += TB)
        (&h[0][0])[i] = 0;
    uint32_t cur = OWM_NONE;
    double mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
    auto flush = [&]() {
        bbox_block_flush(bk, cur, mnx, mny, mxx, mxy, red);  // synchronises the block
        uint32_t *g = hist + (size_t)cur * SORT_SEG_HIST;
        for (int i = threadIdx.x; i < passes * 256; i += TB) {
            const uint32_t c = (&h[0][0])[i];
            if (c) {
                atomicAdd(&g[i], c);
                (&h[0][0])[i] = 0;
            }
        }
        __syncthreads();
        mnx = mny = INFINITY;
        mxx = mxy = -INFINITY;
        cur = OWM_NONE;
    };
    __syncthreads();
    const uint32_t nch = (B.np + CHUNK - 1) / CHUNK;
    const uint32_t shift = 64 - key_bits;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const uint32_t c0 = c * CHUNK, c1 = min(c0 + CHUNK, B.np);
        uint32_t t = owm_find_tile(B.pt_off, B.ntiles, c0);
        uint32_t next = B.pt_off[t + 1];
        const bool single = c1 <= next;
        if (cur != OWM_NONE && (!single || t != cur))
            flush();
        if (single) {
            cur = t;
            const KeyFrame F = frames[t];
            for (uint32_t base = c0 + threadIdx.x; base < c1; base += 8 * TB) {
                double xv[8], yv[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint32_t i = base + j * TB;
                    xv[j] = i < c1 ? X[i] : NAN;
                    yv[j] = i < c1 ? Y[i] : NAN;
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    mnx = fmin(mnx, xv[j]);
                    mxx = fmax(mxx, xv[j]);
                    mny = fmin(mny, yv[j]);
                    mxy = fmax(mxy, yv[j]);
                    if (base + j * TB < c1) {
                        const uint32_t k = sort_key_q(F.qminx, F.qminy, F.qscale, xv[j], yv[j], shift);
                        for (int p = 0; p < passes; ++p)
                            atomicAdd(&h[p][(k >> (8 * p)) & 255u], 1u);
                    }
                }
            }
            continue;
        }
        // a chunk across tiles (rare): per-thread bbox flushes, counts straight
        // to each tile's histogram
        double ax = INFINITY, ay = INFINITY, bx = -INFINITY, by = -INFINITY;
        uint32_t tt = t, nx = next;
        for (uint32_t i = c0 + threadIdx.x; i < c1; i += TB) {
            while (i >= nx) {
                bbox_flush(bk, tt, ax, ay, bx, by);
                ax = ay = INFINITY;
                bx = by = -INFINITY;
                nx = B.pt_off[++tt + 1];
            }
            const double x = X[i], y = Y[i];
            ax = fmin(ax, x);
            bx = fmax(bx, x);
            ay = fmin(ay, y);
            by = fmax(by, y);
            const KeyFrame F = frames[tt];
            const uint32_t k = sort_key_q(F.qminx, F.qminy, F.qscale, x, y, shift);
            uint32_t *g = hist + (size_t)tt * SORT_SEG_HIST;
            for (int p = 0; p < passes; ++p)
                atomicAdd(&g[p * 256 + ((k >> (8 * p)) & 255u)], 1u);
        }
        bbox_flush(bk, tt, ax, ay, bx, by);
    }
    if (cur != OWM_NONE)
        flush();
}

// Each tile's digit histograms of every radix pass of its points' sort keys
// (the top key_bits of each point's Morton code within its tile); the first
// sort pass computes the keys again rather than reading them.
__global__ __launch_bounds__(TB) void k_morton(BatchDev B, const double *__restrict__ X,
                                                const double *__restrict__ Y, uint32_t *hist,
                                                int passes)
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
            const PointRec r = owm_load_pt(T.pts, TD.pt_off + min(e, TD.n - 1));
            v[q] = r.id;
            x[q] = r.x;
            y[q] = r.y;
            z[q] = r.z;
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
        double fx = x[0], fy = y[0];  // the leaf's first point
#pragma unroll
        for (int k = 1; k < NK; ++k)
            if (sub == (uint32_t)k) {
                mine = a[k];
                fx = x[k * QL];
                fy = y[k * QL];
            }
        // its sort key (the records' aux word may hold window classes)
        const uint32_t fk = sort_key(TD, fx, fy, shift);
        const uint32_t jl = e0 / LL + sub;
        const bool store = LL <= 64 ? sub < (uint32_t)NK : (sub == 0 && (seg & 1) == 0);
        if (store && jl < TD.nleaf)
            leaf_store(TD, T, fk, jl, TD.pt_off + jl * LL, min((uint32_t)LL, TD.n - jl * LL), mine);
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
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * (TB / 64);
    const uint32_t shift = 64u - B.key_bits;
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
            uint32_t v[LEAF_G], kq[LEAF_G];
            double x[LEAF_G], y[LEAF_G], z[LEAF_G];
            bool have[LEAF_G];
#pragma unroll
            for (int q = 0; q < LEAF_G; ++q) {
                const uint32_t jl = jl0 + q * nl_grp + sub, e = jl * L + off;
                have[q] = jl < TD.nleaf && e < TD.n;
                // loads clamped to the tile's last point (no branch per load)
                const PointRec r = owm_load_pt(T.pts, TD.pt_off + min(e, TD.n - 1));
                v[q] = r.id;
                x[q] = r.x;
                y[q] = r.y;
                z[q] = r.z;
                kq[q] = sort_key(TD, r.x, r.y, shift);
            }
            // lane m ends with the aggregate of leaf jl0 + m when the groups hold
            // at most 64 leaves (leaf_size >= 4); smaller leaves are stored
            // group by group from each leaf's first lane
            const bool gather = LEAF_G * nl_grp <= 64;
            LeafAgg mine = leaf_agg_empty();
            uint32_t mine_key = 0;  // the key of the leaf's first point
#pragma unroll
            for (int q = 0; q < LEAF_G; ++q) {
                LeafAgg a = leaf_agg_empty();
                if (have[q])
                    leaf_agg_add(a, x[q], y[q], z[q], v[q]);
                leaf_agg_reduce(a, L);
                if (!gather) {
                    const uint32_t jl = jl0 + q * nl_grp + sub;
                    const uint32_t k0 = off == 0 && have[q] ? kq[q] : 0u;
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
                const uint32_t fk = shfl_any(kq[q], first);
                if (take) {
                    mine = a;
                    mine_key = fk;
                }
            }
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
                const PointRec r = owm_load_pt(T.pts, start + k);
                if (k == 0)
                    k0 = sort_key(TG, r.x, r.y, shift);
                leaf_agg_add(a, r.x, r.y, r.z, r.id);
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

// ------------------------------------------------------ aggregate levels --
// Level k >= 1 holds, for every aligned block of 2^k leaves of the batch (by
// batch leaf index), the merged aggregate of its leaves; level 0 is the leaves'
// own aggregates.  An internal node's aggregate merges the at most two blocks
// per level that tile its leaf range, so no bottom-up climb is needed: every
// node is computed on its own, where k_karras finds its range.  A block may
// span tiles, but the blocks that tile a node's range lie inside that range.
// The merge (min/max box, count sum, lowest (z, index)) is exact and
// independent of order.

__device__ __forceinline__ void agg_merge(LeafAgg &a, const LeafAgg &b)
{
    a.mnx = fmin(a.mnx, b.mnx);
    a.mny = fmin(a.mny, b.mny);
    a.mxx = fmax(a.mxx, b.mxx);
    a.mxy = fmax(a.mxy, b.mxy);
    a.cnt += b.cnt;
    if (owm_less(b.bk, b.bid, a.bk, a.bid)) {
        a.bk = b.bk;
        a.bid = b.bid;
    }
}

__device__ __forceinline__ LeafAgg level_load(const TreeOut &T, const AggLevels &V, uint32_t k,
                                              uint32_t i)
{
    LeafAgg a;
    double4 b;
    if (k == 0) {
        b = T.lf_box[i];
        a.cnt = T.lf_cnt[i];
        a.bk = T.lf_zkey[i];
        a.bid = T.lf_zid[i];
    } else {
        const uint32_t o = V.off[k] + i;
        b = V.box[o];
        a.cnt = V.cnt[o];
        a.bk = V.zk[o];
        a.bid = V.zid[o];
    }
    a.mnx = b.x;
    a.mny = b.y;
    a.mxx = b.z;
    a.mxy = b.w;
    return a;
}

__device__ __forceinline__ void level_store(const AggLevels &V, uint32_t k, uint32_t i,
                                            const LeafAgg &a)
{
    if (k > V.nlev || i >= V.size[k])
        return;
    const uint32_t o = V.off[k] + i;
    V.box[o] = make_double4(a.mnx, a.mny, a.mxx, a.mxy);
    V.cnt[o] = a.cnt;
    V.zk[o] = a.bk;
    V.zid[o] = a.bid;
}

constexpr int LV_TB = 256;
constexpr int LV_PER = 8;                    // level-k0 entries per thread
constexpr uint32_t LV_SPAN = LV_TB * LV_PER;  // 2^11 entries per workgroup
constexpr uint32_t LV_STEPS = 11;             // levels made per launch

// Levels k0+1 .. k0+11 from level k0 (n_in entries): three inside a thread
// (8 entries), six across the wave (butterflies), two across the four waves.
__global__ __launch_bounds__(LV_TB) void k_agg_levels(TreeOut T, AggLevels V, uint32_t k0,
                                                       uint32_t n_in)
{
    __shared__ LeafAgg sW[LV_TB / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (uint32_t g = blockIdx.x; g * LV_SPAN < n_in; g += gridDim.x) {
        const uint32_t e0 = g * LV_SPAN + tid * LV_PER;
        LeafAgg a[LV_PER];
#pragma unroll
        for (int j = 0; j < LV_PER; ++j)
            a[j] = e0 + j < n_in ? level_load(T, V, k0, e0 + j) : leaf_agg_empty();
#pragma unroll
        for (int lv = 1; lv <= 3; ++lv) {
#pragma unroll
            for (int i = 0; i < (LV_PER >> lv); ++i) {
                agg_merge(a[2 * i], a[2 * i + 1]);
                a[i] = a[2 * i];
                level_store(V, k0 + lv, (e0 >> lv) + i, a[i]);
            }
        }
        LeafAgg m = a[0];  // level k0+3 entry g * 256 + tid
        const uint32_t q = g * LV_TB + tid;
        leaf_agg_step<1>(m);
        if ((lane & 1) == 0)
            level_store(V, k0 + 4, q >> 1, m);
        leaf_agg_step<2>(m);
        if ((lane & 3) == 0)
            level_store(V, k0 + 5, q >> 2, m);
        leaf_agg_step<4>(m);
        if ((lane & 7) == 0)
            level_store(V, k0 + 6, q >> 3, m);
        leaf_agg_step<8>(m);
        if ((lane & 15) == 0)
            level_store(V, k0 + 7, q >> 4, m);
        leaf_agg_step<16>(m);
        if ((lane & 31) == 0)
            level_store(V, k0 + 8, q >> 5, m);
        leaf_agg_step<32>(m);
        if (lane == 0) {
            level_store(V, k0 + 9, q >> 6, m);
            sW[w] = m;
        }
        __syncthreads();
        if (tid < 2) {
            LeafAgg u = sW[2 * tid];
            agg_merge(u, sW[2 * tid + 1]);
            level_store(V, k0 + 10, g * 2 + tid, u);
        } else if (tid == 2) {
            LeafAgg u = sW[0];
            agg_merge(u, sW[1]);
            agg_merge(u, sW[2]);
            agg_merge(u, sW[3]);
            level_store(V, k0 + 11, g, u);
        }
        __syncthreads();  // sW is rewritten by the next block
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
// Each node's aggregate (bbox, count, lowest (z, index)) merges the aggregate
// levels' blocks that tile its leaf range.
__global__ __launch_bounds__(TB) void k_karras(BatchDev B, TreeOut T, AggLevels V)
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
        const uint32_t left = lo == gamma ? OWM_LEAF_BIT | (TD.leaf_off + gamma) : TD.node_off + gamma;
        const uint32_t right = hi == gamma + 1 ? OWM_LEAF_BIT | (TD.leaf_off + gamma + 1)
                                               : TD.node_off + gamma + 1;
        T.nd_child[g] = make_uint2(left, right);
        // blocks of [leaf_off + lo, leaf_off + hi] level by level (half-open)
        LeafAgg acc = leaf_agg_empty();
        uint32_t ea = TD.leaf_off + (uint32_t)lo, ee = TD.leaf_off + (uint32_t)hi + 1;
        for (uint32_t k = 0; ea < ee; ++k) {
            if (ea & 1)
                agg_merge(acc, level_load(T, V, k, ea++));
            if (ee & 1)
                agg_merge(acc, level_load(T, V, k, --ee));
            ea >>= 1;
            ee >>= 1;
        }
        T.nd_box[g] = make_double4(acc.mnx, acc.mny, acc.mxx, acc.mxy);
        T.nd_cnt[g] = acc.cnt;
        T.nd_zkey[g] = acc.bk;
        T.nd_zid[g] = acc.bid;
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

void launch_bbox_keys(hipStream_t s, const BatchDev &B, const double *x, const double *y,
                      KeyFrame *frames, unsigned long long *bk, uint32_t *hist, int passes,
                      uint32_t key_bits)
{
    const uint32_t gf = B.ntiles < 1 ? 1 : (B.ntiles > 4096 ? 4096 : B.ntiles);
    hipLaunchKernelGGL(k_frame_sample, dim3(gf), dim3(FRAME_RUNS), 0, s, B.pt_off, B.ntiles, x, y,
                       frames);
    hipLaunchKernelGGL(k_bbox_keys, dim3(grid_for(B.np, CHUNK, 2048)), dim3(TB), 0, s, B, x, y,
                       (const KeyFrame *)frames, bk, hist, passes, key_bits);
    hipLaunchKernelGGL(k_bbox_fold, dim3(grid_for(4ull * B.ntiles, TB, 64)), dim3(TB), 0, s, bk,
                       B.ntiles);
}

void launch_morton(hipStream_t s, const BatchDev &B, const double *x, const double *y,
                   uint32_t *hist, int passes)
{
    hipLaunchKernelGGL(k_morton, dim3(grid_for(B.np, CHUNK, 2048)), dim3(TB), 0, s, B, x, y, hist,
                       passes);
}

void launch_leaf_aggs(hipStream_t s, const BatchDev &B, const TreeOut &T)
{
    hipLaunchKernelGGL(k_leaf_aggs, dim3(grid_for(B.ngrp, LEAF_G * TB / 64, 16384)), dim3(TB), 0, s,
                       B, T);
}

void launch_karras(hipStream_t s, const BatchDev &B, const TreeOut &T, const AggLevels &V)
{
    if (B.nnode == 0)
        return;
    hipLaunchKernelGGL(k_karras, dim3(grid_for(B.nnode, TB, 8192)), dim3(TB), 0, s, B, T, V);
}

AggLevels agg_levels_layout(uint64_t nleaf, uint64_t *entries)
{
    AggLevels V{};
    uint64_t o = 0;
    uint32_t k = 0;
    V.size[0] = (uint32_t)nleaf;
    // a node's range needs levels up to floor(log2(nleaf))
    while (k < AGG_MAX_LEV && (nleaf >> k) > 1) {
        ++k;
        V.off[k] = (uint32_t)o;
        V.size[k] = (uint32_t)((nleaf + (1ull << k) - 1) >> k);
        o += V.size[k];
    }
    V.nlev = k;
    *entries = o;
    return V;
}

void launch_agg_levels(hipStream_t s, const TreeOut &T, const AggLevels &V)
{
    uint64_t n_in = V.size[0];
    for (uint32_t k0 = 0; k0 < V.nlev; k0 += LV_STEPS) {
        const uint64_t blocks = (n_in + LV_SPAN - 1) / LV_SPAN;
        hipLaunchKernelGGL(k_agg_levels, dim3(blocks > 8192 ? 8192 : (uint32_t)blocks), dim3(LV_TB),
                           0, s, T, V, k0, (uint32_t)n_in);
        n_in = blocks;
    }
}

}  // namespace owm
