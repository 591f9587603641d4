// Spatial index of the OWM pipeline on gfx950: bbox reduction, 2D Morton
// keys, gather into Morton order with per-leaf aggregates, Karras LBVH over the
// leaves and an atomic-flag bottom-up pass that propagates each node's tight
// bbox, point count and lowest (z, index).
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

__global__ void k_bbox_init(unsigned long long *bk, uint32_t ntiles)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ntiles; i += gridDim.x * blockDim.x) {
        bk[4 * i + 0] = ~0ull;  // min x
        bk[4 * i + 1] = ~0ull;  // min y
        bk[4 * i + 2] = 0ull;   // max x
        bk[4 * i + 3] = 0ull;   // max y
    }
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

// Exact fp64 min/max of x and y per tile (NaNs ignored).
__global__ __launch_bounds__(TB) void k_bbox(BatchDev B, const double *__restrict__ X,
                                              const double *__restrict__ Y, unsigned long long *bk)
{
    __shared__ double red[4][TB / 64];
    const uint32_t nch = (B.np + CHUNK - 1) / CHUNK;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        uint32_t c0 = c * CHUNK, c1 = min(c0 + CHUNK, B.np);
        uint32_t t0 = owm_find_tile(B.pt_off, B.ntiles, c0);
        uint32_t t1 = owm_find_tile(B.pt_off, B.ntiles, c1 - 1);
        double mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
        uint32_t t = t0, next = B.pt_off[t0 + 1];
        for (uint32_t i = c0 + tid; i < c1; i += TB) {
            while (i >= next) {
                bbox_flush(bk, t, mnx, mny, mxx, mxy);
                mnx = mny = INFINITY;
                mxx = mxy = -INFINITY;
                ++t;
                next = B.pt_off[t + 1];
            }
            double x = X[i], y = Y[i];
            mnx = fmin(mnx, x);
            mxx = fmax(mxx, x);
            mny = fmin(mny, y);
            mxy = fmax(mxy, y);
        }
        if (t0 != t1) {
            bbox_flush(bk, t, mnx, mny, mxx, mxy);
            continue;
        }
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
            bbox_flush(bk, t0, mnx, mny, mxx, mxy);
        }
        __syncthreads();
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

// Sort keys (tile bits above the top key_bits of the Morton code) plus the
// digit histograms of every radix pass.  The cloud is also packed into 32-byte
// (x, y, z, 0) records, so that the gather after the sort reads one aligned
// sector per point instead of three scattered 8-byte words.
__global__ __launch_bounds__(TB) void k_morton(BatchDev B, const double *__restrict__ X,
                                                const double *__restrict__ Y,
                                                const double *__restrict__ Z,
                                                uint32_t *__restrict__ keys, double4 *__restrict__ aos,
                                                uint32_t *hist, int passes)
{
    __shared__ uint32_t h[4][256];
    for (int i = threadIdx.x; i < 4 * 256; i += TB)
        (&h[0][0])[i] = 0;
    __syncthreads();
    const uint32_t nch = (B.np + CHUNK - 1) / CHUNK;
    const uint32_t shift = 64 - B.key_bits;
    for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
        uint32_t c0 = c * CHUNK, c1 = min(c0 + CHUNK, B.np);
        uint32_t t = owm_find_tile(B.pt_off, B.ntiles, c0);
        uint32_t next = B.pt_off[t + 1];
        for (uint32_t i = c0 + threadIdx.x; i < c1; i += TB) {
            while (i >= next)
                next = B.pt_off[++t + 1];
            const TileDev &T = B.tiles[t];
            const double x = X[i], y = Y[i];
            unsigned long long m = morton64(T, x, y);
            uint32_t k = T.tile_key | (uint32_t)(m >> shift);
            keys[i] = k;
            aos[i] = make_double4(x, y, Z[i], 0.0);
            for (int p = 0; p < passes; ++p)
                atomicAdd(&h[p][(k >> (8 * p)) & 255u], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * 256; i += TB) {
        uint32_t c = (&h[0][0])[i];
        if (c)
            atomicAdd(&hist[i], c);
    }
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

// Gather into Morton order and per-leaf aggregates.  A wave owns a group of
// 64/leaf_size consecutive leaves of one tile (one leaf when leaf_size > 64,
// its points strided over the lanes): every lane gathers one point per step
// and segmented shuffles reduce each leaf's tight bbox, count (non-NaN x and
// y) and lowest (z, index).
__global__ __launch_bounds__(TB) void k_gather_leaves(BatchDev B, const uint32_t *__restrict__ vals,
                                                       const uint32_t *__restrict__ skeys,
                                                       const double4 *__restrict__ aos, TreeOut T)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nw = gridDim.x * (TB / 64);
    for (uint32_t g = blockIdx.x * (TB / 64) + (threadIdx.x >> 6); g < B.ngrp; g += nw) {
        const TileDev &TD = B.tiles[owm_find_tile(B.grp_off, B.ntiles, g)];
        const uint32_t L = TD.leaf_size;
        const uint32_t nl_grp = L <= 64 ? 64 / L : 1;
        const uint32_t ls = L <= 64 ? L : 64;
        const uint32_t sub = lane / ls, off = lane - sub * ls;
        const uint32_t jl = (g - TD.grp_off) * nl_grp + sub;
        const bool active = sub < nl_grp && jl < TD.nleaf;
        const uint32_t start = TD.pt_off + jl * L;
        const uint32_t npts = active ? min(L, TD.n - jl * L) : 0u;
        double mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
        uint32_t cnt = 0, bid = OWM_NONE;
        unsigned long long bk = OWM_ZKEY_NONE;
        for (uint32_t k = off; k < npts; k += ls) {
            uint32_t p = start + k;
            uint32_t v = vals[p];
            const double4 r = aos[v];
            const double x = r.x, y = r.y, z = r.z;
            T.xs[p] = x;
            T.ys[p] = y;
            T.zs[p] = z;
            T.ids[p] = v;
            if (T.pos)
                T.pos[v] = p;
            if (!isnan(x) && !isnan(y)) {
                ++cnt;
                mnx = fmin(mnx, x);
                mxx = fmax(mxx, x);
                mny = fmin(mny, y);
                mxy = fmax(mxy, y);
                if (owm_eligible(z)) {
                    unsigned long long zk = owm_zkey(z);
                    if (owm_less(zk, v, bk, bid)) {
                        bk = zk;
                        bid = v;
                    }
                }
            }
        }
        seg_reduce(off, ls, [&](uint32_t o, bool take) {
            uint32_t c2 = __shfl_down(cnt, o, 64);
            double a = __shfl_down(mnx, o, 64), b = __shfl_down(mny, o, 64);
            double c = __shfl_down(mxx, o, 64), d = __shfl_down(mxy, o, 64);
            unsigned long long k2 = __shfl_down(bk, o, 64);
            uint32_t i2 = __shfl_down(bid, o, 64);
            if (take) {
                cnt += c2;
                mnx = fmin(mnx, a);
                mny = fmin(mny, b);
                mxx = fmax(mxx, c);
                mxy = fmax(mxy, d);
                if (owm_less(k2, i2, bk, bid)) {
                    bk = k2;
                    bid = i2;
                }
            }
        });
        if (active && off == 0) {
            const uint32_t leaf = TD.leaf_off + jl;
            T.lf_box[leaf] = make_double4(mnx, mny, mxx, mxy);
            T.lf_cnt[leaf] = cnt;
            T.lf_start[leaf] = start;
            T.lf_npts[leaf] = npts;
            T.lf_zkey[leaf] = bk;
            T.lf_zid[leaf] = bid;
            T.lf_key0[leaf] = skeys[start];
            T.lf_parent[leaf] = OWM_NONE;
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

// Aggregates are read with device-coherent loads: the writer may be another
// workgroup (global climb) or another wave of this workgroup (chunk climb).
__device__ __forceinline__ double ld_d(const double *p)
{
    unsigned long long v = __hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)v);
}

__device__ __forceinline__ Agg load_agg(const TreeOut &T, uint32_t ref)
{
    Agg a;
    const double *bp;
    if (ref & OWM_LEAF_BIT) {
        uint32_t l = ref & ~OWM_LEAF_BIT;
        bp = (const double *)(T.lf_box + l);
        a.cnt = __hip_atomic_load(T.lf_cnt + l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        a.zk = __hip_atomic_load(T.lf_zkey + l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        a.zid = __hip_atomic_load(T.lf_zid + l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        bp = (const double *)(T.nd_box + ref);
        a.cnt = __hip_atomic_load(T.nd_cnt + ref, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        a.zk = __hip_atomic_load(T.nd_zkey + ref, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        a.zid = __hip_atomic_load(T.nd_zid + ref, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    a.box = make_double4(ld_d(bp), ld_d(bp + 1), ld_d(bp + 2), ld_d(bp + 3));
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

// Chunk climb: a workgroup owns leaves [A, E] of the batch; a node whose
// leaf range lies inside [A, E] has both children completed by this
// workgroup, so its arrival counter lives in LDS and needs no device-wide
// fence.  A completed node whose parent's range leaves the chunk is exported
// for the global climb.
__global__ __launch_bounds__(TB) void k_bottomup_local(BatchDev B, TreeOut T, uint32_t *exported,
                                                        uint32_t *n_exported, uint32_t cap,
                                                        uint32_t *err)
{
    __shared__ uint32_t sFlag[BU_CHUNK];
    const uint32_t nch = (B.nleaf + BU_CHUNK - 1) / BU_CHUNK;
    for (uint32_t ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        const uint32_t A = ch * BU_CHUNK, E = min(A + BU_CHUNK, B.nleaf) - 1;
        for (uint32_t k = threadIdx.x; k < BU_CHUNK; k += TB)
            sFlag[k] = 0;
        __syncthreads();
        for (uint32_t l = A + threadIdx.x; l <= E; l += TB) {
            const TileDev &TD = B.tiles[owm_find_tile(B.leaf_off, B.ntiles, l)];
            uint32_t c = OWM_LEAF_BIT | l;
            Agg mine = load_agg(T, c);
            uint32_t p = T.lf_parent[l];
            for (uint32_t depth = 0; p != OWM_NONE; ++depth) {
                if (depth > 128) {
                    atomicOr(err, 4u);
                    break;
                }
                uint2 rg = T.nd_range[p];
                uint32_t glo = TD.leaf_off + rg.x, ghi = TD.leaf_off + rg.y;
                if (glo < A || ghi > E) {
                    uint32_t k = atomicAdd(n_exported, 1u);
                    if (k < cap)
                        exported[k] = c;
                    else
                        atomicOr(err, 8u);
                    break;
                }
                uint32_t slot = TD.leaf_off + (p - TD.node_off) - A;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                uint32_t old = atomicAdd(&sFlag[slot], 1u);
                if (old == 0)
                    break;
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                uint2 chd = T.nd_child[p];
                Agg other = load_agg(T, chd.x == c ? chd.y : chd.x);
                mine = chd.x == c ? merge_agg(mine, other) : merge_agg(other, mine);
                store_agg(T, p, mine);
                c = p;
                p = T.nd_parent[p];
            }
        }
        __syncthreads();
    }
}

// Global climb from the exported nodes: arrival counters in global memory
// with device-scope release/acquire fences (only chunk-crossing nodes).
__global__ __launch_bounds__(TB) void k_bottomup_global(TreeOut T, const uint32_t *exported,
                                                         const uint32_t *n_exported, uint32_t cap,
                                                         uint32_t *err)
{
    const uint32_t ne = min(*n_exported, cap);
    for (uint32_t k = blockIdx.x * TB + threadIdx.x; k < ne; k += gridDim.x * TB) {
        uint32_t c = exported[k];
        uint32_t p = (c & OWM_LEAF_BIT) ? T.lf_parent[c & ~OWM_LEAF_BIT] : T.nd_parent[c];
        for (uint32_t depth = 0; p != OWM_NONE; ++depth) {
            if (depth > 128) {
                atomicOr(err, 4u);
                break;
            }
            __threadfence();
            uint32_t old = atomicAdd(&T.nd_flag[p], 1u);
            if (old == 0)
                break;
            __threadfence();
            uint2 chd = T.nd_child[p];
            Agg a = load_agg(T, chd.x), b = load_agg(T, chd.y);
            store_agg(T, p, merge_agg(a, b));
            c = p;
            p = T.nd_parent[p];
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

void launch_bbox_init(hipStream_t s, unsigned long long *bk, uint32_t ntiles)
{
    hipLaunchKernelGGL(k_bbox_init, dim3(grid_for(ntiles, TB, 64)), dim3(TB), 0, s, bk, ntiles);
}

void launch_bbox(hipStream_t s, const BatchDev &B, const double *x, const double *y,
                 unsigned long long *bk)
{
    hipLaunchKernelGGL(k_bbox, dim3(grid_for(B.np, CHUNK, 4096)), dim3(TB), 0, s, B, x, y, bk);
}

void launch_morton(hipStream_t s, const BatchDev &B, const double *x, const double *y,
                   const double *z, uint32_t *keys, double4 *aos, uint32_t *hist, int passes)
{
    hipLaunchKernelGGL(k_morton, dim3(grid_for(B.np, CHUNK, 2048)), dim3(TB), 0, s, B, x, y, z, keys,
                       aos, hist, passes);
}

void launch_gather_leaves(hipStream_t s, const BatchDev &B, const uint32_t *vals,
                          const uint32_t *skeys, const double4 *aos, const TreeOut &T)
{
    hipLaunchKernelGGL(k_gather_leaves, dim3(grid_for(B.ngrp, TB / 64, 16384)), dim3(TB), 0, s, B,
                       vals, skeys, aos, T);
}

void launch_karras(hipStream_t s, const BatchDev &B, const TreeOut &T)
{
    if (B.nnode == 0)
        return;
    hipLaunchKernelGGL(k_karras, dim3(grid_for(B.nnode, TB, 8192)), dim3(TB), 0, s, B, T);
}

uint64_t bottomup_work_words(uint64_t nleaf, uint32_t ntiles)
{
    uint64_t chunks = (nleaf + BU_CHUNK - 1) / BU_CHUNK;
    return 1 + 160 * (chunks + ntiles) + 1024;
}

void launch_bottomup(hipStream_t s, const BatchDev &B, const TreeOut &T, uint32_t *work,
                     uint64_t work_words, uint32_t *err)
{
    if (B.nnode == 0)
        return;
    const uint32_t cap = (uint32_t)(work_words - 1);
    launch_fill_u32(s, T.nd_flag, B.nnode, 0u);
    launch_fill_u32(s, work, 1, 0u);
    const uint32_t nch = (B.nleaf + BU_CHUNK - 1) / BU_CHUNK;
    hipLaunchKernelGGL(k_bottomup_local, dim3(nch < 1 ? 1 : (nch > 16384 ? 16384 : nch)), dim3(TB),
                       0, s, B, T, work + 1, work, cap, err);
    hipLaunchKernelGGL(k_bottomup_global, dim3(64), dim3(TB), 0, s, T, (const uint32_t *)(work + 1),
                       (const uint32_t *)work, cap, err);
}

}  // namespace owm
