// Stage 1 of OWM on gfx950: the exact minimum (and point count) of every
// overlapping window.  Reference: stage1 (CPP-TBB-CPU/src/optim3_func.cpp:541-561,
// the exact -DNOMEMOA build) and the GPU query2D (SYCL-CUDA-GPU/basic/traverse.h:686-721).
//
// Windows are processed in blocks of BX x BY windows, one workgroup per block.
// The block's window edges split x into "classes": class(x) =
// #{edges < x} + #{edges <= x} over the multiset of the block's left and right
// x edges (likewise y), so that points of one class cell lie in exactly the
// same windows and window (i, j) is the class rectangle
// [2#{e <= L_i}, 2#{e < R_i}] x [2#{e <= B_j}, 2#{e < T_j}].  Each point strictly
// inside the block is binned into an LDS class cell (count, lowest (z, index));
// a leaf whose tight bbox lies inside one class cell is binned whole from its
// aggregate without reading its points.  Class cells are then reduced to
// windows separably (along x per class row, then along y per window): every
// window re-uses the cells it shares with its neighbours, which is the
// reference's memoisation of the previous window's minimum taken to its limit.
//
// The leaves each block must read come from a spatial join of leaf bboxes
// against block extents (count, scan, fill), so no tree walk is on the
// critical path of stage 1.
#include "owm_kernels.h"

namespace owm {

constexpr int S1_THREADS = 256;
constexpr int S1_CHUNK = S1_THREADS;  // leaves whose metadata is staged at once
constexpr uint32_t S1_LEAFTAG = 0x80000000u;
constexpr uint32_t S1_POINT = 0xFFFFFFFEu;  // item is a point, classified after the loads
constexpr uint32_t S1_GPOINT = 0xFFFFFFFDu; // item is a point with stamped classes
// Entries of a block's class tables (global class -> the block's class, per
// axis); blocks whose class range is longer (overlaps above ~0.99) read
// coordinates instead of stamped classes.
constexpr uint32_t S1_GT = 512;
template <bool B>
struct S1Tag {
    static constexpr bool value = B;
};

// #{k : E[k] < x} and #{k : E[k] <= x} over a strictly increasing edge list
// E[0..n) spaced ~1/inv apart: estimate, then correct with exact compares.
__device__ __forceinline__ uint32_t cnt_lt(const double *E, uint32_t n, double x, double inv)
{
    double f = (x - E[0]) * inv;
    int k = f >= 0.0 ? (f < (double)n ? (int)f + 1 : (int)n) : 0;
    while (k > 0 && !(E[k - 1] < x))
        --k;
    while (k < (int)n && E[k] < x)
        ++k;
    return (uint32_t)k;
}

__device__ __forceinline__ uint32_t cnt_le(const double *E, uint32_t n, double x, double inv)
{
    double f = (x - E[0]) * inv;
    int k = f >= 0.0 ? (f < (double)n ? (int)f + 1 : (int)n) : 0;
    while (k > 0 && E[k - 1] > x)
        --k;
    while (k < (int)n && E[k] <= x)
        ++k;
    return (uint32_t)k;
}

// #{E < x} + #{E <= x}: the "<=" count only differs when x equals the first
// edge not below it.
__device__ __forceinline__ uint32_t cnt_2(const double *E, uint32_t n, double x, double inv)
{
    uint32_t k = cnt_lt(E, n, x, inv);
    return 2 * k + (k < n && E[k] == x ? 1u : 0u);
}

// Edge lists of one window block along one axis, with the rounding tolerance
// of the arithmetic class below.  Edges are E_i = fl(fl(init + fl(i*disp)) +- h);
// every value involved is at most M = max(|init|, |first edge|, |last edge|) + h
// in magnitude, so E_i = E_0 + i*disp + e with |e| <= 2^-50 M, and
// f = fl(fl(x - E_0) * inv) is within 2^-50 M / disp of (x - E_0) / disp for
// x in the block.  Hence when f is farther than tol = 2^-46 M / disp from every
// integer, x lies strictly between E_floor(f) and E_floor(f)+1 (the bound holds
// four times over), and #{E < x} = #{E <= x} = clamp(floor(f) + 1, 0, n).
struct AxisEdges {
    const double *L, *R;
    uint32_t n;
    double l0, r0, inv, tol;
};

__device__ __forceinline__ uint32_t cnt_2f(const double *E, uint32_t n, double e0, double x,
                                           double inv, double tol)
{
    const double f = (x - e0) * inv, fl = floor(f), fr = f - fl;
    if (fr > tol && fr < 1.0 - tol)
        return 2u * (uint32_t)fmin(fmax(fl + 1.0, 0.0), (double)n);
    return cnt_2(E, n, x, inv);  // near an edge: exact compares
}

// class of a coordinate against the block's left and right edges
__device__ __forceinline__ uint32_t cls_f(const AxisEdges &A, double x)
{
    return cnt_2f(A.L, A.n, A.l0, x, A.inv, A.tol) + cnt_2f(A.R, A.n, A.r0, x, A.inv, A.tol);
}

__device__ __forceinline__ AxisEdges axis_edges(const double *L, const double *R, uint32_t n,
                                                double init, double hd, double inv)
{
    const double m = fmax(fabs(init), fmax(fabs(L[0]), fabs(R[n - 1]))) + hd;
    return AxisEdges{L, R, n, L[0], R[0], inv, ldexp(m, -46) * inv};
}

// Windows [a, b] whose open extent (c_i - h, c_i + h) can hold a point of
// the closed range [lo, hi]: R_i > lo and L_i < hi, with the exact edges.
__device__ __forceinline__ bool win_span(double lo, double hi, double init, double disp, double inv,
                                         double hd, uint32_t n, uint32_t &a, uint32_t &b)
{
    double fa = floor((lo - init - hd) * inv), fb = ceil((hi - init + hd) * inv);
    if (!(fb >= 0.0) || !(fa <= (double)n) || !(fa <= fb + 2.0))
        return false;
    int ia = fa < 0.0 ? 0 : (int)fa, ib = fb > (double)n - 1.0 ? (int)n - 1 : (int)fb;
    // first i with R_i > lo
    while (ia > 0 && __dadd_rn(owm_wcenter(init, disp, ia - 1), hd) > lo)
        --ia;
    while (ia < (int)n && !(__dadd_rn(owm_wcenter(init, disp, ia), hd) > lo))
        ++ia;
    // last i with L_i < hi
    while (ib < (int)n - 1 && __dsub_rn(owm_wcenter(init, disp, ib + 1), hd) < hi)
        ++ib;
    while (ib >= 0 && !(__dsub_rn(owm_wcenter(init, disp, ib), hd) < hi))
        --ib;
    if (ia > ib)
        return false;
    a = (uint32_t)ia;
    b = (uint32_t)ib;
    return true;
}

// Leaf -> window-block join.  mode 0 counts, mode 1 fills the lists.  A lane
// owns a leaf; consecutive leaves are Morton neighbours and mostly share their
// blocks, so the lanes of a wave that name the same block combine into one
// atomic per block (match by ballot), issued by the lowest such lane.
template <int MODE>
__global__ __launch_bounds__(256) void k_s1_bin(BatchDev B, TreeDev T, uint32_t bx, uint32_t by,
                                                uint32_t *__restrict__ cnt,
                                                const uint32_t *__restrict__ start,
                                                uint32_t *__restrict__ list, uint32_t cap,
                                                uint32_t *err)
{
    const uint32_t lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t l0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); l0 < B.nleaf; l0 += stride) {
        const uint32_t l = l0 + lane;
        uint32_t i0 = 0, j0 = 0, wi = 0, nb = 0, boff = 0, nbx = 0;
        if (l < B.nleaf && T.lf_cnt[l] != 0) {
            const TileDev &TD = B.tiles[owm_find_tile(B.leaf_off, B.ntiles, l)];
            if (TD.ncols != 0) {
                const double4 box = T.lf_box[l];
                const double hd = (double)(float)TD.h;
                uint32_t a, b, c, d;
                if (win_span(box.x, box.z, TD.initx, TD.displace, TD.inv_disp, hd, TD.ncols, a, b) &&
                    win_span(box.y, box.w, TD.inity, TD.displace, TD.inv_disp, hd, TD.nrows, c, d)) {
                    i0 = a / bx;
                    j0 = c / by;
                    wi = b / bx - i0 + 1;
                    nb = wi * (d / by - j0 + 1);
                    boff = TD.blk_off;
                    nbx = TD.nbx;
                }
            }
        }
        for (uint32_t k = 0; __ballot(k < nb) != 0ull; ++k) {
            const uint32_t key = k < nb ? boff + (j0 + k / wi) * nbx + (i0 + k % wi) : OWM_NONE;
            // match lanes by block (scalar work only), then every leader issues
            // its group's atomic in one instruction: one round trip per k, not
            // one per distinct block
            unsigned long long pending = __ballot(key != OWM_NONE);
            uint32_t leader_of = 0, rank = 0, group = 0;
            while (pending) {
                const int leader = __ffsll((long long)pending) - 1;
                const uint32_t lk = __builtin_amdgcn_readlane(key, leader);
                const unsigned long long peers = __ballot(key == lk);
                if (key == lk) {
                    leader_of = (uint32_t)leader;
                    rank = (uint32_t)__popcll(peers & below);
                }
                if ((int)lane == leader)
                    group = (uint32_t)__popcll(peers);
                pending &= ~peers;
            }
            const bool lead = key != OWM_NONE && lane == leader_of;
            if (MODE == 0) {
                if (lead)
                    atomicAdd(&cnt[key], group);
            } else {
                uint32_t b0 = lead ? atomicAdd(&cnt[key], group) : 0u;
                b0 = (uint32_t)__shfl((int)b0, (int)leader_of, 64);
                if (key != OWM_NONE) {
                    const uint32_t e = start[key] + b0 + rank;
                    if (e < cap)
                        list[e] = l;
                    else
                        atomicOr(err, 16u);
                }
            }
        }
    }
}

// ---------------------------------------------------- leaf cell summaries --
//
// Global edge classes (owm_gclass): points with equal (Gx, Gy) lie in exactly
// the same windows, so a leaf's points with equal classes can be binned as
// one aggregate (count, lowest (z, index)) by every block; k_lcs builds one
// summary per run of equal classes in the leaf's Morton order (up to
// LCS_SLOTS), and stage 1 reads them instead of the points (leaves with more
// runs stay point by point).

// #{a[k] < g} + #{a[k] <= g} over a strictly increasing array of n classes
__device__ __forceinline__ uint32_t icnt2(const uint32_t *a, uint32_t n, uint32_t g)
{
    uint32_t lo = 0, hi = n;  // first k with a[k] >= g
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < g)
            lo = mid + 1;
        else
            hi = mid;
    }
    return 2 * lo + (lo < n && a[lo] == g ? 1u : 0u);
}

__global__ __launch_bounds__(256) void k_lcs(BatchDev B, TreeDev T, LeafCells LC)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nw = gridDim.x * 4;
    for (uint32_t g = blockIdx.x * 4 + (threadIdx.x >> 6); g < B.ngrp; g += nw) {
        const TileDev &TD = B.tiles[owm_find_tile(B.grp_off, B.ntiles, g)];
        if (!TD.lcs)
            continue;
        const uint32_t L = TD.leaf_size;  // <= 64 when summaries are on
        const uint32_t nl_grp = 64 / L, sub = lane / L, off = lane - sub * L;
        const uint32_t jl = (g - TD.grp_off) * nl_grp + sub;
        const bool active = sub < nl_grp && jl < TD.nleaf;
        const uint32_t npts = active ? min(L, TD.n - jl * L) : 0u;
        const bool in = off < npts;
        const double hd = (double)(float)TD.h, inv = TD.inv_disp, disp = TD.displace;
        uint32_t key = OWM_NONE, id = OWM_NONE, cnt = 0;
        unsigned long long zk = OWM_ZKEY_NONE;
        if (in) {
            const PointRec r = owm_load_pt(T.pts, TD.pt_off + jl * L + off);
            const double x = r.x, y = r.y, z = r.z;
            id = r.id;
            // points outside the outermost edges lie in no window
            if (x > owm_edge_at(TD.initx, disp, hd, 0, false) &&
                x < owm_edge_at(TD.initx, disp, hd, TD.ncols - 1, true) &&
                y > owm_edge_at(TD.inity, disp, hd, 0, false) &&
                y < owm_edge_at(TD.inity, disp, hd, TD.nrows - 1, true)) {
                const uint32_t gx = owm_gclass(x, TD.initx, disp, inv, hd, TD.ncols);
                const uint32_t gy = owm_gclass(y, TD.inity, disp, inv, hd, TD.nrows);
                key = (gy << 16) | gx;
                cnt = 1;
                zk = owm_eligible(z) ? owm_zkey(z) : OWM_ZKEY_NONE;
            }
        }
        // 1. runs of equal keys within a leaf (Morton order): segmented scan
        const uint32_t prev = __shfl_up(key, 1, 64);
        const bool head = in && (off == 0 || prev != key);
        const unsigned long long heads = __ballot(head);
        const uint32_t start = 63 - __clzll((heads & ((2ull << lane) - 1ull)) | 1ull);
#pragma unroll
        for (uint32_t o = 1; o < 64; o <<= 1) {
            const uint32_t c2 = __shfl_up(cnt, o, 64);
            const unsigned long long k2 = __shfl_up(zk, o, 64);
            const uint32_t i2 = __shfl_up(id, o, 64);
            if (in && lane >= start + o) {
                cnt += c2;
                if (owm_less(k2, i2, zk, id)) {
                    zk = k2;
                    id = i2;
                }
            }
        }
        const bool tail = in && (off == npts - 1 || (lane < 63 && ((heads >> (lane + 1)) & 1ull)));
        // 2. every run's tail holds the run's summary; a key seen again later
        // in the leaf simply gets another slot (the bins add up the same)
        const bool emit = tail && key != OWM_NONE;
        const unsigned long long emits = __ballot(emit);
        const unsigned long long seg = (L == 64 ? ~0ull : ((1ull << L) - 1ull)) << (sub * L);
        const uint32_t ne = (uint32_t)__popcll(emits & seg);
        if (active) {
            const uint32_t leaf = TD.leaf_off + jl;
            if (emit && ne <= LCS_SLOTS) {
                const size_t e = (size_t)leaf * LCS_SLOTS + __popcll(emits & seg & ((1ull << lane) - 1ull));
                LC.kc[e] = (unsigned long long)key | ((unsigned long long)cnt << 32);
                LC.id[e] = id;
                LC.zk[e] = zk;
            }
            if (off == 0)
                LC.ne[leaf] = ne <= LCS_SLOTS ? ne : LCS_RAW;
        }
    }
}

template <int BX, int BY, int NT>
struct S1Shape {
    static constexpr int NXC = 4 * BX + 1;
    static constexpr int NYC = 4 * BY + 1;
    static constexpr int NCELL = NXC * NYC;
    // the x-pass results (class row x window column) reuse the cell arrays
    static constexpr int M1 = NYC * BX;
    static constexpr int M1_PER_THREAD = (M1 + NT - 1) / NT;
    static_assert(M1 <= NCELL, "x-pass results fit the cell arrays");
};

constexpr uint32_t S1_ENT_LEAVES = 64 / LCS_SLOTS;  // summary leaves per unit of 64 lanes

// USE_LCS compiles the leaf-summary items in (their registers cost occupancy
// when no tile uses them).
// NT threads per workgroup (256 for blocks up to 16x8; larger blocks take
// more waves to hide the latency of one workgroup per LDS-sized share of a
// CU).  The second bound is waves per SIMD: 4 keeps the 8x8 and the 512- and
// 1024-thread variants within 128 VGPRs; 16x8 blocks take 3 (their LDS admits
// three workgroups per CU), leaf summaries 3 and eight items per lane 2.
template <int BX, int BY, int S1_K, bool USE_LCS, int NT = S1_THREADS>
__global__ __launch_bounds__(NT, NT > 256 ? 4 : ((USE_LCS || BX * BY > 64) ? 3 : (S1_K > 4 ? 2 : 4))) void k_stage1(BatchDev B, TreeDev T, LeafCells LC,
                                                        const uint32_t *__restrict__ blk_start,
                                                        const uint32_t *__restrict__ list,
                                                        uint32_t cap,
                                                        uint32_t *__restrict__ min_ids)
{
    using S = S1Shape<BX, BY, NT>;
    constexpr uint32_t NWV = NT / 64;
    constexpr int NXC = S::NXC;
    __shared__ double sLx[BX], sRx[BX], sLy[BY], sRy[BY];
    __shared__ uint32_t sGLx[BX], sGRx[BX], sGLy[BY], sGRy[BY];  // global classes of the edges
    __shared__ uint32_t sCxl[BX], sCxh[BX], sCyl[BY], sCyh[BY];
    // the block's class of global class GBL + 1 + i (x) and GBB + 1 + i (y)
    __shared__ uint8_t sTx[S1_GT], sTy[S1_GT];
    __shared__ uint32_t sCnt[S::NCELL];
    __shared__ unsigned long long sZ[S::NCELL];
    __shared__ uint32_t sId[S::NCELL];
    __shared__ uint32_t sNraw, sNent;
    // leaf chunk: point leaves (first sorted position, and point count or
    // S1_LEAFTAG | class cell when binned whole from the leaf aggregate), and
    // summary leaves (leaf << 5 | number of summaries)
    __shared__ uint2 sLf[S1_CHUNK];
    // the aggregate of a leaf binned whole: z-key (low, high word), index, count
    // (loaded while the chunk is staged, so that the rounds' point loads are
    // the only global loads in flight)
    __shared__ uint4 sAg[S1_CHUNK];
    __shared__ uint32_t sEnt[USE_LCS ? S1_CHUNK : 1];
    unsigned long long *m1z = sZ;
    uint32_t *m1cnt = sCnt;
    uint32_t *m1id = sId;

    // the wave index is uniform: the per-item unit arithmetic stays scalar
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    __shared__ uint32_t sTile;
    for (uint32_t blk = blockIdx.x; blk < B.nblk; blk += gridDim.x) {
        // the block's tile: every thread tests tile boundaries in parallel
        // (one coalesced load each) instead of one dependent binary search;
        // empty tiles never match, exactly one tile does
        for (uint32_t t = tid; t < B.ntiles; t += NT)
            if (B.blk_off[t] <= blk && blk < B.blk_off[t + 1])
                sTile = t;
        __syncthreads();
        const TileDev &TD = B.tiles[__builtin_amdgcn_readfirstlane(sTile)];
        const uint32_t b = blk - TD.blk_off;
        const uint32_t i0 = (b % TD.nbx) * BX, j0 = (b / TD.nbx) * BY;
        const uint32_t ni = min((uint32_t)BX, TD.ncols - i0), nj = min((uint32_t)BY, TD.nrows - j0);
        const double hd = (double)(float)TD.h;
        const double inv = TD.inv_disp;
        const bool lcs = USE_LCS && TD.lcs != 0;
        // the tree's records carry each point's global classes for these windows
        const bool gcls = TD.gcls != 0;

        if (tid < ni) {
            double c = owm_wcenter(TD.initx, TD.displace, i0 + tid);
            sLx[tid] = __dsub_rn(c, hd);
            sRx[tid] = __dadd_rn(c, hd);
            if (lcs || gcls) {
                sGLx[tid] = owm_gclass(sLx[tid], TD.initx, TD.displace, TD.inv_disp, hd, TD.ncols);
                sGRx[tid] = owm_gclass(sRx[tid], TD.initx, TD.displace, TD.inv_disp, hd, TD.ncols);
            }
        }
        if (tid >= 64 && tid < 64 + nj) {
            double c = owm_wcenter(TD.inity, TD.displace, j0 + tid - 64);
            sLy[tid - 64] = __dsub_rn(c, hd);
            sRy[tid - 64] = __dadd_rn(c, hd);
            if (lcs || gcls) {
                sGLy[tid - 64] = owm_gclass(sLy[tid - 64], TD.inity, TD.displace, TD.inv_disp, hd, TD.nrows);
                sGRy[tid - 64] = owm_gclass(sRy[tid - 64], TD.inity, TD.displace, TD.inv_disp, hd, TD.nrows);
            }
        }
        for (int c = tid; c < S::NCELL; c += NT) {
            sCnt[c] = 0;
            sZ[c] = OWM_ZKEY_NONE;
            sId[c] = OWM_NONE;
        }
        __syncthreads();
        // window (i, j) = classes [2#{e <= L_i}, 2#{e < R_i}] (likewise y)
        if (tid < ni) {
            double l = sLx[tid], r = sRx[tid];
            sCxl[tid] = 2 * (cnt_le(sLx, ni, l, inv) + cnt_le(sRx, ni, l, inv));
            sCxh[tid] = 2 * (cnt_lt(sLx, ni, r, inv) + cnt_lt(sRx, ni, r, inv));
        } else if (tid >= 64 && tid < 64 + nj) {
            double l = sLy[tid - 64], r = sRy[tid - 64];
            sCyl[tid - 64] = 2 * (cnt_le(sLy, nj, l, inv) + cnt_le(sRy, nj, l, inv));
            sCyh[tid - 64] = 2 * (cnt_lt(sLy, nj, r, inv) + cnt_lt(sRy, nj, r, inv));
        }
        const double BL = sLx[0], BR = sRx[ni - 1], BB = sLy[0], BT = sRy[nj - 1];
        const AxisEdges AX = axis_edges(sLx, sRx, ni, TD.initx, hd, inv);
        const AxisEdges AY = axis_edges(sLy, sRy, nj, TD.inity, hd, inv);
        const uint32_t GBL = sGLx[0], GBR = sGRx[ni - 1], GBB = sGLy[0], GBT = sGRy[nj - 1];
        // points are read by their stamped classes when the block's class
        // tables (global class strictly inside the block -> block class) fit;
        // the tables are complete after the barrier that opens every leaf chunk
        const bool gc = gcls && GBR - GBL - 1 <= S1_GT && GBT - GBB - 1 <= S1_GT;
        if (gc) {
            for (uint32_t i = tid; i < GBR - GBL - 1; i += NT)
                sTx[i] = (uint8_t)(icnt2(sGLx, ni, GBL + 1 + i) + icnt2(sGRx, ni, GBL + 1 + i));
            for (uint32_t i = tid; i < GBT - GBB - 1; i += NT)
                sTy[i] = (uint8_t)(icnt2(sGLy, nj, GBB + 1 + i) + icnt2(sGRy, nj, GBB + 1 + i));
        }

        const uint32_t s0 = min(blk_start[blk], cap), s1 = min(blk_start[blk + 1], cap);
        const uint32_t nl = s1 - s0;
        const uint32_t L = TD.leaf_size;
        const uint32_t nl_grp = L <= 64 ? 64 / L : 1;
        const uint32_t ls = L <= 64 ? L : 64;
        const uint32_t nseg = L <= 64 ? 1 : (L + 63) / 64;
        // units per leaf group as a shift where it is a power of two (it is
        // for leaf sizes up to 128): the per-item division is scalar work
        const uint32_t nseg_sh = (nseg & (nseg - 1)) == 0 ? (uint32_t)__builtin_ctz(nseg) : 32u;
        const uint32_t sub = lane / ls, off = lane - sub * ls;
        const uint32_t esub = lane / LCS_SLOTS, eslot = lane - esub * LCS_SLOTS;

        // the leaf chunks, compiled once for stamped classes and once for
        // coordinates, so that each keeps its own registers
        auto chunks = [&](auto gc_tag) {
        constexpr bool GC = decltype(gc_tag)::value;
        // a stamped point takes one load and fewer registers: twice the items
        constexpr int KK = GC ? 2 * S1_K : S1_K;
        for (uint32_t ch = 0; ch < nl; ch += S1_CHUNK) {
            const uint32_t m = min((uint32_t)S1_CHUNK, nl - ch);
            // the class ranges (first chunk) and the previous chunk's rounds
            // are complete before the metadata is overwritten
            if (tid == 0) {
                sNraw = 0;
                sNent = 0;
            }
            __syncthreads();
            if (tid < m) {
                const uint32_t leaf = list[s0 + ch + tid];
                const uint32_t ne = lcs ? LC.ne[leaf] : LCS_RAW;
                if (ne != LCS_RAW) {
                    if (ne)
                        sEnt[atomicAdd(&sNent, 1u)] = (leaf << 5) | ne;
                } else {
                    const uint32_t np = T.lf_npts[leaf], st = T.lf_start[leaf];
                    const double4 bx = T.lf_box[leaf];
                    uint32_t a = st, n = np;
                    // whole leaf inside one class cell: bin its aggregate (a class
                    // cell is never wider than the window step)
                    if (nseg == 1 && np > 0 && bx.z - bx.x < TD.displace && bx.w - bx.y < TD.displace &&
                        bx.x > BL && bx.z < BR && bx.y > BB && bx.w < BT) {
                        uint32_t cx = cls_f(AX, bx.x), cy = cls_f(AY, bx.y);
                        if (cx == cls_f(AX, bx.z) && cy == cls_f(AY, bx.w)) {
                            a = leaf;
                            n = S1_LEAFTAG | (cy * NXC + cx);
                        }
                    }
                    if (n) {
                        const uint32_t k = atomicAdd(&sNraw, 1u);
                        sLf[k] = make_uint2(a, n);
                        if (n & S1_LEAFTAG) {
                            const unsigned long long z = T.lf_zkey[leaf];
                            sAg[k] = make_uint4((uint32_t)z, (uint32_t)(z >> 32), T.lf_zid[leaf], T.lf_cnt[leaf]);
                        }
                    }
                }
            }
            __syncthreads();
            const uint32_t mraw = sNraw, ment = sNent;
            const uint32_t neu = (ment + S1_ENT_LEAVES - 1) / S1_ENT_LEAVES;
            const uint32_t nunits = neu + (mraw + nl_grp - 1) / nl_grp * nseg;
            const uint32_t rounds = (nunits + NWV * KK - 1) / (NWV * KK);

            if constexpr (GC) {
            for (uint32_t r = 0; r < rounds; ++r) {
                // points: z (low, high word), index, classes; aggregates: the
                // z-key (low, high word), index, count.  After binning every
                // item holds its z-key and index.
                uint32_t cell[KK];
                uint4 pr[KK];
                uint32_t dropped = 0;  // bit k: item k lowered its cell's minimum
                // issue every load of the round before using any of them
#pragma unroll
                for (int k = 0; k < KK; ++k) {
                    cell[k] = OWM_NONE;
                    const uint32_t u = (r * KK + k) * NWV + wave;
                    if (u >= nunits)
                        continue;
                    if (USE_LCS && u < neu) {
                        const uint32_t li = u * S1_ENT_LEAVES + esub;
                        if (li >= ment)
                            continue;
                        const uint32_t v = sEnt[li];
                        if (eslot >= (v & 31u))
                            continue;
                        const size_t e = (size_t)(v >> 5) * LCS_SLOTS + eslot;
                        const unsigned long long kc = LC.kc[e], z = LC.zk[e];
                        pr[k] = make_uint4((uint32_t)z, (uint32_t)(z >> 32), LC.id[e], (uint32_t)(kc >> 32));
                        cell[k] = S1_LEAFTAG | (uint32_t)kc;  // classes, mapped below
                        continue;
                    }
                    const uint32_t uu = u - neu;
                    const uint32_t g = nseg_sh < 32 ? uu >> nseg_sh : uu / nseg, seg = uu - g * nseg;
                    const uint32_t li = g * nl_grp + sub;
                    if (sub >= nl_grp || li >= mraw)
                        continue;
                    const uint2 lf = sLf[li];
                    const uint32_t a = lf.x, n = lf.y;
                    if (n & S1_LEAFTAG) {
                        if (seg * 64 + off == 0) {
                            pr[k] = sAg[li];
                            cell[k] = n & ~S1_LEAFTAG;
                        }
                    } else if (seg * 64 + off < n) {
                        // z, index and classes in one 16-byte load
                        pr[k] = *reinterpret_cast<const uint4 *>(T.pts.zi + a + seg * 64 + off);
                        cell[k] = S1_GPOINT;
                    }
                }
#pragma unroll
                for (int k = 0; k < KK; ++k) {
                    if (cell[k] == OWM_NONE)
                        continue;
                    uint32_t cn = pr[k].w;
                    unsigned long long zk = (unsigned long long)pr[k].y << 32 | pr[k].x;
                    if (cell[k] == S1_GPOINT) {
                        if (__builtin_expect(pr[k].w == OWM_GCLS_NONE, 0)) {
                            // not stamped (near an edge, or far out): by the coordinates
                            const uint32_t uu = (r * KK + k) * NWV + wave - neu;
                            const uint32_t g = nseg_sh < 32 ? uu >> nseg_sh : uu / nseg, seg = uu - g * nseg;
                            const double2 p = T.pts.xy[sLf[g * nl_grp + sub].x + seg * 64 + off];
                            if (!(p.x > BL && p.x < BR && p.y > BB && p.y < BT)) {
                                cell[k] = OWM_NONE;
                                continue;
                            }
                            const uint32_t cx = cls_f(AX, p.x), cy = cls_f(AY, p.y);
                            cell[k] = cy * NXC + cx;
                        } else {
                            const uint32_t gx = pr[k].w & 0xFFFFu, gy = pr[k].w >> 16;
                            if (!(gx > GBL && gx < GBR && gy > GBB && gy < GBT)) {
                                cell[k] = OWM_NONE;
                                continue;
                            }
                            const uint32_t tx = sTx[gx - GBL - 1], ty = sTy[gy - GBB - 1];
                            cell[k] = ty * NXC + tx;
                        }
                        cn = 1;
                        const double z = __hiloint2double((int)pr[k].y, (int)pr[k].x);
                        zk = owm_eligible(z) ? owm_zkey(z) : OWM_ZKEY_NONE;
                    } else if (USE_LCS && (cell[k] & S1_LEAFTAG)) {
                        const uint32_t gx = cell[k] & 0xFFFFu, gy = (cell[k] >> 16) & 0x7FFFu;
                        if (!(gx > GBL && gx < GBR && gy > GBB && gy < GBT)) {
                            cell[k] = OWM_NONE;
                            continue;
                        }
                        const uint32_t cx = icnt2(sGLx, ni, gx) + icnt2(sGRx, ni, gx);
                        const uint32_t cy = icnt2(sGLy, nj, gy) + icnt2(sGRy, nj, gy);
                        cell[k] = cy * NXC + cx;
                    }
                    const uint32_t c = cell[k];
                    atomicAdd(&sCnt[c], cn);
                    if (zk != OWM_ZKEY_NONE)
                        dropped |= (atomicMin(&sZ[c], zk) > zk ? 1u : 0u) << k;
                    else
                        cell[k] = OWM_NONE;  // counted, never a minimum
                    pr[k].x = (uint32_t)zk;
                    pr[k].y = (uint32_t)(zk >> 32);
                }
                __syncthreads();
                // the index phases of the coordinate path (below)
#pragma unroll
                for (int k = 0; k < KK; ++k)
                    if (cell[k] != OWM_NONE && sZ[cell[k]] == ((unsigned long long)pr[k].y << 32 | pr[k].x) &&
                        ((dropped >> k) & 1u))
                        sId[cell[k]] = OWM_NONE;
                __syncthreads();
#pragma unroll
                for (int k = 0; k < KK; ++k)
                    if (cell[k] != OWM_NONE && sZ[cell[k]] == ((unsigned long long)pr[k].y << 32 | pr[k].x))
                        atomicMin(&sId[cell[k]], pr[k].z);
            }
            } else {
            for (uint32_t r = 0; r < rounds; ++r) {
                uint32_t cell[KK], idv[KK];
                unsigned long long zk[KK];
                uint32_t dropped = 0;  // bit k: item k lowered its cell's minimum
                double xv[KK], yv[KK], zv[KK];
                uint32_t cn[KK];
                // issue every load of the round before using any of them
#pragma unroll
                for (int k = 0; k < KK; ++k) {
                    cell[k] = OWM_NONE;
                    const uint32_t u = (r * KK + k) * NWV + wave;
                    if (u >= nunits)
                        continue;
                    if (USE_LCS && u < neu) {
                        // a leaf summary: one slot per lane
                        const uint32_t li = u * S1_ENT_LEAVES + esub;
                        if (li >= ment)
                            continue;
                        const uint32_t v = sEnt[li];
                        if (eslot >= (v & 31u))
                            continue;
                        const size_t e = (size_t)(v >> 5) * LCS_SLOTS + eslot;
                        const unsigned long long kc = LC.kc[e];
                        zk[k] = LC.zk[e];
                        idv[k] = LC.id[e];
                        cn[k] = (uint32_t)(kc >> 32);
                        cell[k] = S1_LEAFTAG | (uint32_t)kc;  // classes, mapped below
                        continue;
                    }
                    const uint32_t uu = u - neu;
                    const uint32_t g = nseg_sh < 32 ? uu >> nseg_sh : uu / nseg, seg = uu - g * nseg;
                    const uint32_t li = g * nl_grp + sub;
                    if (sub >= nl_grp || li >= mraw)
                        continue;
                    const uint2 lf = sLf[li];
                    const uint32_t a = lf.x, n = lf.y;
                    if (n & S1_LEAFTAG) {
                        if (seg * 64 + off == 0) {
                            const uint4 ag = sAg[li];
                            cell[k] = n & ~S1_LEAFTAG;
                            cn[k] = ag.w;
                            zk[k] = (unsigned long long)ag.y << 32 | ag.x;
                            idv[k] = ag.z;
                        }
                    } else if (seg * 64 + off < n) {
                        const PointRec r = owm_load_pt(T.pts, a + seg * 64 + off);
                        xv[k] = r.x;
                        yv[k] = r.y;
                        zv[k] = r.z;
                        idv[k] = r.id;
                        cell[k] = S1_POINT;
                    }
                }
#pragma unroll
                for (int k = 0; k < KK; ++k) {
                    if (cell[k] == OWM_NONE)
                        continue;
                    if (cell[k] == S1_POINT) {
                        const double x = xv[k], y = yv[k];
                        if (!(x > BL && x < BR && y > BB && y < BT)) {
                            cell[k] = OWM_NONE;
                            continue;
                        }
                        const uint32_t cx = cls_f(AX, x), cy = cls_f(AY, y);
                        cell[k] = cy * NXC + cx;
                        cn[k] = 1;
                        zk[k] = owm_eligible(zv[k]) ? owm_zkey(zv[k]) : OWM_ZKEY_NONE;
                    } else if (USE_LCS && (cell[k] & S1_LEAFTAG)) {
                        const uint32_t gx = cell[k] & 0xFFFFu, gy = (cell[k] >> 16) & 0x7FFFu;
                        if (!(gx > GBL && gx < GBR && gy > GBB && gy < GBT)) {
                            cell[k] = OWM_NONE;
                            continue;
                        }
                        const uint32_t cx = icnt2(sGLx, ni, gx) + icnt2(sGRx, ni, gx);
                        const uint32_t cy = icnt2(sGLy, nj, gy) + icnt2(sGRy, nj, gy);
                        cell[k] = cy * NXC + cx;
                    }
                    const uint32_t c = cell[k];
                    atomicAdd(&sCnt[c], cn[k]);
                    if (zk[k] != OWM_ZKEY_NONE)
                        dropped |= (atomicMin(&sZ[c], zk[k]) > zk[k] ? 1u : 0u) << k;
                    else
                        cell[k] = OWM_NONE;  // counted, never a minimum
                }
                __syncthreads();
                // A cell whose minimum dropped in this round drops its old
                // index: a lane holding the final minimum that found a larger
                // one before its own update saw the drop (lanes that only tie
                // an earlier minimum found it equal and keep the index).
#pragma unroll
                for (int k = 0; k < KK; ++k)
                    if (cell[k] != OWM_NONE && sZ[cell[k]] == zk[k] && ((dropped >> k) & 1u))
                        sId[cell[k]] = OWM_NONE;
                __syncthreads();
                // No barrier after this phase: a lane that reads a minimum
                // already lowered by the next round skips an index that the
                // next round's reset discards anyway; the next round's resets
                // follow a barrier that every update of this round precedes.
#pragma unroll
                for (int k = 0; k < KK; ++k)
                    if (cell[k] != OWM_NONE && sZ[cell[k]] == zk[k])
                        atomicMin(&sId[cell[k]], idv[k]);
            }
            }
        }
        };
        if (gc)
            chunks(S1Tag<true>{});
        else
            chunks(S1Tag<false>{});
        // the class ranges of a block without leaves, and the last round's
        // index updates, are complete before the window passes
        __syncthreads();

        // separable reduction: x-pass per (class row, window column), held in
        // registers until every cell has been read, then stored over the cells ...
        // Odd classes hold only points lying exactly on an edge.  Window
        // class ranges start and end on even classes, so with every odd class
        // column (row) empty the reductions step over odd columns (rows).
        int ox = 0, oy = 0;
        for (int c = tid; c < S::NCELL; c += NT) {
            const int cy = c / NXC, cx = c - cy * NXC;
            if ((cx | cy) & 1) {
                const bool used = sCnt[c] != 0;
                ox |= (cx & 1) && used;
                oy |= (cy & 1) && used;
            }
        }
        const uint32_t sx = __syncthreads_or(ox) ? 1u : 2u, sy = __syncthreads_or(oy) ? 1u : 2u;
        const uint32_t nyc = 4 * nj + 1, nyr = sy == 1u ? nyc : 2 * nj + 1;
        uint32_t xc[S::M1_PER_THREAD], xi[S::M1_PER_THREAD];
        unsigned long long xz[S::M1_PER_THREAD];
#pragma unroll
        for (int t = 0; t < S::M1_PER_THREAD; ++t) {
            const uint32_t q = tid + t * NT;
            xc[t] = 0;
            xi[t] = OWM_NONE;
            xz[t] = OWM_ZKEY_NONE;
            if (q >= nyr * ni)
                continue;
            const uint32_t rr = q / ni * sy, w = q % ni;
            const uint32_t chi = min(sCxh[w], 4 * ni);
            for (uint32_t c = sCxl[w]; c <= chi; c += sx) {
                const uint32_t cc = rr * NXC + c;
                xc[t] += sCnt[cc];
                if (owm_less(sZ[cc], sId[cc], xz[t], xi[t])) {
                    xz[t] = sZ[cc];
                    xi[t] = sId[cc];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < S::M1_PER_THREAD; ++t) {
            const uint32_t q = tid + t * NT;
            if (q >= nyr * ni)
                continue;
            const uint32_t rr = q / ni * sy, w = q % ni;
            m1cnt[rr * BX + w] = xc[t];
            m1id[rr * BX + w] = xi[t];
            m1z[rr * BX + w] = xz[t];
        }
        __syncthreads();
        // ... then y-pass per window
        for (uint32_t q = tid; q < ni * nj; q += NT) {
            uint32_t v = q / ni, w = q - v * ni;
            uint32_t cnt = 0, id = OWM_NONE;
            unsigned long long z = OWM_ZKEY_NONE;
            const uint32_t rhi = min(sCyh[v], 4 * nj);
            for (uint32_t rr = sCyl[v]; rr <= rhi; rr += sy) {
                uint32_t k = rr * BX + w;
                cnt += m1cnt[k];
                if (owm_less(m1z[k], m1id[k], z, id)) {
                    z = m1z[k];
                    id = m1id[k];
                }
            }
            bool valid = cnt >= TD.min_num_points && z != OWM_ZKEY_NONE;
            min_ids[TD.win_off + (j0 + v) * TD.ncols + (i0 + w)] = valid ? id : OWM_NONE;
        }
        __syncthreads();
    }
}

// ------------------------------------------------ window columns (pairs) --
//
// When a tile's window side is m window steps, every right edge R_i lies
// within rounding of the left edge L_{i+m}, and the left edges cut each axis
// into columns [L_c, L_{c+1}) whose points lie in windows c - m + 1 .. c.
// The tree's records then carry every point's column per axis
// (owm_paxis_fast), and a block of BX x BY windows needs only
// (BX + m - 1) x (BY + m - 1) column cells where the class cells of k_stage1
// take (4 BX + 1) x (4 BY + 1): blocks four times larger fit the LDS, so
// fewer points are read again for block margins and for leaves that straddle
// blocks, and the per-block work is spread over four times the windows.  A
// window's minimum is that of its m x m column cells.  The rest -- points
// within rounding of an edge or between a right edge and the left edge it
// pairs with, which carry no column -- take their windows from exact compares
// with the block's edges and update per-window minima directly, P1_FB at a
// time, with the same three phases (count and minimum z-key, index reset
// where the minimum dropped, index minimum) as the cells.
constexpr uint32_t P1_MMAX = S1_PAIRS_MAX;  // (m + 1)^2 windows of an exception fit 64 bits
constexpr int P1_FB = 256;        // exceptions taken per pass
constexpr int P1_KK = 8;          // stamped points per lane per round
constexpr uint32_t P1_EXC = 0xFFFFFFFCu;  // item is an exception (its windows in pr.w)

template <int BX, int BY, int NT>
__global__ __launch_bounds__(NT, 4) void k_stage1_pairs(BatchDev B, TreeDev T,
                                                        const uint32_t *__restrict__ blk_start,
                                                        const uint32_t *__restrict__ list,
                                                        uint32_t cap, uint32_t *__restrict__ min_ids)
{
    constexpr int NXC = BX + (int)P1_MMAX - 1, NYC = BY + (int)P1_MMAX - 1;
    constexpr int NCELL = NXC * NYC;
    constexpr int NWIN = BX * BY;
    constexpr uint32_t NWV = NT / 64;
    constexpr int M1 = NYC * BX;
    constexpr int M1_PER_THREAD = (M1 + NT - 1) / NT;
    static_assert(M1 <= NCELL, "x-pass results fit the cell arrays");
    static_assert(NT >= 64 + BY && BX <= 64, "one thread per block edge");
    static_assert((P1_MMAX + 1) * (P1_MMAX + 1) <= 64, "an exception's windows fit a 64-bit mask");
    __shared__ double sLx[BX], sRx[BX], sLy[BY], sRy[BY];
    __shared__ uint32_t sCnt[NCELL];
    __shared__ unsigned long long sZ[NCELL];
    __shared__ uint32_t sId[NCELL];
    __shared__ uint32_t wCnt[NWIN];  // the exceptions' per-window counts and minima
    __shared__ unsigned long long wZ[NWIN];
    __shared__ uint32_t wId[NWIN];
    __shared__ uint4 sFb[P1_FB];                // z-key (low, high word), index, windows
    __shared__ unsigned long long sFd[P1_FB];   // windows whose minimum the entry lowered
    __shared__ uint32_t sNfb, sNraw;
    __shared__ uint2 sLf[S1_CHUNK];             // leaf: first sorted position, points
    __shared__ uint32_t sTile;
    unsigned long long *m1z = sZ;
    uint32_t *m1cnt = sCnt;
    uint32_t *m1id = sId;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    for (uint32_t blk = blockIdx.x; blk < B.nblk; blk += gridDim.x) {
        for (uint32_t t = tid; t < B.ntiles; t += NT)
            if (B.blk_off[t] <= blk && blk < B.blk_off[t + 1])
                sTile = t;
        __syncthreads();
        const TileDev &TD = B.tiles[__builtin_amdgcn_readfirstlane(sTile)];
        const uint32_t b = blk - TD.blk_off;
        const uint32_t i0 = (b % TD.nbx) * BX, j0 = (b / TD.nbx) * BY;
        const uint32_t ni = min((uint32_t)BX, TD.ncols - i0), nj = min((uint32_t)BY, TD.nrows - j0);
        const uint32_t m = TD.pairs;
        const int ncx = (int)(ni + m - 1), ncy = (int)(nj + m - 1);
        const double hd = (double)(float)TD.h;
        const double inv = TD.inv_disp;
        if (tid < ni) {
            const double c = owm_wcenter(TD.initx, TD.displace, i0 + tid);
            sLx[tid] = __dsub_rn(c, hd);
            sRx[tid] = __dadd_rn(c, hd);
        }
        if (tid >= 64 && tid < 64 + nj) {
            const double c = owm_wcenter(TD.inity, TD.displace, j0 + tid - 64);
            sLy[tid - 64] = __dsub_rn(c, hd);
            sRy[tid - 64] = __dadd_rn(c, hd);
        }
        for (int c = tid; c < NCELL; c += NT) {
            sCnt[c] = 0;
            sZ[c] = OWM_ZKEY_NONE;
            sId[c] = OWM_NONE;
        }
        for (int w = tid; w < NWIN; w += NT) {
            wCnt[w] = 0;
            wZ[w] = OWM_ZKEY_NONE;
            wId[w] = OWM_NONE;
        }
        if (tid == 0)
            sNfb = 0;
        __syncthreads();
        const double BL = sLx[0], BR = sRx[ni - 1], BB = sLy[0], BT = sRy[nj - 1];

        const uint32_t s0 = min(blk_start[blk], cap), s1 = min(blk_start[blk + 1], cap);
        const uint32_t nl = s1 - s0;
        const uint32_t L = TD.leaf_size;
        const uint32_t nl_grp = L <= 64 ? 64 / L : 1;
        const uint32_t ls = L <= 64 ? L : 64;
        const uint32_t nseg = L <= 64 ? 1 : (L + 63) / 64;
        const uint32_t nseg_sh = (nseg & (nseg - 1)) == 0 ? (uint32_t)__builtin_ctz(nseg) : 32u;
        const uint32_t sub = lane / ls, off = lane - sub * ls;

        for (uint32_t ch = 0; ch < nl; ch += S1_CHUNK) {
            const uint32_t mc = min((uint32_t)S1_CHUNK, nl - ch);
            // the previous chunk's rounds are complete before its leaves go
            if (tid == 0)
                sNraw = 0;
            __syncthreads();
            if (tid < mc) {
                const uint32_t leaf = list[s0 + ch + tid];
                const uint32_t np = T.lf_npts[leaf];
                if (np)
                    sLf[atomicAdd(&sNraw, 1u)] = make_uint2(T.lf_start[leaf], np);
            }
            __syncthreads();
            const uint32_t mraw = sNraw;
            const uint32_t nunits = (mraw + nl_grp - 1) / nl_grp * nseg;
            const uint32_t rounds = (nunits + NWV * P1_KK - 1) / (NWV * P1_KK);
            for (uint32_t r = 0; r < rounds; ++r) {
                // z (low, high word), index, columns; after binning the z-key
                // (low, high word), index and, for an exception, its windows
                uint32_t cell[P1_KK];
                uint4 pr[P1_KK];
                uint32_t dropped = 0;  // bit k: item k lowered its cell's minimum
                uint32_t pend = 0;     // bit k: item k is an exception not yet taken
#pragma unroll
                for (int k = 0; k < P1_KK; ++k) {
                    cell[k] = OWM_NONE;
                    const uint32_t u = (r * P1_KK + k) * NWV + wave;
                    if (u >= nunits)
                        continue;
                    const uint32_t g = nseg_sh < 32 ? u >> nseg_sh : u / nseg, seg = u - g * nseg;
                    const uint32_t li = g * nl_grp + sub;
                    if (sub >= nl_grp || li >= mraw)
                        continue;
                    const uint2 lf = sLf[li];
                    if (seg * 64 + off < lf.y) {
                        pr[k] = *reinterpret_cast<const uint4 *>(T.pts.zi + lf.x + seg * 64 + off);
                        cell[k] = S1_GPOINT;
                    }
                }
#pragma unroll
                for (int k = 0; k < P1_KK; ++k) {
                    if (cell[k] == OWM_NONE)
                        continue;
                    const double z = __hiloint2double((int)pr[k].y, (int)pr[k].x);
                    const unsigned long long zk = owm_eligible(z) ? owm_zkey(z) : OWM_ZKEY_NONE;
                    if (__builtin_expect(pr[k].w == OWM_GCLS_NONE, 0)) {
                        // no column: the windows by exact compares with the edges
                        const uint32_t u = (r * P1_KK + k) * NWV + wave;
                        const uint32_t g = nseg_sh < 32 ? u >> nseg_sh : u / nseg, seg = u - g * nseg;
                        const double2 p = T.pts.xy[sLf[g * nl_grp + sub].x + seg * 64 + off];
                        if (!(p.x > BL && p.x < BR && p.y > BB && p.y < BT)) {
                            cell[k] = OWM_NONE;
                            continue;
                        }
                        const uint32_t wx0 = cnt_le(sRx, ni, p.x, inv), wx1 = cnt_lt(sLx, ni, p.x, inv);
                        const uint32_t wy0 = cnt_le(sRy, nj, p.y, inv), wy1 = cnt_lt(sLy, nj, p.y, inv);
                        if (wx0 >= wx1 || wy0 >= wy1) {
                            cell[k] = OWM_NONE;
                            continue;
                        }
                        pr[k] = make_uint4((uint32_t)zk, (uint32_t)(zk >> 32), pr[k].z,
                                           wx0 | wx1 << 8 | wy0 << 16 | wy1 << 24);
                        cell[k] = P1_EXC;
                        pend |= 1u << k;
                        continue;
                    }
                    const int cx = (int)(pr[k].w & 0xFFFFu) - 1 - (int)i0;
                    const int cy = (int)(pr[k].w >> 16) - 1 - (int)j0;
                    if (cx < 0 || cx >= ncx || cy < 0 || cy >= ncy) {
                        cell[k] = OWM_NONE;
                        continue;
                    }
                    const uint32_t c = (uint32_t)(cy * NXC + cx);
                    atomicAdd(&sCnt[c], 1u);
                    if (zk != OWM_ZKEY_NONE) {
                        dropped |= (atomicMin(&sZ[c], zk) > zk ? 1u : 0u) << k;
                        cell[k] = c;
                    } else {
                        cell[k] = OWM_NONE;  // counted, never a minimum
                    }
                    pr[k].x = (uint32_t)zk;
                    pr[k].y = (uint32_t)(zk >> 32);
                }
                __syncthreads();
                // the index phases of k_stage1
#pragma unroll
                for (int k = 0; k < P1_KK; ++k)
                    if (cell[k] < (uint32_t)NCELL &&
                        sZ[cell[k]] == ((unsigned long long)pr[k].y << 32 | pr[k].x) &&
                        ((dropped >> k) & 1u))
                        sId[cell[k]] = OWM_NONE;
                __syncthreads();
#pragma unroll
                for (int k = 0; k < P1_KK; ++k)
                    if (cell[k] < (uint32_t)NCELL &&
                        sZ[cell[k]] == ((unsigned long long)pr[k].y << 32 | pr[k].x))
                        atomicMin(&sId[cell[k]], pr[k].z);
                // exceptions, P1_FB at a time (the block-wide test is the
                // barrier that orders each pass after the previous one)
                while (__syncthreads_or(pend != 0)) {
#pragma unroll
                    for (int k = 0; k < P1_KK; ++k)
                        if ((pend >> k) & 1u) {
                            const uint32_t e = atomicAdd(&sNfb, 1u);
                            if (e < (uint32_t)P1_FB) {
                                sFb[e] = pr[k];
                                pend &= ~(1u << k);
                            }
                        }
                    __syncthreads();
                    const uint32_t nf = min(sNfb, (uint32_t)P1_FB);
                    for (uint32_t e = tid; e < nf; e += NT) {
                        const uint4 f = sFb[e];
                        const unsigned long long zk = (unsigned long long)f.y << 32 | f.x;
                        const uint32_t x0 = f.w & 255u, x1 = (f.w >> 8) & 255u;
                        const uint32_t y0 = (f.w >> 16) & 255u, y1 = f.w >> 24;
                        unsigned long long low = 0;
                        uint32_t bit = 0;
                        for (uint32_t v = y0; v < y1; ++v)
                            for (uint32_t w = x0; w < x1; ++w, ++bit) {
                                const uint32_t q = v * BX + w;
                                atomicAdd(&wCnt[q], 1u);
                                if (zk != OWM_ZKEY_NONE && atomicMin(&wZ[q], zk) > zk)
                                    low |= 1ull << bit;
                            }
                        sFd[e] = low;
                    }
                    __syncthreads();
                    for (uint32_t e = tid; e < nf; e += NT) {
                        const uint4 f = sFb[e];
                        const unsigned long long zk = (unsigned long long)f.y << 32 | f.x;
                        const unsigned long long low = sFd[e];
                        const uint32_t x0 = f.w & 255u, x1 = (f.w >> 8) & 255u;
                        const uint32_t y0 = (f.w >> 16) & 255u, y1 = f.w >> 24;
                        uint32_t bit = 0;
                        for (uint32_t v = y0; v < y1; ++v)
                            for (uint32_t w = x0; w < x1; ++w, ++bit) {
                                const uint32_t q = v * BX + w;
                                if (((low >> bit) & 1ull) && wZ[q] == zk)
                                    wId[q] = OWM_NONE;
                            }
                    }
                    __syncthreads();
                    for (uint32_t e = tid; e < nf; e += NT) {
                        const uint4 f = sFb[e];
                        const unsigned long long zk = (unsigned long long)f.y << 32 | f.x;
                        if (zk == OWM_ZKEY_NONE)
                            continue;
                        const uint32_t x0 = f.w & 255u, x1 = (f.w >> 8) & 255u;
                        const uint32_t y0 = (f.w >> 16) & 255u, y1 = f.w >> 24;
                        for (uint32_t v = y0; v < y1; ++v)
                            for (uint32_t w = x0; w < x1; ++w) {
                                const uint32_t q = v * BX + w;
                                if (wZ[q] == zk)
                                    atomicMin(&wId[q], f.z);
                            }
                    }
                    __syncthreads();
                    if (tid == 0)
                        sNfb = 0;
                }
            }
        }
        // every round's index updates are complete before the window passes
        __syncthreads();

        // x-pass per (column row, window column), held in registers until
        // every cell has been read, then stored over the cells ...
        uint32_t xc[M1_PER_THREAD], xi[M1_PER_THREAD];
        unsigned long long xz[M1_PER_THREAD];
#pragma unroll
        for (int t = 0; t < M1_PER_THREAD; ++t) {
            const uint32_t q = tid + t * NT;
            xc[t] = 0;
            xi[t] = OWM_NONE;
            xz[t] = OWM_ZKEY_NONE;
            if (q >= (uint32_t)ncy * ni)
                continue;
            const uint32_t rr = q / ni, w = q - rr * ni;
            for (uint32_t c = w; c < w + m; ++c) {
                const uint32_t cc = rr * NXC + c;
                xc[t] += sCnt[cc];
                if (owm_less(sZ[cc], sId[cc], xz[t], xi[t])) {
                    xz[t] = sZ[cc];
                    xi[t] = sId[cc];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < M1_PER_THREAD; ++t) {
            const uint32_t q = tid + t * NT;
            if (q >= (uint32_t)ncy * ni)
                continue;
            const uint32_t rr = q / ni, w = q - rr * ni;
            m1cnt[rr * BX + w] = xc[t];
            m1id[rr * BX + w] = xi[t];
            m1z[rr * BX + w] = xz[t];
        }
        __syncthreads();
        // ... then the y-pass per window, with the window's exceptions
        for (uint32_t q = tid; q < ni * nj; q += NT) {
            const uint32_t v = q / ni, w = q - v * ni;
            const uint32_t qw = v * BX + w;
            uint32_t cnt = wCnt[qw], id = wId[qw];
            unsigned long long z = wZ[qw];
            for (uint32_t rr = v; rr < v + m; ++rr) {
                const uint32_t k = rr * BX + w;
                cnt += m1cnt[k];
                if (owm_less(m1z[k], m1id[k], z, id)) {
                    z = m1z[k];
                    id = m1id[k];
                }
            }
            const bool valid = cnt >= TD.min_num_points && z != OWM_ZKEY_NONE;
            min_ids[TD.win_off + (j0 + v) * TD.ncols + (i0 + w)] = valid ? id : OWM_NONE;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------ exclusive scan ----

constexpr int SC_BLOCK = 256;
constexpr int SC_ITEMS = 16;
constexpr int SC_TILE = SC_BLOCK * SC_ITEMS;
constexpr uint32_t LB_AGG = 1u << 30, LB_INC = 2u << 30, LB_CNT = (1u << 30) - 1;

// out[i] = sum(in[0..i)), out[n] = total; single pass, decoupled look-back.
__global__ __launch_bounds__(SC_BLOCK) void k_scan_excl(const uint32_t *__restrict__ in, uint32_t n,
                                                         uint32_t *__restrict__ out,
                                                         uint32_t *lookback, uint32_t *tile_counter)
{
    __shared__ uint32_t s_tile, s_tmp[SC_BLOCK / 64], s_prefix;
    const uint32_t ntiles = (n + SC_TILE - 1) / SC_TILE;
    if (n == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            out[0] = 0;
        return;
    }
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (;;) {
        if (threadIdx.x == 0)
            s_tile = atomicAdd(tile_counter, 1u);
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= ntiles)
            break;
        const uint32_t i0 = tile * SC_TILE + threadIdx.x * SC_ITEMS;
        uint32_t v[SC_ITEMS], sum = 0;
#pragma unroll
        for (int k = 0; k < SC_ITEMS; ++k) {
            v[k] = i0 + k < n ? in[i0 + k] : 0u;
            sum += v[k];
        }
        uint32_t inc = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t u = __shfl_up(inc, o, 64);
            if (lane >= (uint32_t)o)
                inc += u;
        }
        if (lane == 63)
            s_tmp[w] = inc;
        __syncthreads();
        uint32_t add = 0, total = 0;
        for (uint32_t k = 0; k < SC_BLOCK / 64; ++k) {
            if (k < w)
                add += s_tmp[k];
            total += s_tmp[k];
        }
        if (threadIdx.x == 0) {
            uint32_t *flag = lookback + tile;
            uint32_t prefix = 0;
            if (tile == 0) {
                __hip_atomic_store(flag, LB_INC | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                __hip_atomic_store(flag, LB_AGG | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                int32_t j = (int32_t)tile - 1;
                for (;;) {
                    uint32_t f = __hip_atomic_load(lookback + j, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                    uint32_t st = f & ~LB_CNT;
                    if (st == 0)
                        continue;
                    prefix += f & LB_CNT;
                    if (st == LB_INC)
                        break;
                    --j;
                }
                __hip_atomic_store(flag, LB_INC | (prefix + total), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
            s_prefix = prefix;
            if (tile == ntiles - 1)
                out[n] = prefix + total;
        }
        __syncthreads();
        uint32_t run = s_prefix + add + inc - sum;
#pragma unroll
        for (int k = 0; k < SC_ITEMS; ++k) {
            if (i0 + k < n)
                out[i0 + k] = run;
            run += v[k];
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------- launch --


uint64_t stage1_work_words(uint64_t nblk) { return 2 * (nblk + 1) + (nblk / SC_TILE + 2) + 8; }

template <int BX, int BY, int K, bool LCS, int NT = S1_THREADS>
static void stage1_shape(hipStream_t s, const BatchDev &B, const TreeDev &T, const LeafCells &LC,
                         const uint32_t *start, const uint32_t *list, uint32_t cap, uint32_t *min_ids)
{
    uint32_t grid = B.nblk < 1 ? 1 : (B.nblk > 65536 ? 65536 : B.nblk);
    hipLaunchKernelGGL((k_stage1<BX, BY, K, LCS, NT>), dim3(grid), dim3(NT), 0, s, B, T, LC,
                       start, list, cap, min_ids);
}

bool stage1_shape_ok(const Stage1Cfg &c)
{
    if (c.lcs > 1)
        return false;
    if (c.bx == 8 && c.by == 8)
        return c.k == 4 || (!c.lcs && (c.k == 2 || c.k == 8));
    if (c.bx == 16 && c.by == 8)
        return c.k == 4;
    return !c.lcs && c.k == 4 &&
           ((c.bx == 16 && c.by == 16) || (c.bx == 32 && c.by == 8) || (c.bx == 32 && c.by == 16));
}

static uint64_t al256(uint64_t v) { return (v + 255) / 256 * 256; }

uint64_t leaf_cells_bytes(uint64_t nleaf)
{
    const uint64_t ns = nleaf * LCS_SLOTS;
    return al256(nleaf * 4) + al256(ns * 8) + al256(ns * 4) + al256(ns * 8);
}

LeafCells carve_leaf_cells(uint8_t *base, uint64_t nleaf)
{
    const uint64_t ns = nleaf * LCS_SLOTS;
    LeafCells c;
    uint64_t o = 0;
    c.ne = (uint32_t *)(base + o);
    o += al256(nleaf * 4);
    c.kc = (unsigned long long *)(base + o);
    o += al256(ns * 8);
    c.id = (uint32_t *)(base + o);
    o += al256(ns * 4);
    c.zk = (unsigned long long *)(base + o);
    return c;
}

void launch_leaf_cells(hipStream_t s, const BatchDev &B, const TreeDev &T, const LeafCells &lc)
{
    uint32_t g = (B.ngrp + 3) / 4;
    g = g < 1 ? 1 : (g > 16384 ? 16384 : g);
    hipLaunchKernelGGL(k_lcs, dim3(g), dim3(256), 0, s, B, T, lc);
}

// work: cnt[nblk] | start[nblk+1] | look-back[...] | tile counter, all zero on
// entry.
void launch_stage1_join(hipStream_t s, const BatchDev &B, const TreeDev &T, const Stage1Cfg &cfg,
                        uint32_t *work, uint32_t *list, uint32_t cap, uint32_t *err,
                        uint32_t *total_out)
{
    if (B.nblk == 0)
        return;
    uint32_t *cnt = work;
    uint32_t *start = work + (B.nblk + 1);
    uint32_t *lookback = start + (B.nblk + 1);
    const uint32_t nt = B.nblk / SC_TILE + 2;
    uint32_t *counter = lookback + nt;
    uint32_t g = (B.nleaf + 255) / 256;
    g = g < 1 ? 1 : (g > 8192 ? 8192 : g);
    hipLaunchKernelGGL(k_s1_bin<0>, dim3(g), dim3(256), 0, s, B, T, cfg.bx, cfg.by, cnt,
                       (const uint32_t *)nullptr, (uint32_t *)nullptr, 0u, err);
    uint32_t gs = (B.nblk + SC_TILE - 1) / SC_TILE;
    hipLaunchKernelGGL(k_scan_excl, dim3(gs < 1 ? 1 : gs), dim3(SC_BLOCK), 0, s, cnt, B.nblk, start,
                       lookback, counter);
    launch_fill_u32(s, cnt, B.nblk, 0u);
    hipLaunchKernelGGL(k_s1_bin<1>, dim3(g), dim3(256), 0, s, B, T, cfg.bx, cfg.by, cnt,
                       (const uint32_t *)start, list, cap, err);
    launch_add_u32(s, start + B.nblk, counter + 1, total_out);
}

void launch_stage1_bins(hipStream_t s, const BatchDev &B, const TreeDev &T, const Stage1Cfg &cfg,
                        const LeafCells *lc, const uint32_t *work, const uint32_t *list,
                        uint32_t cap, uint32_t *min_ids)
{
    if (B.nblk == 0)
        return;
    LeafCells LC{};
    if (lc)
        LC = *lc;
    const uint32_t *start = work + (B.nblk + 1);
    const bool use = lc != nullptr;
    if (cfg.pairs) {
        // window columns: every tile's edges pair up and its records carry columns
        uint32_t grid = B.nblk > 65536 ? 65536 : B.nblk;
        if (cfg.bx == 32)
            hipLaunchKernelGGL((k_stage1_pairs<32, 32, 512>), dim3(grid), dim3(512), 0, s, B, T,
                               start, list, cap, min_ids);
        else
            hipLaunchKernelGGL((k_stage1_pairs<16, 16, 512>), dim3(grid), dim3(512), 0, s, B, T,
                               start, list, cap, min_ids);
        return;
    }
    if (cfg.bx == 16 && cfg.by == 8) {
        if (use)
            stage1_shape<16, 8, 4, true>(s, B, T, LC, start, list, cap, min_ids);
        else
            stage1_shape<16, 8, 4, false>(s, B, T, LC, start, list, cap, min_ids);
    } else if (cfg.bx == 16 && cfg.by == 16) {
        stage1_shape<16, 16, 4, false, 512>(s, B, T, LC, start, list, cap, min_ids);
    } else if (cfg.bx == 32 && cfg.by == 8) {
        stage1_shape<32, 8, 4, false, 512>(s, B, T, LC, start, list, cap, min_ids);
    } else if (cfg.bx == 32 && cfg.by == 16) {
        stage1_shape<32, 16, 4, false, 1024>(s, B, T, LC, start, list, cap, min_ids);
    } else if (cfg.k == 2) {
        stage1_shape<8, 8, 2, false>(s, B, T, LC, start, list, cap, min_ids);
    } else if (cfg.k == 8) {
        stage1_shape<8, 8, 8, false>(s, B, T, LC, start, list, cap, min_ids);
    } else if (use) {
        stage1_shape<8, 8, 4, true>(s, B, T, LC, start, list, cap, min_ids);
    } else {
        stage1_shape<8, 8, 4, false>(s, B, T, LC, start, list, cap, min_ids);
    }
}

}  // namespace owm
