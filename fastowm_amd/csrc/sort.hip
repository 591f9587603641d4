// LSD radix sort of u32 keys (optionally with u32 values) for gfx950.
//
// One upfront histogram pass counts the digits of every pass; each pass is
// then a single "onesweep" kernel: a workgroup takes the next tile of keys
// from an atomic counter (so every predecessor tile is already resident),
// ranks its keys stably with wave64 ballot matching, publishes its digit
// counts, resolves its global digit offsets by decoupled look-back over the
// predecessors' published counts, stages the tile in LDS in digit order and
// writes it out so consecutive lanes write consecutive addresses.
//
// Stability: keys are ranked in (wave, item, lane) order, which is input order,
// so equal keys keep their input order and the sort equals std::stable_sort.
//
// Segmented form: the tiles of a batch are contiguous segments of the input
// that are sorted each on its own.  Sort tiles never straddle a segment, every
// segment has its own digit offsets, and a segment's first sort tile starts
// its look-back chain, so the keys hold Morton bits only and every pass
// permutes within a segment.
//
// Record form (the tree sort): every pass moves whole 32-byte point records
// (x, y, z, batch index, key); the first reads the caller's coordinate arrays
// and computes each point's key, so the cloud ends in Morton order without a
// gather by index (random 32-byte reads run at a fraction of HBM bandwidth).
#include "owm_kernels.h"

namespace owm {

constexpr int RB = 8;
constexpr int RADIX = 1 << RB;
constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;
// Keys per thread: every tile costs one same-address atomic (the tile
// counter, which serialises device-wide at roughly 12 ns per increment) and
// one look-back, so large sorts use 8192-key tiles; small sorts keep 4096-key
// tiles for more workgroups.
constexpr uint32_t BIG_SORT = 1u << 21;
// key/value sort tile: 8192 keys per 256 threads from BIG_SORT keys; smaller
// sorts (window minima, seeds) take 2048- or 1024-key tiles so that their few
// hundred thousand keys still spread over every CU
static inline int sort_items(uint32_t n) { return n >= BIG_SORT ? 32 : (n >= (1u << 18) ? 8 : 4); }
constexpr uint32_t SEG_HIST = SORT_SEG_HIST;
static_assert(SEG_HIST == 4 * RADIX, "four passes of 8-bit digits per segment");
constexpr uint32_t LB_AGG = 1u << 30;
constexpr uint32_t LB_INC = 2u << 30;
constexpr uint32_t LB_CNT = (1u << 30) - 1;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o)
            v += t;
    }
    return v;
}

// Exclusive scan of one value per thread over a block of NW waves.
template <int NW = WAVES>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_tmp)
{
    uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t inc = wave_incl_scan(v, lane);
    if (lane == 63)
        s_tmp[w] = inc;
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t k = 0; k < w; ++k)
        add += s_tmp[k];
    __syncthreads();
    return add + inc - v;
}

__global__ __launch_bounds__(BLOCK) void k_sort_hist(const uint32_t *__restrict__ keys,
                                                      const uint32_t *d_n, uint32_t n_host,
                                                      int passes, uint32_t *__restrict__ hist)
{
    __shared__ uint32_t h[4][RADIX];
    uint32_t n = d_n ? *d_n : n_host;
    for (int i = threadIdx.x; i < 4 * RADIX; i += BLOCK)
        (&h[0][0])[i] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        uint32_t k = keys[i];
        for (int p = 0; p < passes; ++p)
            atomicAdd(&h[p][(k >> (p * RB)) & (RADIX - 1)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * RADIX; i += BLOCK) {
        uint32_t c = (&h[0][0])[i];
        if (c)
            atomicAdd(&hist[i], c);
    }
}

// Per segment t (one workgroup each): counts[t][p][d] -> hist[t][p][d] =
// pt[t] + exclusive offsets within the segment (counts may be hist), and
// tile_seg[u] = t for the segment's sort tiles; workgroup 0 also zeroes the
// tile counters.
__global__ __launch_bounds__(BLOCK) void k_sort_scan_seg(const uint32_t *counts, uint32_t *hist,
                                                          const uint32_t *pt,
                                                          const uint32_t *units, uint32_t nseg,
                                                          int passes, uint32_t *counters,
                                                          uint32_t *tile_seg)
{
    __shared__ uint32_t tmp[WAVES];
    for (uint32_t t = blockIdx.x; t < nseg; t += gridDim.x) {
        for (uint32_t u = units[t] + threadIdx.x; u < units[t + 1]; u += BLOCK)
            tile_seg[u] = t;
        const uint32_t *c = counts + (size_t)t * SEG_HIST;
        uint32_t *h = hist + (size_t)t * SEG_HIST;
        const uint32_t b = pt[t];
        for (int p = 0; p < passes; ++p) {
            uint32_t v = c[p * RADIX + threadIdx.x];
            uint32_t e = block_excl_scan(v, tmp);
            h[p * RADIX + threadIdx.x] = b + e;
            __syncthreads();
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < 8)
        counters[threadIdx.x] = 0;
}

// hist[p][d] -> exclusive offsets; also zeroes the per-pass tile counters.
__global__ __launch_bounds__(BLOCK) void k_sort_scan(uint32_t *hist, int passes, uint32_t *counters)
{
    __shared__ uint32_t tmp[WAVES];
    for (int p = 0; p < passes; ++p) {
        uint32_t v = hist[p * RADIX + threadIdx.x];
        uint32_t e = block_excl_scan(v, tmp);
        hist[p * RADIX + threadIdx.x] = e;
        __syncthreads();
    }
    if (threadIdx.x < 8)
        counters[threadIdx.x] = 0;
}

// LDS state of one onesweep tile (a block of NW waves).
template <int NW = WAVES>
struct SweepLds {
    uint32_t whist[NW][RADIX];  // per-wave digit counts, then wave-exclusive prefixes
    uint32_t excl[RADIX];       // digit start within the tile
    uint32_t glob[RADIX];       // digit start in the output minus excl
    uint32_t tmp[NW];
    uint32_t tile;
};

// Takes the next tile from the counter and resolves its segment: input range
// [base, n), digit offsets doff and whether it starts its look-back chain.
// Returns false (uniformly) when no tile is left.
template <int TILE_T, int NW>
__device__ __forceinline__ bool sweep_claim(SweepLds<NW> &L, uint32_t *tile_counter, uint32_t n_all,
                                            const SortSegs &seg, const uint32_t *digit_off,
                                            uint32_t &tile, uint32_t &base, uint32_t &n, bool &first,
                                            const uint32_t *&doff, uint32_t *seg_index = nullptr)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t ntiles = seg.pt ? seg.units[seg.nseg] : (n_all + TILE_T - 1) / TILE_T;
    if (tid == 0)
        L.tile = atomicAdd(tile_counter, 1u);
    for (int i = tid; i < NW * RADIX; i += NW * 64)
        (&L.whist[0][0])[i] = 0;
    __syncthreads();
    tile = L.tile;
    if (tile >= ntiles)
        return false;
    doff = digit_off;
    if (seg.pt) {
        const uint32_t t = seg.tile_seg ? seg.tile_seg[tile] : owm_find_tile(seg.units, seg.nseg, tile);
        base = seg.pt[t] + (tile - seg.units[t]) * TILE_T;
        n = seg.pt[t + 1];
        first = tile == seg.units[t];
        doff = digit_off + (size_t)t * seg.digit_stride;
        if (seg_index)
            *seg_index = t;
    } else {
        base = tile * TILE_T;
        n = n_all;
        first = tile == 0;
        if (seg_index)
            *seg_index = 0;
    }
    return true;
}

// Stable in-wave ranks: item k of lane l of wave w is element
// base + w*64*ITEMS_T + k*64 + l; rank[k] counts the wave's earlier elements of
// the same digit, and L.whist[w][d] ends as the wave's count of digit d.  FULL:
// the tile holds ITEMS_T keys per thread (no element is past n).  PACK: rank[k]
// is d << 16 | rank, and rank may be key itself (each key is read before its
// rank is written); HI: the key's byte above the digit goes to bits 24-31.
template <int ITEMS_T, bool FULL = false, bool PACK = false, int NW = WAVES, bool HI = false>
__device__ __forceinline__ void sweep_rank(SweepLds<NW> &L, const uint32_t (&key)[ITEMS_T],
                                           uint32_t base, uint32_t n, int shift,
                                           uint32_t (&rank)[ITEMS_T])
{
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < ITEMS_T; ++k) {
        const uint32_t idx = base + w * (64 * ITEMS_T) + k * 64 + lane;
        const bool valid = FULL || idx < n;
        const uint32_t d = (key[k] >> shift) & (RADIX - 1);
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t before = __popcll(m & lt_mask);
        const uint32_t old = L.whist[w][d];
        if (valid && before == 0)
            L.whist[w][d] = old + (uint32_t)__popcll(m);
        const uint32_t hi = HI ? ((key[k] >> (shift + RB)) & 255u) << 24 : 0u;
        rank[k] = PACK ? hi | d << 16 | (old + before) : old + before;
    }
}

// After sweep_rank and a barrier: wave-exclusive prefixes per digit, the tile's
// digit counts published and its global digit bases resolved by decoupled
// look-back (thread d < RADIX handles digit d).  Ends with a barrier; then a key
// of digit d and rank r in wave w sits at tile slot
// s = excl[d] + whist[w][d] + r and goes to output position glob[d] + s (glob
// holds the digit's output start minus its tile start, modulo 2^32).
template <int NW>
__device__ __forceinline__ void sweep_offsets(SweepLds<NW> &L, uint32_t tile, bool first,
                                              const uint32_t *doff, uint32_t *lookback)
{
    const uint32_t d = threadIdx.x;
    uint32_t total = 0, prefix = 0;
    if (d < RADIX) {
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t c = L.whist[w][d];
            L.whist[w][d] = total;
            total += c;
        }
        uint32_t *flag = lookback + (size_t)tile * RADIX + d;
        if (first) {
            __hip_atomic_store(flag, LB_INC | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(flag, LB_AGG | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // walk back over the predecessors four flags at a time (their
            // loads in flight together), in order: add aggregates until an
            // inclusive prefix; at a predecessor still ranking (it holds an
            // earlier tile, so it is resident) poll again from there.  The
            // walk ends at the latest at the segment's first tile (inclusive),
            // so indices below it are loaded only as spare reads.
            int32_t j = (int32_t)tile - 1;
            for (bool done = false; !done;) {
                uint32_t f[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    f[q] = __hip_atomic_load(lookback + (size_t)max(j - q, 0) * RADIX + d,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                int q = 0;
                for (; q < 4; ++q) {
                    const uint32_t st = f[q] & ~LB_CNT;
                    if (st == 0)
                        break;
                    prefix += f[q] & LB_CNT;
                    if (st == LB_INC) {
                        done = true;
                        break;
                    }
                }
                j -= q;
            }
            __hip_atomic_store(flag, LB_INC | (prefix + total), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // waves past the digits add zeros after them, so digit prefixes are exact
    const uint32_t ex = block_excl_scan<NW>(total, L.tmp);
    if (d < RADIX) {
        L.excl[d] = ex;
        L.glob[d] = doff[d] + prefix - ex;
    }
    __syncthreads();
}

template <bool KEYS_ONLY, bool IOTA, int ITEMS_T>
__global__ __launch_bounds__(BLOCK) void k_onesweep(const uint32_t *__restrict__ keys_in,
                                                     const uint32_t *__restrict__ vals_in,
                                                     uint32_t *__restrict__ keys_out,
                                                     uint32_t *__restrict__ vals_out,
                                                     const uint32_t *d_n, uint32_t n_host, int shift,
                                                     const uint32_t *__restrict__ digit_off,
                                                     uint32_t *lookback, uint32_t *tile_counter,
                                                     int write_keys, SortSegs seg)
{
    constexpr int TILE_T = BLOCK * ITEMS_T;
    __shared__ SweepLds<> L;
    __shared__ uint32_t s_keys[TILE_T];
    __shared__ uint32_t s_vals[KEYS_ONLY ? 1 : TILE_T];

    const uint32_t n_all = d_n ? *d_n : n_host;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t tile, base, n;
    bool first;
    const uint32_t *doff;
    while (sweep_claim<TILE_T>(L, tile_counter, n_all, seg, digit_off, tile, base, n, first, doff)) {
        uint32_t key[ITEMS_T], val[ITEMS_T], rank[ITEMS_T];
#pragma unroll
        for (int k = 0; k < ITEMS_T; ++k) {
            uint32_t idx = base + w * (64 * ITEMS_T) + k * 64 + lane;
            bool valid = idx < n;
            key[k] = valid ? keys_in[idx] : 0u;
            if (!KEYS_ONLY)
                val[k] = valid ? (IOTA ? idx : vals_in[idx]) : 0u;
        }
        sweep_rank<ITEMS_T>(L, key, base, n, shift, rank);
        __syncthreads();
        sweep_offsets(L, tile, first, doff, lookback);
#pragma unroll
        for (int k = 0; k < ITEMS_T; ++k) {
            uint32_t idx = base + w * (64 * ITEMS_T) + k * 64 + lane;
            if (idx < n) {
                uint32_t dd = (key[k] >> shift) & (RADIX - 1);
                uint32_t slot = L.excl[dd] + L.whist[w][dd] + rank[k];
                s_keys[slot] = key[k];
                if (!KEYS_ONLY)
                    s_vals[slot] = val[k];
            }
        }
        __syncthreads();
        const uint32_t cnt = min((uint32_t)TILE_T, n - base);
        for (uint32_t s = tid; s < cnt; s += BLOCK) {
            uint32_t k = s_keys[s];
            uint32_t dd = (k >> shift) & (RADIX - 1);
            uint32_t dst = L.glob[dd] + s;
            if (write_keys)
                keys_out[dst] = k;
            if (!KEYS_ONLY)
                vals_out[dst] = s_vals[s];
        }
        __syncthreads();
    }
}

// One pass of the segmented tree sort (the segments are the tiles of a
// batch).  Whole 32-byte point records move: a workgroup loads its sort
// tile's records into registers (the first pass reads the caller's x, y, z,
// numbers the points and computes their keys from their tile's quantisation),
// ranks them by the pass's digit (stable: item k of lane l of wave w is tile
// position w*64*ITEMS + k*64 + l, ranked in that order), resolves the digits'
// output bases by decoupled look-back, and then writes the tile out STAGE
// slots at a time: the stage's points go to LDS in slot (digit) order, then
// consecutive lanes write consecutive slots' 16-byte (x, y) pairs and (z, id,
// aux) triples, so each digit's run leaves as two contiguous streams.
// MODE REC_POS: pos[id] = the point's output position.  MODE REC_STAMP (the
// last pass of a tree built with its query): in tiles with gcls set, aux
// leaves as the point's global window-edge classes (owm_gstamp, computed from
// the staged coordinates as the slot is written) instead of its key; MODE
// REC_COLS: its window columns (owm_cstamp, tiles whose window side is a
// whole number of steps).
constexpr int REC_PLAIN = 0, REC_POS = 1, REC_STAMP = 2, REC_COLS = 3;
template <bool FIRST, int MODE, int NW, int ITEMS, int STAGE>
__global__ __launch_bounds__(NW * 64) void k_sort_records(PointsIn in, PointsOut out,
                                                          uint32_t *__restrict__ pos,
                                                          const TileDev *__restrict__ tiles,
                                                          uint32_t key_shift, uint32_t n_all,
                                                          int shift,
                                                          const uint32_t *__restrict__ digit_off,
                                                          uint32_t *lookback, uint32_t *tile_counter,
                                                          SortSegs seg)
{
    constexpr int TILE_T = NW * 64 * ITEMS;
    constexpr bool POS = MODE == REC_POS, STAMP = MODE == REC_STAMP || MODE == REC_COLS;
    static_assert(TILE_T % STAGE == 0, "a sort tile is a whole number of stages");
    __shared__ SweepLds<NW> L;
    __shared__ uint4 s_xy[STAGE], s_zi[STAGE];  // a stage of points in slot order
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t p0 = w * (64 * ITEMS) + lane;  // tile position of the thread's item 0
    uint4 *const oxy = reinterpret_cast<uint4 *>(out.xy);
    uint4 *const ozi = reinterpret_cast<uint4 *>(out.zi);
    uint32_t tile, base, n, t;
    bool first;
    const uint32_t *doff;
    while (sweep_claim<TILE_T, NW>(L, tile_counter, n_all, seg, digit_off, tile, base, n, first, doff,
                                   &t)) {
        const uint32_t cnt = min((uint32_t)TILE_T, n - base);
        double x[ITEMS], y[ITEMS], z[ITEMS];
        uint32_t id[ITEMS], key[ITEMS];
        // loads clamped to the tile's last point (items past it are unused)
        if constexpr (FIRST) {
            const TileDev &TD = tiles[t];
#pragma unroll
            for (int k = 0; k < ITEMS; ++k) {
                const uint32_t i = base + min(p0 + k * 64, cnt - 1);
                x[k] = in.x[i];
                y[k] = in.y[i];
                z[k] = in.z[i];  // the point's index i is recomputed when staged
            }
#pragma unroll
            for (int k = 0; k < ITEMS; ++k)
                key[k] = sort_key(TD, x[k], y[k], key_shift);
        } else {
#pragma unroll
            for (int k = 0; k < ITEMS; ++k) {
                const PointRec r = owm_load_pt(in.rec, base + min(p0 + k * 64, cnt - 1));
                x[k] = r.x;
                y[k] = r.y;
                z[k] = r.z;
                id[k] = r.id;
                key[k] = r.aux;
            }
        }
        uint32_t slot[ITEMS];
        sweep_rank<ITEMS, false, false, NW>(L, key, 0, cnt, shift, slot);
        __syncthreads();
        sweep_offsets(L, tile, first, doff, lookback);
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const uint32_t d = (key[k] >> shift) & (RADIX - 1);
            // items past the tile's end take the unused slots past cnt
            slot[k] = p0 + k * 64 < cnt ? L.excl[d] + L.whist[w][d] + slot[k] : p0 + k * 64;
        }
        // the last pass of a tree built with its query stamps classes
        const TileDev &TS = tiles[__builtin_amdgcn_readfirstlane(t)];
        const bool gstamp = STAMP && TS.gcls;
        for (uint32_t lo = 0; lo < cnt; lo += STAGE) {
#pragma unroll
            for (int k = 0; k < ITEMS; ++k) {
                const uint32_t r = slot[k] - lo;
                if (r < (uint32_t)STAGE) {
                    const unsigned long long bx = (unsigned long long)__double_as_longlong(x[k]);
                    const unsigned long long by = (unsigned long long)__double_as_longlong(y[k]);
                    const unsigned long long bz = (unsigned long long)__double_as_longlong(z[k]);
                    s_xy[r] = make_uint4((uint32_t)bx, (uint32_t)(bx >> 32), (uint32_t)by,
                                         (uint32_t)(by >> 32));
                    const uint32_t pid = FIRST ? base + min(p0 + k * 64, cnt - 1) : id[k];
                    s_zi[r] = make_uint4((uint32_t)bz, (uint32_t)(bz >> 32), pid, key[k]);
                }
            }
            __syncthreads();
            // each slot's (x, y) pair and (z, id, aux) triple: consecutive
            // lanes write consecutive 16-byte elements of both streams
            const uint32_t m = min((uint32_t)STAGE, cnt - lo);
            for (uint32_t r = tid; r < m; r += NW * 64) {
                uint4 v = s_zi[r];
                const uint32_t d = (v.w >> shift) & (RADIX - 1);
                const uint32_t dst = L.glob[d] + lo + r;
                const uint4 p = s_xy[r];
                oxy[dst] = p;
                if (gstamp) {
                    const double px = __hiloint2double((int)p.y, (int)p.x);
                    const double py = __hiloint2double((int)p.w, (int)p.z);
                    v.w = MODE == REC_COLS ? owm_cstamp(TS, px, py) : owm_gstamp(TS, px, py);
                }
                ozi[dst] = v;
                if (POS)
                    pos[v.z] = dst;
            }
            __syncthreads();
        }
    }
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

uint32_t sort_tiles(uint32_t n)
{
    const uint32_t tile = (uint32_t)(BLOCK * sort_items(n));
    return (n + tile - 1) / tile;
}

// One range per blockIdx.y.
__global__ void k_fill_ranges(FillRanges r)
{
    uint32_t *p = r.ptr[blockIdx.y];
    const uint64_t n = r.len[blockIdx.y];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

void launch_fill_ranges(hipStream_t s, const FillRanges &r)
{
    uint64_t mx = 0;
    for (uint32_t k = 0; k < r.count; ++k)
        mx = r.len[k] > mx ? r.len[k] : mx;
    if (r.count == 0 || mx == 0)
        return;
    uint64_t blocks = (mx + 255) / 256;
    if (blocks > 1024)
        blocks = 1024;
    hipLaunchKernelGGL(k_fill_ranges, dim3((unsigned)blocks, r.count), dim3(256), 0, s, r);
}

void launch_fill_u32(hipStream_t s, uint32_t *p, uint64_t n, uint32_t v)
{
    if (n == 0)
        return;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096)
        blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

void launch_sort_hist(hipStream_t s, const uint32_t *keys, const uint32_t *d_n, uint32_t n_max,
                      int passes, uint32_t *hist)
{
    uint32_t blocks = (n_max + BLOCK * 16 - 1) / (BLOCK * 16);
    blocks = blocks < 1 ? 1 : (blocks > 2048 ? 2048 : blocks);
    hipLaunchKernelGGL(k_sort_hist, dim3(blocks), dim3(BLOCK), 0, s, keys, d_n, n_max, passes,
                       hist);
}

void launch_sort_scan(hipStream_t s, uint32_t *hist, int passes, uint32_t *counters)
{
    hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(BLOCK), 0, s, hist, passes, counters);
}

void launch_sort_scan_seg(hipStream_t s, const uint32_t *counts, uint32_t *hist, const uint32_t *pt,
                          const uint32_t *units, uint32_t nseg, int passes, uint32_t *counters,
                          uint32_t *tile_seg)
{
    const uint32_t g = nseg < 1 ? 1 : (nseg > 4096 ? 4096 : nseg);
    hipLaunchKernelGGL(k_sort_scan_seg, dim3(g), dim3(BLOCK), 0, s, counts, hist, pt, units, nseg,
                       passes, counters, tile_seg);
}


template <int IT>
static void onesweep_items(hipStream_t s, const uint32_t *keys_in, const uint32_t *vals_in,
                           uint32_t *keys_out, uint32_t *vals_out, const uint32_t *d_n,
                           uint32_t n_max, int shift, const uint32_t *digit_off, uint32_t *lookback,
                           uint32_t *tile_counter, bool keys_only, bool iota, int wk, uint32_t grid,
                           const SortSegs &seg)
{
    if (keys_only)
        hipLaunchKernelGGL((k_onesweep<true, false, IT>), dim3(grid), dim3(BLOCK), 0, s, keys_in,
                           vals_in, keys_out, vals_out, d_n, n_max, shift, digit_off, lookback,
                           tile_counter, wk, seg);
    else if (iota)
        hipLaunchKernelGGL((k_onesweep<false, true, IT>), dim3(grid), dim3(BLOCK), 0, s, keys_in,
                           vals_in, keys_out, vals_out, d_n, n_max, shift, digit_off, lookback,
                           tile_counter, wk, seg);
    else
        hipLaunchKernelGGL((k_onesweep<false, false, IT>), dim3(grid), dim3(BLOCK), 0, s, keys_in,
                           vals_in, keys_out, vals_out, d_n, n_max, shift, digit_off, lookback,
                           tile_counter, wk, seg);
}

void launch_onesweep(hipStream_t s, const uint32_t *keys_in, const uint32_t *vals_in,
                     uint32_t *keys_out, uint32_t *vals_out, const uint32_t *d_n, uint32_t n_max,
                     int shift, const uint32_t *digit_off, uint32_t *lookback,
                     uint32_t *tile_counter, bool keys_only, bool iota, bool write_keys,
                     uint32_t grid_cap, const SortSegs &seg, uint32_t seg_tiles)
{
    // the tile size follows n_max; a device count d_n <= n_max uses the same
    uint32_t tiles = seg.pt ? seg_tiles : sort_tiles(n_max);
    uint32_t grid = tiles < 1 ? 1 : tiles;
    if (grid_cap && grid > grid_cap)
        grid = grid_cap;
    int wk = write_keys ? 1 : 0;
    const int it = sort_items(n_max);
    if (it == 32)
        onesweep_items<32>(s, keys_in, vals_in, keys_out, vals_out, d_n, n_max, shift, digit_off,
                           lookback, tile_counter, keys_only, iota, wk, grid, seg);
    else if (it == 8)
        onesweep_items<8>(s, keys_in, vals_in, keys_out, vals_out, d_n, n_max, shift, digit_off,
                          lookback, tile_counter, keys_only, iota, wk, grid, seg);
    else
        onesweep_items<4>(s, keys_in, vals_in, keys_out, vals_out, d_n, n_max, shift, digit_off,
                          lookback, tile_counter, keys_only, iota, wk, grid, seg);
}

// Tree-sort shapes: 8192-point sort tiles (8 waves x 16 points, two stages
// of 4096 records in LDS) for sorts of >= BIG_SORT points whose segments
// average >= 64K points; 4096-point tiles (8 waves x 8, one stage) otherwise,
// so that small sorts spread over more workgroups and short segments waste
// less of their last tile.
static inline bool big_records(uint32_t n, uint32_t nseg)
{
    return n >= BIG_SORT && n / (nseg ? nseg : 1u) >= 65536u;
}
uint32_t record_tile_keys(uint32_t n, uint32_t nseg) { return big_records(n, nseg) ? 8192u : 4096u; }

template <int NW, int ITEMS, int STAGE>
static void sort_records_shape(hipStream_t s, const BatchDev &B, const PointsIn &in, const PointsOut &out,
                               uint32_t *pos, uint32_t n, int shift, uint32_t key_shift,
                               const uint32_t *digit_off, uint32_t *lookback, uint32_t *tile_counter,
                               bool first, int stamp, const SortSegs &seg, uint32_t grid)
{
    const dim3 b(NW * 64);
    const int mode = pos ? REC_POS : (stamp == 2 ? REC_COLS : (stamp ? REC_STAMP : REC_PLAIN));
#define OWM_REC(FIRST, MODE)                                                                        \
    hipLaunchKernelGGL((k_sort_records<FIRST, MODE, NW, ITEMS, STAGE>), dim3(grid), b, 0, s, in, out, \
                       pos, B.tiles, key_shift, n, shift, digit_off, lookback, tile_counter, seg)
    if (first) {
        if (mode == REC_POS)
            OWM_REC(true, REC_POS);
        else if (mode == REC_STAMP)
            OWM_REC(true, REC_STAMP);
        else if (mode == REC_COLS)
            OWM_REC(true, REC_COLS);
        else
            OWM_REC(true, REC_PLAIN);
    } else {
        if (mode == REC_POS)
            OWM_REC(false, REC_POS);
        else if (mode == REC_STAMP)
            OWM_REC(false, REC_STAMP);
        else if (mode == REC_COLS)
            OWM_REC(false, REC_COLS);
        else
            OWM_REC(false, REC_PLAIN);
    }
#undef OWM_REC
}

void launch_sort_records(hipStream_t s, const BatchDev &B, const PointsIn &in, const PointsOut &out,
                         uint32_t *pos, uint32_t n, int shift, int key_bits,
                         const uint32_t *digit_off, uint32_t *lookback, uint32_t *tile_counter,
                         bool first, int stamp, const SortSegs &seg, uint32_t seg_tiles)
{
    const uint32_t grid = seg_tiles < 1 ? 1 : seg_tiles;
    const uint32_t key_shift = 64u - (uint32_t)key_bits;
    if (big_records(n, seg.nseg))
        sort_records_shape<8, 16, 4096>(s, B, in, out, pos, n, shift, key_shift, digit_off, lookback,
                                        tile_counter, first, stamp, seg, grid);
    else
        sort_records_shape<8, 8, 4096>(s, B, in, out, pos, n, shift, key_shift, digit_off, lookback,
                                       tile_counter, first, stamp, seg, grid);
}

}  // namespace owm
