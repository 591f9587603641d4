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
// Record form (the tree sort): every pass moves the points' x, y, z and batch
// index with the keys, the first reading the caller's coordinate arrays and
// the last writing the tree's (without keys: the tree recomputes the few it
// needs), so the cloud ends in Morton order without a gather by index (random
// 32-byte reads run at a fraction of HBM bandwidth).
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

// Per segment t (one workgroup each): hist[t][p][d] -> pt[t] + exclusive
// offsets within the segment; workgroup 0 also zeroes the tile counters.
__global__ __launch_bounds__(BLOCK) void k_sort_scan_seg(uint32_t *hist, const uint32_t *pt,
                                                          uint32_t nseg, int passes,
                                                          uint32_t *counters)
{
    __shared__ uint32_t tmp[WAVES];
    for (uint32_t t = blockIdx.x; t < nseg; t += gridDim.x) {
        uint32_t *h = hist + (size_t)t * SEG_HIST;
        const uint32_t b = pt[t];
        for (int p = 0; p < passes; ++p) {
            uint32_t v = h[p * RADIX + threadIdx.x];
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
                                            const uint32_t *&doff)
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
        const uint32_t t = owm_find_tile(seg.units, seg.nseg, tile);
        base = seg.pt[t] + (tile - seg.units[t]) * TILE_T;
        n = seg.pt[t + 1];
        first = tile == seg.units[t];
        doff = digit_off + (size_t)t * seg.digit_stride;
    } else {
        base = tile * TILE_T;
        n = n_all;
        first = tile == 0;
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
            int32_t j = (int32_t)tile - 1;
            for (;;) {
                const uint32_t f = __hip_atomic_load(lookback + (size_t)j * RADIX + d, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t st = f & ~LB_CNT;
                if (st == 0)
                    continue;  // predecessor still ranking (it holds an earlier tile, so it is resident)
                prefix += f & LB_CNT;
                if (st == LB_INC)
                    break;
                --j;
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

// One tile of a record-moving pass (see k_sort_records).  FULL: the tile holds
// TILE_T points, so no load or store needs a bound check; otherwise loads are
// clamped to the tile's last point and the items past it take the unused tile
// slots past cnt, so that the LDS writes need no check either.
// K16: 16-bit keys sorted by their low byte (the first of two passes): a key is
// its digit and the byte above it, kept from the ranking, so the keys are not
// read a second time for the key/index staging.
template <bool IOTA, bool LAST, bool POS, bool K16, int ITEMS_T, int NW, bool FULL>
__device__ __forceinline__ void records_tile(SweepLds<NW> &L, double *s_buf, uint8_t *s_dig,
                                             const uint32_t *__restrict__ keys_in, const RecIn &in,
                                             uint32_t *__restrict__ keys_out, const RecOut &out,
                                             uint32_t tile, uint32_t base, uint32_t cnt, bool first,
                                             const uint32_t *doff, uint32_t *lookback, int shift)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t p0 = w * (64 * ITEMS_T) + lane;  // tile position of the thread's item 0
    unsigned long long *const s_u64 = reinterpret_cast<unsigned long long *>(s_buf);
    // element of item k; o is p0 passed through an empty asm, so that every
    // field computes its own addresses (otherwise the compiler keeps the 64-bit
    // addresses of all items live from one field to the next, and spills)
    auto fresh = [&]() {
        uint32_t o = p0;
        asm volatile("" : "+v"(o));
        return o;
    };
    auto at = [&](uint32_t o, int k) { return base + (FULL ? o + k * 64 : min(o + k * 64, cnt - 1)); };
    double v[ITEMS_T];
    auto load = [&](const double *__restrict__ src) {
        const uint32_t o = fresh();
#pragma unroll
        for (int k = 0; k < ITEMS_T; ++k)
            v[k] = src[at(o, k)];
    };
    // per item its tile slot (16 bits, two per word); K16: the key's high
    // byte (four per word)
    uint32_t slot2[ITEMS_T / 2], hi4[K16 ? ITEMS_T / 4 : 1];
    {
        // keys, then (in place) [high byte << 24 |] digit << 16 | in-wave rank
        uint32_t kr[ITEMS_T];
        const uint32_t o = fresh();
#pragma unroll
        for (int k = 0; k < ITEMS_T; ++k)
            kr[k] = keys_in[at(o, k)];
        load(in.x);
        sweep_rank<ITEMS_T, FULL, true, NW, K16>(L, kr, 0, cnt, shift, kr);
        __syncthreads();
        sweep_offsets(L, tile, first, doff, lookback);
#pragma unroll
        for (int k = 0; k < ITEMS_T; ++k) {
            const uint32_t dd = (kr[k] >> 16) & 255u;
            if constexpr (K16) {
                const uint32_t h = kr[k] >> 24;
                hi4[k / 4] = (k & 3) ? hi4[k / 4] | h << (8 * (k & 3)) : h;
            }
            uint32_t sl = L.excl[dd] + L.whist[w][dd] + (kr[k] & 0xFFFFu);
            if (!FULL)
                sl = p0 + k * 64 < cnt ? sl : p0 + k * 64;
            s_dig[sl] = (uint8_t)dd;
            s_buf[sl] = v[k];
            if (k & 1)
                slot2[k / 2] |= sl << 16;
            else
                slot2[k / 2] = sl;
        }
    }
    __syncthreads();
    auto slot_of = [&](int k) { return (slot2[k / 2] >> (16 * (k & 1))) & 0xFFFFu; };
    // the staged field goes out (then a barrier); the next one, already
    // loaded, comes in (then a barrier).  Destinations are looked up per slot
    // (digit, then the digit's base) in a rolled loop: unrolled, the compiler
    // hoists every lookup and address and runs out of registers.
    auto store = [&](double *__restrict__ to) {
#pragma unroll 2
        for (uint32_t s = tid; s < cnt; s += NW * 64)
            to[L.glob[s_dig[s]] + s] = s_buf[s];
        __syncthreads();
    };
    auto stage = [&]() {
#pragma unroll
        for (int k = 0; k < ITEMS_T; ++k)
            s_buf[slot_of(k)] = v[k];
        __syncthreads();
    };
    load(in.y);
    store(out.x);
    stage();
    load(in.z);
    store(out.y);
    stage();
    if constexpr (LAST) {
        // the last pass writes no keys: the index alone
        const uint32_t o = fresh();
        uint32_t id[ITEMS_T];
#pragma unroll
        for (int k = 0; k < ITEMS_T; ++k)
            id[k] = IOTA ? base + o + k * 64 : in.id[at(o, k)];
        store(out.z);
        uint32_t *const s_u32 = reinterpret_cast<uint32_t *>(s_buf);
#pragma unroll
        for (int k = 0; k < ITEMS_T; ++k)
            s_u32[slot_of(k)] = id[k];
        __syncthreads();
        for (uint32_t s = tid; s < cnt; s += NW * 64) {
            const uint32_t v = s_u32[s], dst = L.glob[s_dig[s]] + s;
            out.id[dst] = v;
            if (POS)
                out.pos[v] = dst;
        }
        __syncthreads();
        return;
    }
    // key and index travel together as one 8-byte word
    {
        unsigned long long kv[ITEMS_T];
        const uint32_t o = fresh();
#pragma unroll
        for (int k = 0; k < ITEMS_T; ++k) {
            const uint32_t id = IOTA ? base + o + k * 64 : in.id[at(o, k)];
            uint32_t key;
            if constexpr (K16)
                key = ((hi4[k / 4] >> (8 * (k & 3))) & 255u) << 8 | s_dig[slot_of(k)];
            else
                key = keys_in[at(o, k)];
            kv[k] = ((unsigned long long)id << 32) | key;
        }
        store(out.z);
#pragma unroll
        for (int k = 0; k < ITEMS_T; ++k)
            s_u64[slot_of(k)] = kv[k];
        __syncthreads();
    }
    for (uint32_t s = tid; s < cnt; s += NW * 64) {
        const unsigned long long kv = s_u64[s];
        const uint32_t dst = L.glob[s_dig[s]] + s;
        keys_out[dst] = (uint32_t)kv;
        out.id[dst] = (uint32_t)(kv >> 32);
    }
    __syncthreads();
}

// One pass of the segmented tree sort that moves whole point records with the
// keys, so that the last pass leaves the cloud in Morton order and no gather
// by index follows.  Records are x, y, z (8 bytes each) and the point's batch
// index.  The first pass reads the caller's coordinate arrays and numbers the
// points itself (IOTA); the last writes the tree's arrays (and, with POS, each
// point's sorted position at its index; LAST: no keys are written).  Ranks come
// from the keys alone; each
// field is then staged through one LDS buffer in digit order and written out,
// so consecutive lanes write consecutive addresses.  The loads of a field are
// issued before the stores of the previous one (those of x before the
// ranking), so that their latency hides behind that work.
template <bool IOTA, bool LAST, bool POS, bool K16, int ITEMS_T, int NW>
__global__ __launch_bounds__(NW * 64, 8 / NW) void k_sort_records(const uint32_t *__restrict__ keys_in,
                                                                   RecIn in, uint32_t *__restrict__ keys_out,
                                                                   RecOut out, uint32_t n_all, int shift,
                                                                   const uint32_t *__restrict__ digit_off,
                                                                   uint32_t *lookback,
                                                                   uint32_t *tile_counter, SortSegs seg)
{
    constexpr int TILE_T = NW * 64 * ITEMS_T;
    __shared__ SweepLds<NW> L;
    __shared__ double s_buf[TILE_T];
    __shared__ uint8_t s_dig[TILE_T];  // digit of every tile slot
    uint32_t tile, base, n;
    bool first;
    const uint32_t *doff;
    while (sweep_claim<TILE_T, NW>(L, tile_counter, n_all, seg, digit_off, tile, base, n, first, doff)) {
        if (n - base >= (uint32_t)TILE_T)
            records_tile<IOTA, LAST, POS, K16, ITEMS_T, NW, true>(L, s_buf, s_dig, keys_in, in, keys_out, out,
                                                       tile, base, TILE_T, first, doff, lookback, shift);
        else
            records_tile<IOTA, LAST, POS, K16, ITEMS_T, NW, false>(L, s_buf, s_dig, keys_in, in, keys_out, out,
                                                        tile, base, n - base, first, doff, lookback,
                                                        shift);
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

void launch_sort_scan_seg(hipStream_t s, uint32_t *hist, const uint32_t *pt, uint32_t nseg,
                          int passes, uint32_t *counters)
{
    const uint32_t g = nseg < 1 ? 1 : (nseg > 4096 ? 4096 : nseg);
    hipLaunchKernelGGL(k_sort_scan_seg, dim3(g), dim3(BLOCK), 0, s, hist, pt, nseg, passes, counters);
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

// Record sort tiles: 16384 points (8 waves x 32) for sorts of >= BIG_SORT
// points whose segments average >= 64K points, where each digit's run per tile
// is long enough (64 points) that the scattered writes mostly fill whole cache
// lines; 4096 (4 waves x 16) otherwise (every segment ends in a partial tile).
static inline bool big_records(uint32_t n, uint32_t nseg)
{
    return n >= BIG_SORT && n / (nseg ? nseg : 1u) >= 65536u;
}
uint32_t record_tile_keys(uint32_t n, uint32_t nseg) { return big_records(n, nseg) ? 8 * 64 * 32 : 4 * 64 * 16; }

template <int IT, int NW>
static void sort_records_shape(hipStream_t s, const uint32_t *keys_in, const RecIn &in,
                               uint32_t *keys_out, const RecOut &out, uint32_t n, int shift,
                               const uint32_t *digit_off, uint32_t *lookback, uint32_t *tile_counter,
                               bool first, bool last, bool k16, const SortSegs &seg, uint32_t grid)
{
    const dim3 b(NW * 64);
#define OWM_REC(IOTA, LAST, POS, K16)                                                              \
    hipLaunchKernelGGL((k_sort_records<IOTA, LAST, POS, K16, IT, NW>), dim3(grid), b, 0, s, keys_in, \
                       in, keys_out, out, n, shift, digit_off, lookback, tile_counter, seg)
    if (first && last && out.pos)
        OWM_REC(true, true, true, false);  // a single pass
    else if (first && last)
        OWM_REC(true, true, false, false);
    else if (first && k16)
        OWM_REC(true, false, false, true);
    else if (first)
        OWM_REC(true, false, false, false);
    else if (!last)
        OWM_REC(false, false, false, false);
    else if (out.pos)
        OWM_REC(false, true, true, false);
    else
        OWM_REC(false, true, false, false);
#undef OWM_REC
}

void launch_sort_records(hipStream_t s, const uint32_t *keys_in, const RecIn &in,
                         uint32_t *keys_out, const RecOut &out, uint32_t n, int shift,
                         int key_bits, const uint32_t *digit_off, uint32_t *lookback,
                         uint32_t *tile_counter, bool first, bool last, const SortSegs &seg,
                         uint32_t seg_tiles)
{
    const uint32_t grid = seg_tiles < 1 ? 1 : seg_tiles;
    const bool k16 = key_bits <= 16 && shift == 0;
    if (big_records(n, seg.nseg))
        sort_records_shape<32, 8>(s, keys_in, in, keys_out, out, n, shift, digit_off, lookback,
                                  tile_counter, first, last, k16, seg, grid);
    else
        sort_records_shape<16, 4>(s, keys_in, in, keys_out, out, n, shift, digit_off, lookback,
                                  tile_counter, first, last, k16, seg, grid);
}

}  // namespace owm
