// LSD radix sort of u32 keys (optionally with u32 values) for gfx950.
//
// One upfront histogram pass counts the digits of every pass; each pass is
// then a single "onesweep" kernel: a workgroup takes the next 4096-key tile
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
static inline int sort_items(uint32_t n) { return n >= BIG_SORT ? 32 : 16; }
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

// Exclusive scan of one value per thread over a 256-thread block.
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
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_whist[WAVES][RADIX];
    __shared__ uint32_t s_excl[RADIX];
    __shared__ uint32_t s_glob[RADIX];
    __shared__ uint32_t s_tmp[WAVES];
    __shared__ uint32_t s_keys[TILE_T];
    __shared__ uint32_t s_vals[KEYS_ONLY ? 1 : TILE_T];

    const uint32_t n_all = d_n ? *d_n : n_host;
    const uint32_t ntiles = seg.pt ? seg.units[seg.nseg] : (n_all + TILE_T - 1) / TILE_T;
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;

    for (;;) {
        if (tid == 0)
            s_tile = atomicAdd(tile_counter, 1u);
        for (int i = tid; i < WAVES * RADIX; i += BLOCK)
            (&s_whist[0][0])[i] = 0;
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= ntiles)
            break;
        // the tile's segment: [base, n) of the input, digit offsets doff
        uint32_t base, n;
        bool first;
        const uint32_t *doff = digit_off;
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

        uint32_t key[ITEMS_T], val[ITEMS_T], rank[ITEMS_T];
#pragma unroll
        for (int k = 0; k < ITEMS_T; ++k) {
            uint32_t idx = base + w * (64 * ITEMS_T) + k * 64 + lane;
            bool valid = idx < n;
            key[k] = valid ? keys_in[idx] : 0u;
            if (!KEYS_ONLY)
                val[k] = valid ? (IOTA ? idx : vals_in[idx]) : 0u;
        }
#pragma unroll
        for (int k = 0; k < ITEMS_T; ++k) {
            uint32_t idx = base + w * (64 * ITEMS_T) + k * 64 + lane;
            bool valid = idx < n;
            uint32_t d = (key[k] >> shift) & (RADIX - 1);
            unsigned long long m = __ballot(valid);
#pragma unroll
            for (int b = 0; b < RB; ++b) {
                bool bit = (d >> b) & 1u;
                unsigned long long bb = __ballot(bit);
                m &= bit ? bb : ~bb;
            }
            uint32_t before = __popcll(m & lt_mask);
            uint32_t old = s_whist[w][d];
            if (valid && before == 0)
                s_whist[w][d] = old + (uint32_t)__popcll(m);
            rank[k] = old + before;
        }
        __syncthreads();

        // per digit: wave-exclusive prefixes, tile total, look-back, global base
        const uint32_t d = tid;
        uint32_t c0 = s_whist[0][d], c1 = s_whist[1][d], c2 = s_whist[2][d], c3 = s_whist[3][d];
        uint32_t total = c0 + c1 + c2 + c3;
        s_whist[0][d] = 0;
        s_whist[1][d] = c0;
        s_whist[2][d] = c0 + c1;
        s_whist[3][d] = c0 + c1 + c2;
        uint32_t *flag = lookback + (size_t)tile * RADIX + d;
        uint32_t prefix = 0;
        if (first) {
            __hip_atomic_store(flag, LB_INC | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(flag, LB_AGG | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int32_t j = (int32_t)tile - 1;
            for (;;) {
                uint32_t f = __hip_atomic_load(lookback + (size_t)j * RADIX + d, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                uint32_t st = f & ~LB_CNT;
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
        s_glob[d] = doff[d] + prefix;
        uint32_t ex = block_excl_scan(total, s_tmp);
        s_excl[d] = ex;
        __syncthreads();

#pragma unroll
        for (int k = 0; k < ITEMS_T; ++k) {
            uint32_t idx = base + w * (64 * ITEMS_T) + k * 64 + lane;
            if (idx < n) {
                uint32_t dd = (key[k] >> shift) & (RADIX - 1);
                uint32_t slot = s_excl[dd] + s_whist[w][dd] + rank[k];
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
            uint32_t dst = s_glob[dd] + (s - s_excl[dd]);
            if (write_keys)
                keys_out[dst] = k;
            if (!KEYS_ONLY)
                vals_out[dst] = s_vals[s];
        }
        __syncthreads();
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

uint32_t sort_tile_keys(uint32_t n) { return (uint32_t)(BLOCK * sort_items(n)); }

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
    if (sort_items(n_max) == 32)
        onesweep_items<32>(s, keys_in, vals_in, keys_out, vals_out, d_n, n_max, shift, digit_off,
                           lookback, tile_counter, keys_only, iota, wk, grid, seg);
    else
        onesweep_items<16>(s, keys_in, vals_in, keys_out, vals_out, d_n, n_max, shift, digit_off,
                           lookback, tile_counter, keys_only, iota, wk, grid, seg);
}

}  // namespace owm
