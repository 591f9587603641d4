// OWM stages 1-3 on gfx950.
//
// OWM stages 2 and 3 on gfx950 (stage 1 is in stage1.hip).
//
// Stage 2 (sort + stage2, owm_optim3.cpp:158-188, optim3_func.cpp:610-625):
// the window minima mark a point bitmap (first and second window), and an
// ordered compaction of the second-window bits gives the LLPs ascending, as
// the reference's sorted run-length test does.  The seed merge is the same
// bitmap with the stage-3 additions marked, compacted again.
//
// Stage 3 (stage3, optim3_func.cpp:627-662): LLPs mark their Bsize cell; one
// wave per unmarked cell walks the LBVH for the lowest point strictly inside.
#include "owm_kernels.h"

namespace owm {

__device__ __forceinline__ void node_box(const TreeDev &T, uint32_t ref, double4 &box, uint32_t &cnt)
{
    if (ref & OWM_LEAF_BIT) {
        uint32_t l = ref & ~OWM_LEAF_BIT;
        box = T.lf_box[l];
        cnt = T.lf_cnt[l];
    } else {
        box = T.nd_box[ref];
        cnt = T.nd_cnt[ref];
    }
}

// ------------------------------------------------------------ compaction --

constexpr int CP_BLOCK = 256;
constexpr int CP_ITEMS = 16;
constexpr int CP_TILE = CP_BLOCK * CP_ITEMS;
constexpr uint32_t LB_AGG = 1u << 30, LB_INC = 2u << 30, LB_CNT = (1u << 30) - 1;

__device__ __forceinline__ uint32_t block_excl_scan_cp(uint32_t v, uint32_t *s_tmp, uint32_t &total)
{
    uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t u = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o)
            inc += u;
    }
    if (lane == 63)
        s_tmp[w] = inc;
    __syncthreads();
    uint32_t add = 0, tot = 0;
    for (uint32_t k = 0; k < CP_BLOCK / 64; ++k) {
        if (k < w)
            add += s_tmp[k];
        tot += s_tmp[k];
    }
    __syncthreads();
    total = tot;
    return add + inc - v;
}

// Decoupled look-back over compaction tiles taken in order (a tile's
// predecessors are all held by running workgroups): publish this tile's
// count, add the predecessors' counts back to the first inclusive prefix,
// publish the inclusive prefix, return the exclusive one.
__device__ __forceinline__ uint32_t tile_prefix(uint32_t *lookback, uint32_t tile, uint32_t total)
{
    uint32_t *flag = lookback + tile;
    uint32_t prefix = 0;
    if (tile == 0) {
        __hip_atomic_store(flag, LB_INC | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return 0;
    }
    __hip_atomic_store(flag, LB_AGG | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int32_t j = (int32_t)tile - 1;
    for (;;) {
        uint32_t f = __hip_atomic_load(lookback + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t st = f & ~LB_CNT;
        if (st == 0)
            continue;
        prefix += f & LB_CNT;
        if (st == LB_INC)
            break;
        --j;
    }
    __hip_atomic_store(flag, LB_INC | (prefix + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return prefix;
}

// Entries != OWM_NONE of in[0..n), in order, written from out[*d_out_base].
__global__ __launch_bounds__(CP_BLOCK) void k_compact_valid(const uint32_t *__restrict__ in,
                                                            uint32_t n, uint32_t *__restrict__ out,
                                                            const uint32_t *d_out_base,
                                                            uint32_t *d_count, uint32_t *lookback,
                                                            uint32_t *tile_counter)
{
    __shared__ uint32_t s_tile, s_tmp[CP_BLOCK / 64], s_prefix;
    const uint32_t ntiles = (n + CP_TILE - 1) / CP_TILE;
    const uint32_t out_base = d_out_base ? *d_out_base : 0u;
    if (n == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            *d_count = 0;
        return;
    }
    for (;;) {
        if (threadIdx.x == 0)
            s_tile = atomicAdd(tile_counter, 1u);
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= ntiles)
            break;
        const uint32_t i0 = tile * CP_TILE + threadIdx.x * CP_ITEMS;
        uint32_t vals[CP_ITEMS];
        uint32_t flags = 0, cnt = 0;
#pragma unroll
        for (int k = 0; k < CP_ITEMS; ++k) {
            const uint32_t i = i0 + k;
            const uint32_t v = i < n ? in[i] : OWM_NONE;
            const bool f = v != OWM_NONE;
            vals[k] = v;
            flags |= (f ? 1u : 0u) << k;
            cnt += f;
        }
        uint32_t total;
        uint32_t excl = block_excl_scan_cp(cnt, s_tmp, total);
        if (threadIdx.x == 0) {
            const uint32_t prefix = tile_prefix(lookback, tile, total);
            s_prefix = prefix;
            if (tile == ntiles - 1)
                *d_count = prefix + total;
        }
        __syncthreads();
        uint32_t pos = out_base + s_prefix + excl;
#pragma unroll
        for (int k = 0; k < CP_ITEMS; ++k)
            if ((flags >> k) & 1u)
                out[pos++] = vals[k];
        __syncthreads();
    }
}

// --------------------------------------------------- stage 2 by bitmaps --
//
// A point is an LLP when it is the minimum of two valid windows (of one, on
// tiles whose Overlap is 0: the reference keeps every valid minimum there).
// Windows are listed row by row, so neighbouring windows often share their
// minimum: only the first window of such a run acts, and a run of two or more
// marks the point in `llp` at once.  A window alone with its minimum marks it
// in `seen` and, when `seen` already held it (another window elsewhere),
// in `llp`.  Compacting `llp` in bit order then lists the LLPs ascending,
// which is what sorting the minima and keeping runs of two or more gives.
constexpr int S2_ITEMS = 4;

__global__ __launch_bounds__(256) void k_s2_mark(BatchDev B, const uint32_t *__restrict__ min_ids,
                                                 uint32_t nwin, uint32_t np, uint32_t any_need1,
                                                 uint32_t *__restrict__ seen,
                                                 uint32_t *__restrict__ llp)
{
    const uint32_t step = gridDim.x * 256u * S2_ITEMS;
    for (uint32_t base = blockIdx.x * 256u * S2_ITEMS + threadIdx.x; base < nwin; base += step) {
        uint32_t v[S2_ITEMS], act[S2_ITEMS], old[S2_ITEMS];
#pragma unroll
        for (int k = 0; k < S2_ITEMS; ++k) {
            const uint32_t i = base + k * 256u;
            v[k] = i < nwin ? min_ids[i] : OWM_NONE;
            const uint32_t prev = i > 0 && i < nwin ? min_ids[i - 1] : OWM_NONE;
            const uint32_t next = i + 1 < nwin ? min_ids[i + 1] : OWM_NONE;
            act[k] = 0;  // not a point, or a later window of a run
            if (v[k] < np && prev != v[k]) {
                uint32_t need = 2;
                if (any_need1)
                    need = B.tiles[owm_find_tile(B.pt_off, B.ntiles, v[k])].stage2_need;
                act[k] = need == 1 || next == v[k] ? 1u : 2u;
            }
        }
#pragma unroll
        for (int k = 0; k < S2_ITEMS; ++k)
            old[k] = act[k] == 2 ? atomicOr(seen + (v[k] >> 5), 1u << (v[k] & 31u)) : 0u;
#pragma unroll
        for (int k = 0; k < S2_ITEMS; ++k) {
            const uint32_t bit = 1u << (v[k] & 31u);
            if (act[k] == 1 || (act[k] == 2 && (old[k] & bit)))
                atomicOr(llp + (v[k] >> 5), bit);
        }
    }
}

// Points listed in ids[off .. off + n) (off, n read on the device; off may be
// null = 0) marked in a bitmap.
__global__ __launch_bounds__(256) void k_bits_set(const uint32_t *__restrict__ ids,
                                                  const uint32_t *d_off, const uint32_t *d_n,
                                                  uint32_t np, uint32_t *__restrict__ bits)
{
    const uint32_t off = d_off ? *d_off : 0u, n = *d_n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t v = ids[off + i];
        if (v < np)
            atomicOr(bits + (v >> 5), 1u << (v & 31u));
    }
}

// The set bits of words[0 .. nwords), in order, as point indices: ordered
// compaction with one look-back per tile of CP_BLOCK x 4 x BC_VEC words (BC_VEC
// uint4 loads per lane).  The look-back chain is serial, so a large bitmap
// takes large tiles (few links) and a small one small tiles (enough
// workgroups).
template <int BC_VEC>
__global__ __launch_bounds__(CP_BLOCK) void k_bits_compact(const uint32_t *__restrict__ words,
                                                           uint32_t nwords,
                                                           uint32_t *__restrict__ out,
                                                           uint32_t *d_count, uint32_t *lookback,
                                                           uint32_t *tile_counter)
{
    constexpr uint32_t BC_TILE = CP_BLOCK * BC_VEC * 4;
    __shared__ uint32_t s_tile, s_tmp[CP_BLOCK / 64], s_prefix;
    const uint32_t ntiles = (nwords + BC_TILE - 1) / BC_TILE;
    if (nwords == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            *d_count = 0;
        return;
    }
    for (;;) {
        if (threadIdx.x == 0)
            s_tile = atomicAdd(tile_counter, 1u);
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= ntiles)
            break;
        const uint32_t w0 = tile * BC_TILE + threadIdx.x * (BC_VEC * 4);
        uint4 wv[BC_VEC];
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < BC_VEC; ++k) {
            const uint32_t w = w0 + 4 * k;
            if (w + 4 <= nwords) {
                wv[k] = *reinterpret_cast<const uint4 *>(words + w);
            } else {
                wv[k].x = w < nwords ? words[w] : 0u;
                wv[k].y = w + 1 < nwords ? words[w + 1] : 0u;
                wv[k].z = w + 2 < nwords ? words[w + 2] : 0u;
                wv[k].w = 0u;
            }
            cnt += __popc(wv[k].x) + __popc(wv[k].y) + __popc(wv[k].z) + __popc(wv[k].w);
        }
        uint32_t total;
        const uint32_t excl = block_excl_scan_cp(cnt, s_tmp, total);
        if (threadIdx.x == 0) {
            const uint32_t prefix = tile_prefix(lookback, tile, total);
            s_prefix = prefix;
            if (tile == ntiles - 1)
                *d_count = prefix + total;
        }
        __syncthreads();
        uint32_t pos = s_prefix + excl;
        if (cnt) {
#pragma unroll
            for (int k = 0; k < BC_VEC; ++k) {
                const uint32_t m4[4] = {wv[k].x, wv[k].y, wv[k].z, wv[k].w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    uint32_t m = m4[q];
                    while (m) {
                        out[pos++] = ((w0 + 4 * k + q) << 5) | (uint32_t)__builtin_ctz(m);
                        m &= m - 1u;
                    }
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- stage 3 --

__global__ void k_s3_mark(BatchDev B, TreeDev T, const uint32_t *__restrict__ llp,
                          const uint32_t *d_nllp, uint32_t *__restrict__ cell_flag)
{
    const uint32_t nl = *d_nllp;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nl; i += gridDim.x * blockDim.x) {
        uint32_t id = llp[i];
        const TileDev &TD = B.tiles[owm_find_tile(B.pt_off, B.ntiles, id)];
        if (TD.bsize == 0)
            continue;
        double x, y;
        if (T.X) {
            x = T.X[id];
            y = T.Y[id];
        } else {
            const double2 xy = T.pts.xy[T.pos[id]];
            x = xy.x;
            y = xy.y;
        }
        if (isnan(x) || isnan(y))
            continue;
        const double hb = (double)(float)TD.hb;
        double fx = floor((x - (TD.initxg - hb)) / (double)TD.bsize);
        double fy = floor((y - (TD.inityg - hb)) / (double)TD.bsize);
        if (!(fx > -4.0 && fx < 70000.0 && fy > -4.0 && fy < 70000.0))
            continue;
        for (int dj = -1; dj <= 1; ++dj) {
            int jj = (int)fy + dj;
            if (jj < 0 || jj >= (int)TD.nrowsg)
                continue;
            double cy = owm_gcenter(TD.inityg, TD.bsize, (uint32_t)jj);
            if (!(y > __dsub_rn(cy, hb) && y < __dadd_rn(cy, hb)))
                continue;
            for (int di = -1; di <= 1; ++di) {
                int ii = (int)fx + di;
                if (ii < 0 || ii >= (int)TD.ncolsg)
                    continue;
                double cx = owm_gcenter(TD.initxg, TD.bsize, (uint32_t)ii);
                if (x > __dsub_rn(cx, hb) && x < __dadd_rn(cx, hb))
                    cell_flag[TD.cell_off + (uint32_t)jj * TD.ncolsg + (uint32_t)ii] = 1u;
            }
        }
    }
}

constexpr int S3_STACK = 1024;
constexpr uint32_t S3_PART = 256;  // partially covered leaves a wave defers per cell

// One wave per cell without an LLP: lowest (z, index) strictly inside.
__global__ __launch_bounds__(256) void k_s3_query(BatchDev B, TreeDev T,
                                                   const uint32_t *__restrict__ cell_flag,
                                                   uint32_t *__restrict__ cell_res, uint32_t *err)
{
    __shared__ uint32_t sStack[4][S3_STACK];
    __shared__ uint32_t sPart[4][S3_PART];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    uint32_t *stk = sStack[w];
    uint32_t *part = sPart[w];
    const uint32_t nw = gridDim.x * 4;
    for (uint32_t c = blockIdx.x * 4 + w; c < B.ncell; c += nw) {
        if (cell_flag[c]) {
            if (lane == 0)
                cell_res[c] = OWM_NONE;
            continue;
        }
        const TileDev &TD = B.tiles[owm_find_tile(B.cell_off, B.ntiles, c)];
        const uint32_t lc = c - TD.cell_off;
        const uint32_t ii = lc % TD.ncolsg, jj = lc / TD.ncolsg;
        const double hb = (double)(float)TD.hb;
        const double cx = owm_gcenter(TD.initxg, TD.bsize, ii);
        const double cy = owm_gcenter(TD.inityg, TD.bsize, jj);
        const double b0 = __dsub_rn(cx, hb), b1 = __dsub_rn(cy, hb);
        const double b2 = __dadd_rn(cx, hb), b3 = __dadd_rn(cy, hb);
        uint32_t cnt = 0, bid = OWM_NONE;
        unsigned long long bk = OWM_ZKEY_NONE;
        uint32_t nstack = (TD.n > 0 && TD.root != OWM_NONE) ? 1u : 0u;
        uint32_t npart = 0;  // leaves the cell's edge crosses, scanned by the whole wave below
        if (lane == 0)
            stk[0] = TD.root;
        __builtin_amdgcn_wave_barrier();
        while (nstack > 0) {
            uint32_t m = min(min(64u, nstack), S3_STACK - nstack);
            if (m == 0) {
                if (lane == 0)
                    atomicOr(err, 2u);
                break;
            }
            uint32_t node = lane < m ? stk[nstack - 1 - lane] : 0u;
            __builtin_amdgcn_wave_barrier();
            nstack -= m;
            bool push = false, crossed = false;
            if (lane < m) {
                double4 bx;
                uint32_t nc;
                node_box(T, node, bx, nc);
                bool dis = nc == 0 || bx.z <= b0 || bx.x >= b2 || bx.w <= b1 || bx.y >= b3;
                bool ins = bx.x > b0 && bx.z < b2 && bx.y > b1 && bx.w < b3;
                if (!dis) {
                    if (ins) {
                        unsigned long long k;
                        uint32_t id;
                        if (node & OWM_LEAF_BIT) {
                            k = T.lf_zkey[node & ~OWM_LEAF_BIT];
                            id = T.lf_zid[node & ~OWM_LEAF_BIT];
                        } else {
                            k = T.nd_zkey[node];
                            id = T.nd_zid[node];
                        }
                        cnt += nc;
                        if (owm_less(k, id, bk, bid)) {
                            bk = k;
                            bid = id;
                        }
                    } else if (node & OWM_LEAF_BIT) {
                        crossed = true;
                    } else {
                        push = true;
                    }
                }
            }
            // crossed leaves go to the wave's list; past its capacity a lane
            // scans its leaf itself
            const unsigned long long cm = __ballot(crossed);
            if (crossed) {
                const uint32_t at = npart + (uint32_t)__popcll(cm & lt_mask);
                if (at < S3_PART) {
                    part[at] = node & ~OWM_LEAF_BIT;
                } else {
                    const uint32_t l = node & ~OWM_LEAF_BIT;
                    const uint32_t s0 = T.lf_start[l], np = T.lf_npts[l];
                    for (uint32_t p = s0; p < s0 + np; ++p) {
                        const PointRec r = owm_load_pt(T.pts, p);
                        const double x = r.x, y = r.y;
                        if (x > b0 && x < b2 && y > b1 && y < b3) {
                            ++cnt;
                            const double z = r.z;
                            if (owm_eligible(z)) {
                                const unsigned long long k = owm_zkey(z);
                                const uint32_t id = r.id;
                                if (owm_less(k, id, bk, bid)) {
                                    bk = k;
                                    bid = id;
                                }
                            }
                        }
                    }
                }
            }
            npart = min(npart + (uint32_t)__popcll(cm), S3_PART);
            unsigned long long pm = __ballot(push);
            if (push) {
                uint2 ch = T.nd_child[node];
                uint32_t base = nstack + 2 * __popcll(pm & lt_mask);
                stk[base] = ch.y;
                stk[base + 1] = ch.x;
            }
            nstack += 2 * __popcll(pm);
            __builtin_amdgcn_wave_barrier();
        }
        // crossed leaves: one point per lane, four leaves' loads in flight
        for (uint32_t i = 0; i < npart; i += 4) {
            uint32_t s0[4], np[4], mx = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t l = i + q < npart ? part[i + q] : OWM_NONE;
                s0[q] = l != OWM_NONE ? T.lf_start[l] : 0u;
                np[q] = l != OWM_NONE ? T.lf_npts[l] : 0u;
                mx = max(mx, np[q]);
            }
            for (uint32_t o = 0; o < mx; o += 64) {
                double x[4], y[4], z[4];
                uint32_t id[4];
                bool in[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    in[q] = o + lane < np[q];
                    const PointRec r = owm_load_pt(T.pts, s0[q] + (in[q] ? o + lane : 0u));
                    x[q] = r.x;
                    y[q] = r.y;
                    z[q] = r.z;
                    id[q] = r.id;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (in[q] && x[q] > b0 && x[q] < b2 && y[q] > b1 && y[q] < b3) {
                        ++cnt;
                        if (owm_eligible(z[q])) {
                            const unsigned long long k = owm_zkey(z[q]);
                            if (owm_less(k, id[q], bk, bid)) {
                                bk = k;
                                bid = id[q];
                            }
                        }
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            cnt += __shfl_xor(cnt, o, 64);
            unsigned long long k2 = __shfl_xor(bk, o, 64);
            uint32_t i2 = __shfl_xor(bid, o, 64);
            if (owm_less(k2, i2, bk, bid)) {
                bk = k2;
                bid = i2;
            }
        }
        if (lane == 0)
            cell_res[c] = (cnt > 0 && bk != OWM_ZKEY_NONE) ? bid : OWM_NONE;
    }
}

__global__ void k_seed_ranges(BatchDev B, const uint32_t *__restrict__ seeds, const uint32_t *d_ns,
                              uint32_t *__restrict__ seed_off)
{
    const uint32_t ns = *d_ns;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t <= B.ntiles; t += gridDim.x * blockDim.x) {
        if (t == B.ntiles) {
            seed_off[t] = ns;
            continue;
        }
        uint32_t v = B.pt_off[t], lo = 0, hi = ns;
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (seeds[mid] < v)
                lo = mid + 1;
            else
                hi = mid;
        }
        seed_off[t] = lo;
    }
}

// Batch indices of the merged seeds -> indices within their tiles (after the
// ranges: the seeds are sorted by batch index, so the tile of seed i is the one
// whose point range holds it).
__global__ void k_seeds_local(BatchDev B, uint32_t *__restrict__ seeds, const uint32_t *d_ns)
{
    const uint32_t ns = *d_ns;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x) {
        const uint32_t v = seeds[i];
        seeds[i] = v - B.pt_off[owm_find_tile(B.pt_off, B.ntiles, v)];
    }
}

__global__ void k_add_u32(const uint32_t *a, const uint32_t *b, uint32_t *out) { *out = *a + *b; }

// ----------------------------------------------------------------- launch --

static uint32_t cp_grid(uint32_t n)
{
    uint32_t t = (n + CP_TILE - 1) / CP_TILE;
    return t < 1 ? 1 : (t > 4096 ? 4096 : t);
}

// seen and llp (np/32 words each) start at zero.
void launch_stage2_mark(hipStream_t s, const BatchDev &B, const uint32_t *min_ids, uint32_t nwin,
                        uint32_t np, uint32_t any_need1, uint32_t *seen, uint32_t *llp)
{
    uint32_t g = (nwin + 256 * S2_ITEMS - 1) / (256 * S2_ITEMS);
    g = g < 1 ? 1 : (g > 8192 ? 8192 : g);
    hipLaunchKernelGGL(k_s2_mark, dim3(g), dim3(256), 0, s, B, min_ids, nwin, np, any_need1, seen,
                       llp);
}

void launch_bits_set(hipStream_t s, const uint32_t *ids, const uint32_t *d_off, const uint32_t *d_n,
                     uint32_t n_max, uint32_t np, uint32_t *bits)
{
    uint32_t g = (n_max + 255) / 256;
    g = g < 1 ? 1 : (g > 4096 ? 4096 : g);
    hipLaunchKernelGGL(k_bits_set, dim3(g), dim3(256), 0, s, ids, d_off, d_n, np, bits);
}

// lookback (nwords / 4096 + 2 words) and tile_counter start at zero; words
// 16-byte aligned.  Tiles of 4096 words up to 2M words, then 8192, then 16384.
void launch_bits_compact(hipStream_t s, const uint32_t *words, uint32_t nwords, uint32_t *out,
                         uint32_t *d_count, uint32_t *lookback, uint32_t *tile_counter)
{
    const uint32_t tile = nwords <= (2u << 20) ? 4096u : nwords <= (8u << 20) ? 8192u : 16384u;
    uint32_t g = (nwords + tile - 1) / tile;
    g = g < 1 ? 1 : (g > 4096 ? 4096 : g);
    if (tile == 4096u)
        hipLaunchKernelGGL(k_bits_compact<4>, dim3(g), dim3(CP_BLOCK), 0, s, words, nwords, out,
                           d_count, lookback, tile_counter);
    else if (tile == 8192u)
        hipLaunchKernelGGL(k_bits_compact<8>, dim3(g), dim3(CP_BLOCK), 0, s, words, nwords, out,
                           d_count, lookback, tile_counter);
    else
        hipLaunchKernelGGL(k_bits_compact<16>, dim3(g), dim3(CP_BLOCK), 0, s, words, nwords, out,
                           d_count, lookback, tile_counter);
}

void launch_compact_valid(hipStream_t s, const uint32_t *in, uint32_t n, uint32_t *out,
                          const uint32_t *d_out_base, uint32_t *d_count, uint32_t *lookback,
                          uint32_t *tile_counter)
{
    // lookback ((n + CP_TILE - 1) / CP_TILE + 1 words) and tile_counter start at zero
    hipLaunchKernelGGL(k_compact_valid, dim3(cp_grid(n)), dim3(CP_BLOCK), 0, s, in, n, out,
                       d_out_base, d_count, lookback, tile_counter);
}

void launch_stage3(hipStream_t s, const BatchDev &B, const TreeDev &T, const uint32_t *llp,
                   const uint32_t *n_llp, uint32_t n_llp_max, uint32_t *cell_flag,
                   uint32_t *cell_res, uint32_t *err)
{
    if (B.ncell == 0)
        return;
    // cell_flag (B.ncell words) starts at zero
    uint32_t g = (n_llp_max + 255) / 256;
    g = g < 1 ? 1 : (g > 4096 ? 4096 : g);
    hipLaunchKernelGGL(k_s3_mark, dim3(g), dim3(256), 0, s, B, T, llp, n_llp, cell_flag);
    uint32_t gq = (B.ncell + 3) / 4;
    gq = gq < 1 ? 1 : (gq > 16384 ? 16384 : gq);
    hipLaunchKernelGGL(k_s3_query, dim3(gq), dim3(256), 0, s, B, T, cell_flag, cell_res, err);
}

void launch_seed_ranges(hipStream_t s, const BatchDev &B, const uint32_t *seeds,
                        const uint32_t *d_nseeds, uint32_t *seed_off)
{
    uint32_t g = (B.ntiles + 1 + 255) / 256;
    hipLaunchKernelGGL(k_seed_ranges, dim3(g), dim3(256), 0, s, B, seeds, d_nseeds, seed_off);
}

void launch_seeds_local(hipStream_t s, const BatchDev &B, uint32_t *seeds, const uint32_t *d_nseeds,
                        uint32_t n_max)
{
    if (B.ntiles < 2)
        return;  // one tile: batch and tile indices coincide
    uint32_t g = (n_max + 255) / 256;
    g = g < 1 ? 1 : (g > 4096 ? 4096 : g);
    hipLaunchKernelGGL(k_seeds_local, dim3(g), dim3(256), 0, s, B, seeds, d_nseeds);
}

void launch_add_u32(hipStream_t s, const uint32_t *a, const uint32_t *b, uint32_t *out)
{
    hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(1), 0, s, a, b, out);
}

}  // namespace owm
