// Host orchestration of the OWM pipeline behind the C ABI of include/owm.h.
//
// A run is a batch of tiles (a single tile is a batch of one).  The host
// derives each tile's parameters exactly as the reference driver does
// (CPP-TBB-CPU/application/owm_optim3.cpp:92-132), lays every per-tile range
// out in one concatenated index space, uploads the tile table, and enqueues
// the device chain: Morton keys -> record radix sort (the cloud itself ends in
// Morton order) -> leaves -> Karras -> bottom-up -> stage 1 -> sort + stage 2
// -> stage 3 -> merge + sort of the seeds.  Only the bbox (when the caller has none) and the final seed count
// come back to the host.
#include "owm_host.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

namespace owm {

// ------------------------------------------------------------ utilities --

void set_err(owm_ctx *ctx, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx)
        ctx->err = buf;
}

owm_status hip_fail(owm_ctx *ctx, hipError_t e, const char *what)
{
    set_err(ctx, "%s: %s", what, hipGetErrorString(e));
    if (e == hipErrorOutOfMemory)
        return OWM_E_NOMEM;
    return OWM_E_DEVICE;
}

static inline uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

// Buffers a chain needs zeroed, cleared together by one launch right after
// they are carved (every one of them is zero before its first use, and none is
// written in between).
struct ZeroList {
    FillRanges r;
    void add(hipStream_t s, void *p, uint64_t words)
    {
        if (!p || !words)
            return;
        if (r.count == FILL_MAX)
            launch(s);
        r.ptr[r.count] = (uint32_t *)p;
        r.len[r.count++] = words;
    }
    void launch(hipStream_t s)
    {
        launch_fill_ranges(s, r);
        r.count = 0;
    }
};

// Regions of the pinned readback buffer for n_tiles tiles.
struct IoLayout {
    uint64_t ptoff, bbox, frames, counts, seed_off, bytes;
    explicit IoLayout(uint32_t n_tiles)
    {
        const uint64_t off = align_up((n_tiles + 1) * 4ull, 256);
        ptoff = 0;
        bbox = off;
        frames = bbox + align_up(n_tiles * 32ull, 256);
        counts = frames + align_up(n_tiles * (uint64_t)sizeof(KeyFrame), 256);
        seed_off = counts + 256;
        bytes = seed_off + off;
    }
};

static const char *PHASE_NAMES[PH_COUNT] = {
    "bbox", "morton", "sort_tree_pass", "leaf_aggs", "karras", "agg_levels", "stage1",
    "stage2_mark", "stage2_compact", "stage3", "merge_seeds", "stage1_join",
    "stage1_summaries", "xyz_parse", "qtree_stats"};

const char *phase_name(int ph) { return PHASE_NAMES[ph]; }

int Profiler::begin(hipStream_t s, int ph)
{
    if (!on || !((mask >> ph) & 1u))
        return -1;
    hipEvent_t ev[2];
    for (auto &e : ev) {
        if (pool.empty()) {
            if (hipEventCreate(&e) != hipSuccess)
                return -1;
        } else {
            e = pool.back();
            pool.pop_back();
        }
    }
    (void)hipEventRecord(ev[0], s);
    pending.push_back({ph, ev[0], ev[1], 0, 0, 1});
    return (int)pending.size() - 1;
}

void Profiler::end(hipStream_t s, int idx, uint64_t b, uint64_t it, uint32_t launches)
{
    if (idx < 0 || idx >= (int)pending.size())
        return;
    (void)hipEventRecord(pending[idx].b, s);
    pending[idx].bytes = b;
    pending[idx].items = it;
    pending[idx].launches = launches;
}

void Profiler::collect()
{
    for (auto &r : pending) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
            ms[r.ph] += t;
            calls[r.ph] += r.launches;
            bytes[r.ph] += r.bytes;
            items[r.ph] += r.items;
        }
        pool.push_back(r.a);
        pool.push_back(r.b);
    }
    pending.clear();
}

void Profiler::add_work(int ph, uint64_t b, uint64_t it)
{
    if (!on || !((mask >> ph) & 1u))
        return;
    bytes[ph] += b;
    items[ph] += it;
}

void Profiler::reset()
{
    for (int k = 0; k < PH_COUNT; ++k) {
        ms[k] = 0;
        calls[k] = bytes[k] = items[k] = 0;
    }
}

Profiler::~Profiler()
{
    for (auto &r : pending) {
        pool.push_back(r.a);
        pool.push_back(r.b);
    }
    for (auto e : pool)
        (void)hipEventDestroy(e);
}

DevBuf::~DevBuf() { release(); }

void DevBuf::release()
{
    if (ptr && !borrowed)
        (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    borrowed = false;
}

hipError_t DevBuf::reserve(uint64_t need, uint64_t cap)
{
    if (need <= bytes)
        return hipSuccess;
    if (borrowed)
        return hipErrorOutOfMemory;  // a borrowed region never grows
    release();
    uint64_t want = std::max(need, std::min(need + need / 4, cap));
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        e = hipMalloc(&ptr, need);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        want = need;
    }
    bytes = want;
    return hipSuccess;
}

// --------------------------------------------------------- parameters ----

static double round2d(double z) { return std::round(z * 100.0) / 100.0; }

// Parameter derivation of the reference driver (owm_optim3.cpp:92-132), with
// the reference's integer types; values it would wrap or divide by zero on are
// rejected with OWM_E_ARG.
owm_status derive_params(const owm_params &P, uint64_t n, const double *bbox, Derived &D,
                         std::string &why)
{
    if (P.wsize == 0 || P.wsize > 65535u) {
        why = "wsize must be in [1, 65535]";
        return OWM_E_ARG;
    }
    if (P.bsize > 65535u) {
        why = "bsize must be <= 65535";
        return OWM_E_ARG;
    }
    if (!(P.overlap >= 0.0 && P.overlap < 1.0)) {
        why = "overlap must be in [0, 1)";
        return OWM_E_ARG;
    }
    for (int k = 0; k < 4; ++k)
        D.bbox[k] = bbox[k];
    const double minx = bbox[0], maxx = bbox[1], miny = bbox[2], maxy = bbox[3];
    const double width = round2d(maxx - minx), height = round2d(maxy - miny);
    if (!(width > 0.0) || !(height > 0.0) || !std::isfinite(width) || !std::isfinite(height)) {
        why = "bbox width and height (rounded to cm) must be positive and finite";
        return OWM_E_ARG;
    }
    D.width = width;
    D.height = height;
    D.density = (double)n / (width * height);
    const uint32_t Wsize = P.wsize;
    const double mnp = 0.5 * D.density * Wsize * Wsize;
    if (!(mnp < 4294967296.0)) {
        why = "0.5*density*wsize^2 overflows the reference's u32";
        return OWM_E_ARG;
    }
    D.min_num_points = (uint16_t)(uint32_t)mnp;
    D.displace = round2d(Wsize * (1 - P.overlap));
    if (!(D.displace > 0.0)) {
        why = "window displacement round2d(wsize*(1-overlap)) is 0";
        return OWM_E_ARG;
    }
    double c, r;
    if (P.overlap > 0.0) {
        c = std::round((width + 2 * Wsize * P.overlap) / D.displace) - 1;
        r = std::round((height + 2 * Wsize * P.overlap) / D.displace) - 1;
    } else {
        c = std::floor(width / Wsize) + 1;
        r = std::floor(height / Wsize) + 1;
    }
    if (!(c >= 1.0 && c <= 65535.0 && r >= 1.0 && r <= 65535.0)) {
        why = "window grid must be 1..65535 columns and rows";
        return OWM_E_ARG;
    }
    D.ncols = (uint32_t)c;
    D.nrows = (uint32_t)r;
    if (P.bsize > 0) {
        const uint32_t Bsize = P.bsize;
        double cg = std::floor(width / Bsize) + 1, rg = std::floor(height / Bsize) + 1;
        if (!(cg >= 1.0 && cg <= 65535.0 && rg >= 1.0 && rg <= 65535.0)) {
            why = "stage-3 grid must be 1..65535 columns and rows";
            return OWM_E_ARG;
        }
        D.ncolsg = (uint32_t)cg;
        D.nrowsg = (uint32_t)rg;
    } else {
        D.ncolsg = D.nrowsg = 0;
    }
    D.h = (int32_t)(P.wsize / 2);
    D.hb = (int32_t)(P.bsize / 2);
    D.initx = minx - D.h + D.displace;
    D.inity = miny - D.h + D.displace;
    D.initxg = minx + D.hb;
    D.inityg = miny + D.hb;
    return OWM_OK;
}

void fill_query(TileDev &T, const owm_params &P, const Derived &D)
{
    T.ncols = D.ncols;
    T.nrows = D.nrows;
    T.ncolsg = D.ncolsg;
    T.nrowsg = D.nrowsg;
    T.min_num_points = D.min_num_points;
    T.stage2_need = P.overlap != 0.0 ? 2u : 1u;
    T.h = D.h;
    T.hb = D.hb;
    T.bsize = P.bsize;
    T.displace = D.displace;
    T.inv_disp = 1.0 / D.displace;
    T.initx = D.initx;
    T.inity = D.inity;
    T.initxg = D.initxg;
    T.inityg = D.inityg;
    T.gcls = 0;
    T.pairs = 0;
    // the constants of the global window classes (the device's edge formula:
    // centre init + i*displace, then -+ the half side, each step rounded)
    const double hd = (double)(float)D.h;
    auto edge = [&](double init, uint32_t i, bool right) {
        const double c = init + (double)(int)i * D.displace;
        return right ? c + hd : c - hd;
    };
    const double init[2] = {D.initx, D.inity};
    const uint32_t n[2] = {D.ncols, D.nrows};
    double m = 0.0;
    for (int a = 0; a < 2; ++a) {
        T.ge0[a] = edge(init[a], 0, false);
        T.gc[a] = (edge(init[a], 0, true) - T.ge0[a]) * T.inv_disp;
        const double last = n[a] ? edge(init[a], n[a] - 1, true) : 0.0;
        m = std::max(m, std::max(std::fabs(init[a]), std::max(std::fabs(T.ge0[a]), std::fabs(last))) + hd);
    }
    T.gtol = std::ldexp(m, -46) * T.inv_disp;
}

// For a tree built in the same chain as its query: the last sort pass stamps
// every point's global window classes wherever they fit 16 bits per axis.
static uint64_t join_estimate(const Prepared &P);
static void join_floor(owm_ctx *ctx, Prepared &P);

// A tree built in the same chain as its query stamps every point's window
// classes, or, when every tile's window side is m <= S1_PAIRS_MAX window
// steps, its window columns: stage 1 then runs on 32x32-window blocks of
// column cells (16x16 when there are too few large blocks to fill the device;
// k_stage1_pairs), re-planned here.  OWM_NO_PAIRS=1 keeps the
// classes (a switch for comparisons).
static void stamp_classes(owm_ctx *ctx, Prepared &P)
{
    P.L.stamp = 0;
    bool pairs = !getenv("OWM_NO_PAIRS");
    bool any = false;
    for (TileDev &T : P.td) {
        T.gcls = T.ncols != 0 && T.nrows != 0 && T.ncols <= OWM_GCLS_MAX_WINDOWS &&
                 T.nrows <= OWM_GCLS_MAX_WINDOWS;
        if (T.gcls)
            P.L.stamp = 1;
        if (T.ncols == 0 || T.nrows == 0)
            continue;  // no windows, no blocks
        any = true;
        const double m = std::round(T.gc[0]);
        pairs = pairs && T.gcls && m >= 1.0 && m <= (double)S1_PAIRS_MAX &&
                std::fabs(T.gc[0] - m) < 1e-6 && std::fabs(T.gc[1] - m) < 1e-6;
    }
    if (!pairs || !any)
        return;
    for (TileDev &T : P.td) {
        T.pairs = T.gcls ? (uint32_t)std::round(T.gc[0]) : 0u;
        T.lcs = 0;
    }
    P.L.stamp = 2;
    // 32x32-window blocks where there are enough of them to occupy every
    // workgroup slot (two 512-thread workgroups per CU), else 16x16
    P.cfg = Stage1Cfg{32, 32, 4, 0, 1};
    P.Q.lcs_leaves = 0;
    plan_query(P.td, P.cfg, P.Q);
    if (P.Q.nblk < S1_PAIRS_BIG_BLOCKS) {
        P.cfg = Stage1Cfg{16, 16, 4, 0, 1};
        plan_query(P.td, P.cfg, P.Q);
    }
    P.Q.s1_cap = join_estimate(P);
    join_floor(ctx, P);
}

int bit_length(uint64_t v)
{
    int b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b;
}

// ------------------------------------------------------------- planning --

// Per-tile tree layout: leaves, internal nodes, Morton quantisation.
void plan_tree(int sort_bits, std::vector<TileDev> &tiles, const std::vector<uint64_t> &n,
               const std::vector<uint32_t> &leaf, const std::vector<std::array<double, 4>> &qbox,
               TreeLayout &L, const std::vector<KeyFrame> *frames)
{
    L.counts = nullptr;
    const uint32_t T = (uint32_t)tiles.size();
    L.ntiles = T;
    // The tree sort is segmented by tile, so the keys hold Morton bits only:
    // sort_bits of them (16, 24 or 32: two to four 8-bit passes).
    L.tile_bits = 0;
    L.stamp = 0;  // set by the callers that plan the query in the same chain
    L.key_bits = std::max(std::min(sort_bits, 32) & ~1, 4);
    L.pt_off.assign(T + 1, 0);
    L.leaf_off.assign(T + 1, 0);
    L.node_off.assign(T + 1, 0);
    L.grp_off.assign(T + 1, 0);
    for (uint32_t t = 0; t < T; ++t) {
        TileDev &D = tiles[t];
        D.pt_off = L.pt_off[t];
        D.n = (uint32_t)n[t];
        D.leaf_size = leaf[t];
        D.nleaf = (uint32_t)((n[t] + leaf[t] - 1) / leaf[t]);
        D.leaf_off = L.leaf_off[t];
        D.node_off = L.node_off[t];
        D.root = D.nleaf == 0 ? OWM_NONE : (D.nleaf == 1 ? (OWM_LEAF_BIT | D.leaf_off) : D.node_off);
        if (frames) {
            D.qminx = (*frames)[t].qminx;
            D.qminy = (*frames)[t].qminy;
            D.qscale = (*frames)[t].qscale;
        } else {
            const auto &b = qbox[t];
            double side = std::max(b[1] - b[0], b[3] - b[2]);
            if (!(side > 0.0) || !std::isfinite(side))
                side = 1.0;
            D.qminx = b[0];
            D.qminy = b[2];
            D.qscale = 4294967296.0 / (side * (1.0 + 1e-12));
        }
        L.pt_off[t + 1] = L.pt_off[t] + D.n;
        L.leaf_off[t + 1] = L.leaf_off[t] + D.nleaf;
        L.node_off[t + 1] = L.node_off[t] + (D.nleaf > 0 ? D.nleaf - 1 : 0);
        const uint32_t per = leaf[t] <= 64 ? 64 / leaf[t] : 1;
        D.grp_off = L.grp_off[t];
        L.grp_off[t + 1] = L.grp_off[t] + (D.nleaf + per - 1) / per;
    }
    L.ngrp = L.grp_off[T];
    L.np = L.pt_off[T];
    // sort tiles of each tile's segment (a sort tile never straddles tiles)
    const uint64_t tk = record_tile_keys((uint32_t)std::min<uint64_t>(L.np, 0xFFFFFFFFu), T);
    L.sort_units.assign(T + 1, 0);
    for (uint32_t t = 0; t < T; ++t)
        L.sort_units[t + 1] = L.sort_units[t] + (uint32_t)((n[t] + tk - 1) / tk);
    L.nleaf = L.leaf_off[T];
    L.nnode = L.node_off[T];
}

void plan_query(std::vector<TileDev> &tiles, const Stage1Cfg &cfg, QueryLayout &Q)
{
    const uint32_t T = (uint32_t)tiles.size();
    Q.win_off.assign(T + 1, 0);
    Q.cell_off.assign(T + 1, 0);
    Q.blk_off.assign(T + 1, 0);
    for (uint32_t t = 0; t < T; ++t) {
        TileDev &D = tiles[t];
        D.win_off = Q.win_off[t];
        D.cell_off = Q.cell_off[t];
        D.blk_off = Q.blk_off[t];
        D.nbx = (D.ncols + cfg.bx - 1) / cfg.bx;
        uint32_t nby = (D.nrows + cfg.by - 1) / cfg.by;
        Q.win_off[t + 1] = Q.win_off[t] + D.ncols * D.nrows;
        Q.cell_off[t + 1] = Q.cell_off[t] + D.ncolsg * D.nrowsg;
        Q.blk_off[t + 1] = Q.blk_off[t] + D.nbx * nby;
    }
    Q.nwin = Q.win_off[T];
    Q.ncell = Q.cell_off[T];
    Q.nblk = Q.blk_off[T];
}

// ------------------------------------------------------------- the tree --

uint64_t tree_bytes(const TreeLayout &L)
{
    uint64_t b = 0;
    b += align_up(L.np * 16ull, 256) * 2;                 // pts: xy, (z, id, key)
    b += align_up(L.np * 4ull, 256);                      // pos
    b += align_up(L.nleaf * 32ull, 256);                  // lf_box
    b += align_up(L.nleaf * 4ull, 256) * 5;               // cnt start npts zid key0
    b += align_up(L.nleaf * 8ull, 256);                   // lf_zkey
    b += align_up(std::max<uint64_t>(L.nnode, 1) * 32ull, 256);  // nd_box
    b += align_up(std::max<uint64_t>(L.nnode, 1) * 4ull, 256) * 2;  // cnt zid
    b += align_up(std::max<uint64_t>(L.nnode, 1) * 8ull, 256) * 2;  // zkey child
    return b;
}

void carve_tree(uint8_t *base, const TreeLayout &L, TreeOut &T)
{
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) {
        uint8_t *p = base + o;
        o += align_up(bytes, 256);
        return p;
    };
    const uint64_t nn = std::max<uint64_t>(L.nnode, 1);
    T.pts.xy = (double2 *)take(L.np * 16ull);
    T.pts.zi = (PointZ *)take(L.np * 16ull);
    T.pos = (uint32_t *)take(L.np * 4ull);
    T.lf_box = (double4 *)take(L.nleaf * 32ull);
    T.lf_cnt = (uint32_t *)take(L.nleaf * 4ull);
    T.lf_start = (uint32_t *)take(L.nleaf * 4ull);
    T.lf_npts = (uint32_t *)take(L.nleaf * 4ull);
    T.lf_zid = (uint32_t *)take(L.nleaf * 4ull);
    T.lf_key0 = (uint32_t *)take(L.nleaf * 4ull);
    T.lf_zkey = (unsigned long long *)take(L.nleaf * 8ull);
    T.nd_box = (double4 *)take(nn * 32ull);
    T.nd_cnt = (uint32_t *)take(nn * 4ull);
    T.nd_zid = (uint32_t *)take(nn * 4ull);
    T.nd_zkey = (unsigned long long *)take(nn * 8ull);
    T.nd_child = (uint2 *)take(nn * 8ull);
}

TreeDev tree_view(const TreeOut &T)
{
    TreeDev V;
    V.pts = PointsDev{T.pts.xy, T.pts.zi};
    V.pos = T.pos;
    V.lf_box = T.lf_box;
    V.lf_cnt = T.lf_cnt;
    V.lf_start = T.lf_start;
    V.lf_npts = T.lf_npts;
    V.lf_zid = T.lf_zid;
    V.lf_zkey = T.lf_zkey;
    V.nd_box = T.nd_box;
    V.nd_cnt = T.nd_cnt;
    V.nd_zid = T.nd_zid;
    V.nd_zkey = T.nd_zkey;
    V.nd_child = T.nd_child;
    V.X = V.Y = nullptr;
    return V;
}

// Scratch of the tree build (keys, values, histograms, look-back flags).
// nseg segments: one histogram per segment, and up to one partial sort tile more
// per segment.
// lb_tiles: look-back rows per pass (0: the key/value sort's tile count, plus
// one partial tile per segment)
uint64_t sort_scratch_bytes(uint64_t n, int passes, uint32_t nseg, uint64_t lb_tiles)
{
    const uint64_t tiles = lb_tiles ? lb_tiles : sort_tiles((uint32_t)std::max<uint64_t>(n, 1)) + nseg;
    return align_up(n * 4ull, 256) * 4 + align_up((uint64_t)nseg * SORT_SEG_HIST * 4, 256) +
           align_up(64, 256) + align_up((uint64_t)passes * tiles * 256 * 4, 256);
}

// Aggregate levels of nleaf leaves: box, count, z key, z index per entry.
static uint64_t agg_levels_bytes(uint64_t nleaf)
{
    uint64_t e = 0;
    agg_levels_layout(nleaf, &e);
    e = std::max<uint64_t>(e, 1);
    return align_up(e * 32, 256) + 2 * align_up(e * 4, 256) + align_up(e * 8, 256);
}

// Histograms, tile counters and look-back rows of the tree sort.
static uint64_t tree_sort_bytes(const TreeLayout &L, int passes)
{
    const uint64_t lb_tiles = L.sort_units[L.ntiles] + 1;
    return align_up((uint64_t)L.ntiles * SORT_SEG_HIST * 4, 256) + align_up(64, 256) +
           align_up((uint64_t)passes * lb_tiles * 256 * 4, 256) + align_up(lb_tiles * 4, 256);
}

// Device scratch of one tree build: the tree sort's counters, the point
// records between its passes (one set for two passes, two for more: the
// first pass reads the caller's arrays, the last writes the tree's), and the
// aggregate levels.
uint64_t build_scratch_bytes(const TreeLayout &L)
{
    const int passes = (L.tile_bits + L.key_bits + 7) / 8;
    const int rec_sets = passes > 2 ? 2 : passes - 1;
    return tree_sort_bytes(L, passes) + rec_sets * 2 * align_up(L.np * 16ull, 256) +
           agg_levels_bytes(L.nleaf);
}

SortBufs carve_sort(uint8_t *base, uint64_t n, int passes, uint32_t nseg, uint64_t lb_tiles)
{
    SortBufs S;
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) {
        uint8_t *p = base + o;
        o += align_up(bytes, 256);
        return p;
    };
    const uint64_t tiles = lb_tiles ? lb_tiles : sort_tiles((uint32_t)std::max<uint64_t>(n, 1)) + nseg;
    S.k0 = (uint32_t *)take(n * 4ull);
    S.v0 = (uint32_t *)take(n * 4ull);
    S.k1 = (uint32_t *)take(n * 4ull);
    S.v1 = (uint32_t *)take(n * 4ull);
    S.hist = (uint32_t *)take((uint64_t)nseg * SORT_SEG_HIST * 4);
    S.counters = (uint32_t *)take(64);  // per-pass tile counters
    S.lookback = (uint32_t *)take((uint64_t)passes * tiles * 256 * 4);
    return S;
}

// Runs the LSD passes; keys in S.k0 (values implicit iota when iota).  Returns
// which buffer pair holds the result (0 -> k0/v0, 1 -> k1/v1).
int run_sort(owm_ctx *ctx, int phase, hipStream_t s, SortBufs &S, uint32_t n_max,
             const uint32_t *d_n, int passes, bool keys_only, bool iota)
{
    // S.hist and S.lookback were zeroed with the chain's other buffers
    launch_sort_hist(s, S.k0, d_n, n_max, passes, S.hist);
    const uint32_t tiles = sort_tiles(std::max<uint32_t>(n_max, 1));
    launch_sort_scan(s, S.hist, passes, S.counters);
    uint32_t *kin = S.k0, *vin = S.v0, *kout = S.k1, *vout = S.v1;
    const int pi = ctx->prof.begin(s, phase);
    for (int p = 0; p < passes; ++p) {
        launch_onesweep(s, kin, vin, kout, vout, d_n, n_max, 8 * p, S.hist + 256 * p,
                        S.lookback + (uint64_t)p * tiles * 256, S.counters + p, keys_only,
                        iota && p == 0, true, 0);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    // algorithmic bytes per key and pass: 8 for keys only, 16 for key/value
    // pairs (12 in a first pass with iota values)
    const uint64_t per = keys_only ? 8ull : 16ull;
    ctx->prof.end(s, pi, per * n_max * passes - (iota && !keys_only ? 4ull * n_max : 0ull),
                  (uint64_t)n_max * passes, (uint32_t)passes);
    return (passes & 1) ? 1 : 0;
}

// Builds the tree of a batch into T (device memory sized by tree_bytes).
owm_status build_tree(owm_ctx *ctx, hipStream_t s, const TreeLayout &L,
                             const BatchDev &B, const double *dx, const double *dy,
                             const double *dz, const TreeOut &T, Arena &scratch, uint32_t *guard)
{
    if (L.np == 0)
        return OWM_OK;
    const int passes = (L.tile_bits + L.key_bits + 7) / 8;
    // sort tiles of the segments, and one spare look-back row
    const uint32_t tiles = L.sort_units[L.ntiles];
    uint32_t *hist = (uint32_t *)scratch.take((uint64_t)L.ntiles * SORT_SEG_HIST * 4);
    uint32_t *counters = (uint32_t *)scratch.take(64);
    uint32_t *lookback = (uint32_t *)scratch.take((uint64_t)passes * (tiles + 1) * 256 * 4);
    uint32_t *tile_seg = (uint32_t *)scratch.take((tiles + 1) * 4ull);
    uint64_t lv_entries = 0;
    AggLevels V = agg_levels_layout(L.nleaf, &lv_entries);
    lv_entries = std::max<uint64_t>(lv_entries, 1);
    V.box = (double4 *)scratch.take(lv_entries * 32);
    V.cnt = (uint32_t *)scratch.take(lv_entries * 4);
    V.zid = (uint32_t *)scratch.take(lv_entries * 4);
    V.zk = (unsigned long long *)scratch.take(lv_entries * 8);
    bool have = hist && counters && lookback && tile_seg && V.box && V.cnt && V.zid && V.zk;
    // records between passes: the cloud -> rec[0] -> rec[1] -> rec[0] ... -> the tree
    PointsOut rec[2] = {};
    for (int k = 0; k < (passes > 2 ? 2 : passes - 1); ++k) {
        rec[k].xy = (double2 *)scratch.take(L.np * 16ull);
        rec[k].zi = (PointZ *)scratch.take(L.np * 16ull);
        have = have && rec[k].xy && rec[k].zi;
    }
    if (!have) {
        set_err(ctx, "scratch arena exhausted (tree build)");
        return OWM_E_INTERNAL;
    }
    ZeroList z;
    z.add(s, guard, 64);
    if (!L.counts)
        z.add(s, hist, (uint64_t)L.ntiles * SORT_SEG_HIST);
    z.add(s, lookback, (uint64_t)passes * tiles * 256);
    z.launch(s);
    if (!L.counts) {
        int pi = ctx->prof.begin(s, PH_MORTON);
        launch_morton(s, B, dx, dy, hist, passes);
        // algorithmic bytes per point: x and y read
        ctx->prof.end(s, pi, 16ull * L.np, L.np);
    }
    SortSegs seg;
    seg.pt = B.pt_off;
    seg.units = B.sort_units;
    seg.nseg = L.ntiles;
    seg.digit_stride = SORT_SEG_HIST;
    seg.tile_seg = tile_seg;
    // digit counts (taken here, or in the bbox pass) -> the passes' offsets
    launch_sort_scan_seg(s, L.counts ? L.counts : hist, hist, seg.pt, seg.units, seg.nseg, passes,
                         counters, tile_seg);
    int pi = 0;
    uint64_t sort_bytes = 0;
    pi = ctx->prof.begin(s, PH_SORT_TREE);
    for (int p = 0; p < passes; ++p) {
        const bool last = p == passes - 1;
        const PointsIn in = p == 0 ? PointsIn{dx, dy, dz, PointsDev{}}
                                   : PointsIn{nullptr, nullptr, nullptr, rec[(p - 1) & 1]};
        const PointsOut out = last ? T.pts : rec[p & 1];
        launch_sort_records(s, B, in, out, last ? T.pos : nullptr, (uint32_t)L.np, 8 * p,
                            L.tile_bits + L.key_bits, hist + 256 * p,
                            lookback + (uint64_t)p * tiles * 256, counters + p, p == 0,
                            last ? L.stamp : 0, seg, tiles);
        // algorithmic bytes per point and pass: a 32-byte record written, and
        // read (the first pass reads x, y and z, and numbers the points)
        sort_bytes += (p == 0 ? 56ull : 64ull) * L.np + (last && T.pos ? 4ull * L.np : 0ull);
    }
    ctx->prof.end(s, pi, sort_bytes, (uint64_t)L.np * passes, (uint32_t)passes);
    pi = ctx->prof.begin(s, PH_LEAVES);
    launch_leaf_aggs(s, B, T);
    // algorithmic bytes: each point's two 16-byte streams read; 48 per leaf written
    ctx->prof.end(s, pi, 32ull * L.np + 48ull * L.nleaf, L.np);
    pi = ctx->prof.begin(s, PH_LEVELS);
    launch_agg_levels(s, T, V);
    // algorithmic bytes: 48 per leaf read, 48 per level entry written
    ctx->prof.end(s, pi, 48ull * (L.nleaf + lv_entries), L.nleaf);
    pi = ctx->prof.begin(s, PH_KARRAS);
    launch_karras(s, B, T, V);
    // per leaf: the first keys the split searches read (24 B); per node: its
    // children and aggregate written (8 + 48 B).  The aggregate-level blocks
    // that tile a node's range (at most two per level) are read on top.
    ctx->prof.end(s, pi, 24ull * L.nleaf + 56ull * L.nnode, L.nnode);
    HIP_TRY(hipGetLastError(), "tree build launch");
    return OWM_OK;
}

// ------------------------------------------------------------- queries --

uint64_t query_bytes(const QueryLayout &Q, uint64_t np, uint64_t n_in)
{
    const uint64_t nwin = std::max<uint64_t>(std::max<uint64_t>(Q.nwin, n_in), 1);
    const uint64_t nseed_max = std::max<uint64_t>(Q.nwin, n_in) + Q.ncell + 1;
    const uint64_t nwords = (np + 31) / 32 + 1;
    uint64_t b = 0;
    b += align_up(nwin * 4ull, 256);                                    // window minima
    b += align_up(nseed_max * 4ull, 256) * 2;                            // LLPs + adds, seeds
    b += align_up(nwords * 4ull, 256) * 2;                               // point bitmaps
    b += align_up(std::max<uint64_t>(Q.ncell, 1) * 4ull, 256) * 2;      // cell flag / res
    b += align_up((nseed_max / 4096 + 2) * 4ull, 256);                   // stage-3 compaction
    b += align_up((nwords / 4096 + 2) * 4ull, 256) * 2;                  // bitmap compactions
    b += align_up(1024, 256);                                           // counters
    b += align_up(stage1_work_words(Q.nblk) * 4ull, 256);              // stage-1 join
    b += align_up((Q.ntiles + 1) * 4ull, 256);                          // seed_off
    if (Q.lcs_leaves)
        b += align_up(leaf_cells_bytes(Q.lcs_leaves), 256);             // leaf cell summaries
    return b;
}

// Enqueues stages `from`..`to` (1 = stage 1, 2 = sort + stage 2, 3 = stage 3,
// 4 = merge of LLPs and stage-3 adds).  from == 2 takes the window minima
// from h_in (n_in values), from == 3 takes the LLPs from h_in.  Results
// (device): see QueryOut; cnt = [n_llp, n_add, n_seeds, error flags, ...].
owm_status enqueue_query(owm_ctx *ctx, hipStream_t s, const BatchDev &B, const TreeDev &TV,
                         const QueryLayout &Q, uint64_t np, const Stage1Cfg &cfg,
                         uint32_t any_need1, Arena &scratch, QueryOut &out, int from, int to,
                         const uint32_t *h_in, uint64_t n_in, const RunMem *mem)
{
    const uint64_t nwin = from == 2 ? n_in : Q.nwin;
    const uint64_t nw = std::max<uint64_t>(nwin, 1);
    const uint64_t nseed_max = std::max<uint64_t>(nwin, from == 3 ? n_in : 0) + Q.ncell + 1;
    const uint32_t nwords = (uint32_t)((np + 31) / 32);
    uint32_t *mins = (uint32_t *)scratch.take(nw * 4ull);
    uint32_t *seeds_a = (uint32_t *)scratch.take(nseed_max * 4ull);
    uint32_t *seeds_b = (uint32_t *)scratch.take(nseed_max * 4ull);
    uint32_t *seen = (uint32_t *)scratch.take((nwords + 1) * 4ull);
    uint32_t *llpmap = (uint32_t *)scratch.take((nwords + 1) * 4ull);
    uint32_t *cell_flag = (uint32_t *)scratch.take(std::max<uint64_t>(Q.ncell, 1) * 4ull);
    uint32_t *cell_res = (uint32_t *)scratch.take(std::max<uint64_t>(Q.ncell, 1) * 4ull);
    uint32_t *lb_cells = (uint32_t *)scratch.take((nseed_max / 4096 + 2) * 4ull);
    uint32_t *lb_llp = (uint32_t *)scratch.take((nwords / 4096 + 2) * 4ull);
    uint32_t *lb_seeds = (uint32_t *)scratch.take((nwords / 4096 + 2) * 4ull);
    uint32_t *cnt = (uint32_t *)scratch.take(1024);
    uint32_t *seed_off = (uint32_t *)scratch.take((Q.ntiles + 1) * 4ull);
    uint32_t *s1work = (uint32_t *)scratch.take(stage1_work_words(Q.nblk) * 4ull);
    uint8_t *lcs_mem = Q.lcs_leaves ? scratch.take(leaf_cells_bytes(Q.lcs_leaves)) : nullptr;
    if (!mins || !seeds_a || !seeds_b || !seen || !llpmap || !cell_flag || !cell_res ||
        !lb_cells || !lb_llp || !lb_seeds || !cnt || !seed_off || !s1work ||
        (Q.lcs_leaves && !lcs_mem)) {
        set_err(ctx, "scratch arena exhausted (query)");
        return OWM_E_INTERNAL;
    }
    out.cnt = cnt;
    out.seed_off = seed_off;
    // cnt = [n_llp, n_add, n_seeds, error flags, stage-2 tile counter,
    //        stage-3 tile counter, join total, merge tile counter, ...]
    ZeroList z;
    z.add(s, cnt, 16);
    z.add(s, s1work, stage1_work_words(Q.nblk));
    z.add(s, seen, nwords + 1);
    z.add(s, llpmap, nwords + 1);
    z.add(s, lb_llp, nwords / 4096 + 2);
    z.add(s, lb_cells, nseed_max / 4096 + 2);
    z.add(s, lb_seeds, nwords / 4096 + 2);
    z.add(s, cell_flag, std::max<uint64_t>(Q.ncell, 1));
    z.launch(s);
    if (from <= 1) {
        const uint64_t list_words = mem ? mem->s1_words : ctx->s1_list.bytes / 4;
        const uint32_t s1cap = (uint32_t)std::min<uint64_t>(Q.s1_cap, list_words);
        uint32_t *list = mem ? mem->s1_list : (uint32_t *)ctx->s1_list.ptr;
        int pi = ctx->prof.begin(s, PH_S1_JOIN);
        launch_stage1_join(s, B, TV, cfg, s1work, list, s1cap, cnt + 3, cnt + 6);
        // per leaf: bbox + count read by both join passes; per block: count, start
        ctx->prof.end(s, pi, 72ull * B.nleaf + 12ull * Q.nblk, B.nleaf);
        LeafCells LC{};
        if (lcs_mem) {
            LC = carve_leaf_cells(lcs_mem, Q.lcs_leaves);
            pi = ctx->prof.begin(s, PH_LCS);
            launch_leaf_cells(s, B, TV, LC);
            ctx->prof.end(s, pi, 28ull * np + 4ull * B.nleaf, np);  // points read once
        }
        pi = ctx->prof.begin(s, PH_STAGE1);
        launch_stage1_bins(s, B, TV, cfg, lcs_mem ? &LC : nullptr, s1work, list, s1cap, mins);
        // the algorithmic stage-1 traffic: every point streamed once per covering
        // window row (24 B x Wsize/Displace) and one minimum written per window
        ctx->prof.end(s, pi, Q.s1_alg_bytes + 4ull * Q.nwin, np);
        out.min_ids = mins;
    }
    if (to <= 1) {
        HIP_TRY(hipGetLastError(), "stage 1 launch");
        return OWM_OK;
    }
    if (from <= 2) {
        if (from == 2 && nwin)
            HIP_TRY(hipMemcpyAsync(mins, h_in, nwin * 4ull, hipMemcpyHostToDevice, s), "min_ids H2D");
        int pi = ctx->prof.begin(s, PH_S2_MARK);
        launch_stage2_mark(s, B, mins, (uint32_t)nwin, (uint32_t)np, any_need1, seen, llpmap);
        // per window: its minimum read, one bitmap word read and written
        ctx->prof.end(s, pi, 12ull * nwin, nwin);
        pi = ctx->prof.begin(s, PH_STAGE2);
        launch_bits_compact(s, llpmap, nwords, seeds_a, cnt + 0, lb_llp, cnt + 4);
        ctx->prof.end(s, pi, 4ull * nwords, nwords);
    } else {
        if (n_in)
            HIP_TRY(hipMemcpyAsync(seeds_a, h_in, n_in * 4ull, hipMemcpyHostToDevice, s), "llp H2D");
        uint32_t n32 = (uint32_t)n_in;
        HIP_TRY(hipMemcpyAsync(cnt + 0, &n32, 4, hipMemcpyHostToDevice, s), "n_llp H2D");
        HIP_TRY(hipStreamSynchronize(s), "n_llp H2D");
    }
    out.llp = seeds_a;
    if (to <= 2) {
        HIP_TRY(hipGetLastError(), "stage 2 launch");
        return OWM_OK;
    }
    if (Q.ncell > 0) {
        int pi = ctx->prof.begin(s, PH_STAGE3);
        launch_stage3(s, B, TV, seeds_a, cnt + 0, (uint32_t)(nseed_max - Q.ncell), cell_flag,
                      cell_res, cnt + 3);
        launch_compact_valid(s, cell_res, (uint32_t)Q.ncell, seeds_a, cnt + 0, cnt + 1, lb_cells,
                             cnt + 5);
        ctx->prof.end(s, pi, 12ull * Q.ncell, Q.ncell);
    }
    out.cell_res = cell_res;
    launch_add_u32(s, cnt + 0, cnt + 1, cnt + 2);
    if (to <= 3) {
        HIP_TRY(hipGetLastError(), "stage 3 launch");
        return OWM_OK;
    }
    // merge: the stage-3 additions (and LLPs that came from the host) join the
    // LLP bitmap, whose compaction lists the seeds ascending and distinct
    int pi = ctx->prof.begin(s, PH_MERGE);
    if (from == 3)
        launch_bits_set(s, seeds_a, nullptr, cnt + 0, (uint32_t)nseed_max, (uint32_t)np, llpmap);
    if (Q.ncell > 0)
        launch_bits_set(s, seeds_a, cnt + 0, cnt + 1, (uint32_t)Q.ncell, (uint32_t)np, llpmap);
    launch_bits_compact(s, llpmap, nwords, seeds_b, cnt + 2, lb_seeds, cnt + 7);
    ctx->prof.end(s, pi, 4ull * nwords + 8ull * Q.ncell, nwords,
                  1u + (from == 3 ? 1u : 0u) + (Q.ncell > 0 ? 1u : 0u));
    out.seeds = seeds_b;
    launch_seed_ranges(s, B, seeds_b, cnt + 2, seed_off);
    launch_seeds_local(s, B, seeds_b, cnt + 2, (uint32_t)nseed_max);
    HIP_TRY(hipGetLastError(), "query launch");
    return OWM_OK;
}

// ------------------------------------------------------------- context --

uint8_t *Arena::take(uint64_t bytes)
{
    bytes = align_up(std::max<uint64_t>(bytes, 1), 256);
    if (used + bytes > buf.bytes)
        return nullptr;
    uint8_t *p = (uint8_t *)buf.ptr + used;
    used += bytes;
    return p;
}

owm_status Arena::ensure(owm_ctx *ctx, uint64_t bytes)
{
    used = 0;
    if (bytes <= buf.bytes)
        return OWM_OK;
    if (ctx->budget && ctx->held_bytes() - buf.bytes + bytes > ctx->budget) {
        set_err(ctx, "work needs %llu device bytes, budget is %llu", (unsigned long long)bytes,
                (unsigned long long)ctx->budget);
        return OWM_E_BUDGET;
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
        return hip_fail(ctx, e, "sync before arena growth");
    e = buf.reserve(bytes, ctx->room_for(buf));
    if (this == &ctx->arena)
        ctx->graphs.clear();  // captured chains point into the old arena
    if (e != hipSuccess)
        return hip_fail(ctx, e, "arena allocation");
    return OWM_OK;
}


// -------------------------------------------------------------- batches --

static owm_status check_tiles(owm_ctx *ctx, const owm_tile *tiles, uint32_t n_tiles)
{
    if (n_tiles == 0) {
        set_err(ctx, "n_tiles must be > 0");
        return OWM_E_ARG;
    }
    uint64_t total = 0;
    for (uint32_t t = 0; t < n_tiles; ++t) {
        total += tiles[t].n;
        uint32_t L = tiles[t].params.leaf_size;
        if (L > 1024) {
            set_err(ctx, "tile %u: leaf_size must be <= 1024", t);
            return OWM_E_ARG;
        }
    }
    if (total >= (1ull << 30)) {
        set_err(ctx, "a batch holds at most 2^30-1 points (got %llu)", (unsigned long long)total);
        return OWM_E_ARG;
    }
    return OWM_OK;
}

uint32_t pick_leaf(owm_ctx *ctx, uint64_t n, uint32_t requested)
{
    if (requested)
        return requested;
    return ctx->tuner.leaf_for(n);
}

Stage1Cfg pick_stage1(owm_ctx *ctx, uint64_t mean_tile)
{
    Stage1Cfg c = ctx->tuner.stage1_for(mean_tile);
    // "bx,by,k,lcs" forces one stage-1 configuration (for experiments)
    if (const char *ov = getenv("OWM_S1_SHAPE")) {
        Stage1Cfg f{};
        if (sscanf(ov, "%u,%u,%u,%u", &f.bx, &f.by, &f.k, &f.lcs) == 4 && stage1_shape_ok(f))
            c = f;
    }
    return c;
}

// Per-tile bbox (caller's, else an exact device reduction), leaf sizes and the
// tree layout of a batch whose points are in device memory.
owm_status prepare_tree(owm_ctx *ctx, const double *dx, const double *dy, const owm_tile *tiles,
                        uint32_t n_tiles, Prepared &P, const RunMem *mem)
{
    owm_status st = check_tiles(ctx, tiles, n_tiles);
    if (st != OWM_OK)
        return st;
    hipStream_t s = ctx->stream;
    P.td.assign(n_tiles, TileDev{});
    P.n.resize(n_tiles);
    P.leaf.resize(n_tiles);
    P.box.resize(n_tiles);
    bool need_bbox = false;
    uint64_t np = 0;
    for (uint32_t t = 0; t < n_tiles; ++t) {
        memset(&P.td[t], 0, sizeof(TileDev));
        P.n[t] = tiles[t].n;
        np += P.n[t];
        P.leaf[t] = pick_leaf(ctx, P.n[t], tiles[t].params.leaf_size);
        if (tiles[t].params.flags & OWM_F_HAVE_BBOX) {
            for (int k = 0; k < 4; ++k)
                P.box[t][k] = tiles[t].params.bbox[k];
        } else {
            need_bbox = true;
        }
        if (tiles[t].offset != (t ? tiles[t - 1].offset + tiles[t - 1].n : 0)) {
            set_err(ctx, "tile %u: tiles must be consecutive slices of the cloud", t);
            return OWM_E_ARG;
        }
    }
    const int sort_bits = (int)ctx->tuner.sort_bits_for(np / std::max<uint32_t>(n_tiles, 1));
    // With the context's own arenas the bbox pass also takes the sort keys'
    // digit counts (over frames sampled from each tile), so the build does not
    // read x and y again; the no-allocation path keeps its borrowed region
    // free between the passes and takes them in the build.
    const bool fuse = need_bbox && !mem;
    std::vector<KeyFrame> frames;
    uint32_t *counts = nullptr;
    if (need_bbox) {
        // tile offsets first (the boxes are placeholders until the pass below)
        for (uint32_t t = 0; t < n_tiles; ++t)
            if (!(tiles[t].params.flags & OWM_F_HAVE_BBOX))
                P.box[t] = {0.0, 1.0, 0.0, 1.0};
        plan_tree(sort_bits, P.td, P.n, P.leaf, P.box, P.L);
        const int passes = (P.L.tile_bits + P.L.key_bits + 7) / 8;
        const uint64_t bk_bytes = bbox_key_words(n_tiles) * 8;
        const uint64_t fr_bytes = (uint64_t)n_tiles * sizeof(KeyFrame);
        const uint64_t cnt_words = (uint64_t)n_tiles * SORT_SEG_HIST;
        uint64_t bytes = align_up(bk_bytes, 256) + align_up((n_tiles + 1) * 4ull, 256);
        if (fuse)
            bytes += align_up(fr_bytes, 256) + align_up(cnt_words * 4, 256);
        // the bbox pass's scratch: the context's small arena, or the start of
        // a borrowed region (released again once the boxes are on the host)
        Arena &SA = mem ? *mem->arena : ctx->small;
        if (mem)
            SA.used = 0;
        else if ((st = ctx->small.ensure(ctx, bytes)) != OWM_OK)
            return st;
        unsigned long long *bk = (unsigned long long *)SA.take(bk_bytes);
        uint32_t *ptoff = (uint32_t *)SA.take((n_tiles + 1) * 4ull);
        KeyFrame *d_frames = fuse ? (KeyFrame *)SA.take(fr_bytes) : nullptr;
        counts = fuse ? (uint32_t *)SA.take(cnt_words * 4) : nullptr;
        if (!bk || !ptoff || (fuse && (!d_frames || !counts))) {
            set_err(ctx, "scratch exhausted (bbox pass)");
            return OWM_E_BUDGET;
        }
        const IoLayout io(n_tiles);
        uint8_t *h = mem ? mem->h_io : ctx->pinned(ctx->h_io, ctx->h_io_cap, io.bytes);
        if (!h) {
            set_err(ctx, "pinned host allocation failed");
            return OWM_E_NOMEM;
        }
        memcpy(h + io.ptoff, P.L.pt_off.data(), (n_tiles + 1) * 4ull);
        HIP_TRY(hipMemcpyAsync(ptoff, h + io.ptoff, (n_tiles + 1) * 4ull, hipMemcpyHostToDevice, s),
                "upload offsets");
        BatchDev B0{};
        B0.ntiles = n_tiles;
        B0.np = (uint32_t)np;
        B0.pt_off = ptoff;
        launch_bbox_init(s, bk, n_tiles);
        if (fuse)
            launch_fill_u32(s, counts, cnt_words, 0u);
        int pi = ctx->prof.begin(s, PH_BBOX);
        if (np && fuse)
            launch_bbox_keys(s, B0, dx, dy, d_frames, bk, counts, passes, (uint32_t)P.L.key_bits);
        else if (np)
            launch_bbox(s, B0, dx, dy, bk);
        ctx->prof.end(s, pi, 16ull * np, np);
        const unsigned long long *hk = (const unsigned long long *)(h + io.bbox);
        HIP_TRY(hipMemcpyAsync(h + io.bbox, bk, n_tiles * 32ull, hipMemcpyDeviceToHost, s), "bbox D2H");
        if (fuse)
            HIP_TRY(hipMemcpyAsync(h + io.frames, d_frames, fr_bytes, hipMemcpyDeviceToHost, s),
                    "frames D2H");
        HIP_TRY(hipStreamSynchronize(s), "bbox sync");
        if (fuse) {
            frames.resize(n_tiles);
            memcpy(frames.data(), h + io.frames, fr_bytes);
        }
        ctx->prof.collect();
        auto dec = [](unsigned long long k) {
            uint64_t b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
            double v;
            memcpy(&v, &b, 8);
            return v;
        };
        for (uint32_t t = 0; t < n_tiles; ++t)
            if (!(tiles[t].params.flags & OWM_F_HAVE_BBOX)) {
                if (P.n[t] == 0 || hk[4 * t] == ~0ull || hk[4 * t + 1] == ~0ull) {
                    set_err(ctx, "tile %u: no finite points to take a bbox from", t);
                    return OWM_E_ARG;
                }
                P.box[t] = {dec(hk[4 * t]), dec(hk[4 * t + 2]), dec(hk[4 * t + 1]),
                            dec(hk[4 * t + 3])};
            }
        if (mem)
            SA.used = 0;
    }
    plan_tree(sort_bits, P.td, P.n, P.leaf, P.box, P.L, fuse ? &frames : nullptr);
    P.L.counts = counts;
    return OWM_OK;
}

// Query fields of every tile (derived parameters) and the query layout.
// Expected size of the leaf/block join: a leaf of side ~sqrt(L/density)
// meets about (side + 2h + 2D) / (block side) + 1 blocks per axis.
static uint64_t join_estimate(const Prepared &P)
{
    double est = 4096.0;
    for (const TileDev &D : P.td) {
        if (D.nleaf == 0 || D.ncols == 0)
            continue;
        const double dens =
            std::max(D.n / std::max(1e-9, (double)D.ncols * D.nrows * D.displace * D.displace), 1e-9);
        const double side = std::sqrt(D.leaf_size / dens);
        const double ext = side + 2.0 * D.h + 2.0 * D.displace;
        const double fx = ext / (P.cfg.bx * D.displace) + 1.0, fy = ext / (P.cfg.by * D.displace) + 1.0;
        est += 1.5 * D.nleaf * fx * fy + 64;
    }
    return (uint64_t)std::min(est, 4.0e9);
}

// A join that overflowed before sets the floor (and keeps captured chains
// valid); OWM_S1_LIST_CAP overrides the first guess (used to exercise the
// re-run).
static void join_floor(owm_ctx *ctx, Prepared &P)
{
    if (ctx->s1_need_max)
        P.Q.s1_cap = std::max<uint64_t>(P.Q.s1_cap, ctx->s1_need_max + ctx->s1_need_max / 8 + 1024);
    if (const char *ov = getenv("OWM_S1_LIST_CAP"))
        P.Q.s1_cap = std::max<uint64_t>(1, strtoull(ov, nullptr, 10));
}

owm_status prepare_query(owm_ctx *ctx, const owm_params *params, uint32_t stride_bytes,
                         const std::vector<std::array<double, 4>> &box, uint64_t np_total,
                         Prepared &P)
{
    const uint32_t n_tiles = (uint32_t)P.td.size();
    P.any_need1 = false;
    double s1_alg = 0.0;
    for (uint32_t t = 0; t < n_tiles; ++t) {
        const owm_params &pp = *(const owm_params *)((const uint8_t *)params + (size_t)t * stride_bytes);
        Derived D;
        std::string why;
        const double *b = (pp.flags & OWM_F_HAVE_BBOX) ? pp.bbox : box[t].data();
        owm_status st = derive_params(pp, P.td[t].n, b, D, why);
        if (st != OWM_OK) {
            set_err(ctx, "tile %u: %s", t, why.c_str());
            return st;
        }
        fill_query(P.td[t], pp, D);
        P.any_need1 |= pp.overlap == 0.0;
        s1_alg += 24.0 * (double)P.td[t].n * ((double)pp.wsize / D.displace);
    }
    P.Q.s1_alg_bytes = (uint64_t)s1_alg;
    P.cfg = pick_stage1(ctx, np_total / std::max<uint32_t>(n_tiles, 1));
    P.Q.ntiles = n_tiles;
    plan_query(P.td, P.cfg, P.Q);
    P.Q.s1_cap = join_estimate(P);
    // Leaf cell summaries pay off when a leaf spans few window-edge classes:
    // estimate the classes a leaf of side sqrt(L / density) meets (edges every
    // D, or every D/2 when left and right edges do not coincide).
    P.Q.lcs_leaves = 0;
    for (uint32_t t = 0; t < n_tiles; ++t) {
        TileDev &D = P.td[t];
        D.lcs = 0;
        if (!P.cfg.lcs || D.nleaf == 0 || D.ncols == 0 || D.leaf_size > 64 || D.ncols > 16383 ||
            D.nrows > 8191 || P.L.nleaf >= (1ull << 27))
            continue;
        const double dens = std::max(D.n / std::max(1e-9, (double)D.ncols * D.nrows * D.displace *
                                                             D.displace), 1e-9);
        const double side = std::sqrt(D.leaf_size / dens);
        const double q = 2.0 * D.h / D.displace;
        const double spacing = std::fabs(q - std::round(q)) < 1e-6 ? D.displace : D.displace / 2;
        // a Morton-ordered leaf enters a class about twice on average
        const double runs = 2.0 * (side / spacing + 1.0) * (side / spacing + 1.0);
        if (runs <= LCS_SLOTS) {
            D.lcs = 1;
            P.Q.lcs_leaves = P.L.nleaf;
        }
    }
    join_floor(ctx, P);
    if (P.Q.nwin >= (1ull << 31) || P.Q.ncell >= (1ull << 31)) {
        set_err(ctx, "too many windows or cells in one batch");
        return OWM_E_ARG;
    }
    return OWM_OK;
}

// Uploads the tile table and offset arrays; fills B's pointers.
owm_status upload_table(owm_ctx *ctx, Arena &A, const Prepared &P, BatchDev &B,
                        const RunMem *mem)
{
    hipStream_t s = ctx->stream;
    const uint32_t n_tiles = (uint32_t)P.td.size();
    // tiles, then the eight offset arrays, as staged by write_table
    uint8_t *d = A.take(table_bytes(n_tiles));
    if (!d) {
        set_err(ctx, "scratch arena exhausted (table)");
        return OWM_E_INTERNAL;
    }
    const uint8_t *h = nullptr;
    if (mem) {
        write_table(mem->h_tab, P.td, P.L, P.Q);
        h = mem->h_tab;
    } else if (ctx->stage_table(P.td, P.L, P.Q)) {
        h = ctx->h_tab;
    } else {
        set_err(ctx, "pinned host allocation failed");
        return OWM_E_NOMEM;
    }
    HIP_TRY(hipMemcpyAsync(d, h, table_bytes(n_tiles), hipMemcpyHostToDevice, s), "upload table");
    TileDev *d_tiles = (TileDev *)d;
    uint32_t *d_off[8];
    uint64_t o = align_up((uint64_t)n_tiles * sizeof(TileDev), 256);
    for (int k = 0; k < 8; ++k, o += align_up((n_tiles + 1) * 4ull, 256))
        d_off[k] = (uint32_t *)(d + o);
    B = BatchDev{};
    B.ntiles = n_tiles;
    B.np = (uint32_t)P.L.np;
    B.nleaf = (uint32_t)P.L.nleaf;
    B.nnode = (uint32_t)P.L.nnode;
    B.nwin = (uint32_t)P.Q.nwin;
    B.ncell = (uint32_t)P.Q.ncell;
    B.nblk = (uint32_t)P.Q.nblk;
    B.key_bits = (uint32_t)P.L.key_bits;
    B.tiles = d_tiles;
    B.pt_off = d_off[0];
    B.leaf_off = d_off[1];
    B.node_off = d_off[2];
    B.win_off = d_off[3];
    B.cell_off = d_off[4];
    B.blk_off = d_off[5];
    B.grp_off = d_off[6];
    B.sort_units = d_off[7];
    B.ngrp = (uint32_t)P.L.ngrp;
    return OWM_OK;
}

uint64_t table_bytes(uint32_t n_tiles)
{
    return align_up((uint64_t)n_tiles * sizeof(TileDev), 256) + 8 * align_up((n_tiles + 1) * 4ull, 256);
}

// Reads the counts, checks the device guards, copies the seeds (tile-local).
owm_status ensure_s1_list(owm_ctx *ctx, uint64_t cap)
{
    if (cap * 4 <= ctx->s1_list.bytes)
        return OWM_OK;
    if (ctx->budget && ctx->held_bytes() - ctx->s1_list.bytes + cap * 4 > ctx->budget) {
        set_err(ctx, "stage-1 join list needs %llu bytes over the budget", (unsigned long long)(cap * 4));
        return OWM_E_BUDGET;
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream), "sync before list growth");
    HIP_TRY(ctx->s1_list.reserve(cap * 4, ctx->room_for(ctx->s1_list)), "stage-1 join list");
    return OWM_OK;
}

owm_status check_guards(owm_ctx *ctx, const uint32_t *hc, uint64_t *s1_needed)
{
    *s1_needed = 0;
    uint32_t f = hc[3];
    if (f & 16u) {
        *s1_needed = hc[6];
        f &= ~16u;
    }
    if (f) {
        set_err(ctx, "device capacity guard tripped (flags %u)", f);
        return OWM_E_INTERNAL;
    }
    return OWM_OK;
}

owm_status read_seeds(owm_ctx *ctx, const QueryOut &qo, const Prepared &P, uint32_t *seeds,
                      uint64_t cap, uint64_t *seed_off_host, uint64_t *s1_needed,
                      const RunMem *mem)
{
    hipStream_t s = ctx->stream;
    const uint32_t n_tiles = (uint32_t)P.td.size();
    const IoLayout io(n_tiles);
    uint8_t *h = mem ? mem->h_io : ctx->pinned(ctx->h_io, ctx->h_io_cap, io.bytes);
    if (!h) {
        set_err(ctx, "pinned host allocation failed");
        return OWM_E_NOMEM;
    }
    const uint32_t *hc = (const uint32_t *)(h + io.counts);
    const uint32_t *hoff = (const uint32_t *)(h + io.seed_off);
    HIP_TRY(hipMemcpyAsync(h + io.counts, qo.cnt, 16 * 4, hipMemcpyDeviceToHost, s), "counts D2H");
    HIP_TRY(hipMemcpyAsync(h + io.seed_off, qo.seed_off, (n_tiles + 1) * 4ull, hipMemcpyDeviceToHost, s),
            "ranges D2H");
    HIP_TRY(hipStreamSynchronize(s), "pipeline");
    ctx->prof.collect();
    owm_status st = check_guards(ctx, hc, s1_needed);
    if (st != OWM_OK || *s1_needed)
        return st;
    const uint64_t ns = hc[2];
    for (uint32_t t = 0; t <= n_tiles; ++t)
        seed_off_host[t] = hoff[t];
    if (ns > cap) {
        set_err(ctx, "seed buffer holds %llu, %llu needed", (unsigned long long)cap,
                (unsigned long long)ns);
        return OWM_E_CAPACITY;
    }
    // indices are tile-local already (launch_seeds_local)
    if (ns)
        HIP_TRY(hipMemcpy(seeds, qo.seeds, ns * 4ull, hipMemcpyDeviceToHost), "seeds D2H");
    return OWM_OK;
}

// Key of a captured batch chain: everything its launches depend on (device
// addresses, layout, tile table, per-launch configuration).
static void graph_key(owm_ctx *ctx, const double *dx, const double *dy, const double *dz,
                      const Prepared &P, std::vector<uint8_t> &key)
{
    auto put = [&key](const void *p, size_t n) {
        const uint8_t *b = (const uint8_t *)p;
        key.insert(key.end(), b, b + n);
    };
    key.clear();
    const void *ptrs[6] = {dx, dy, dz, ctx->arena.buf.ptr, ctx->s1_list.ptr, P.L.counts};
    put(ptrs, sizeof(ptrs));
    const uint64_t nums[6] = {ctx->arena.buf.bytes, ctx->s1_list.bytes, P.Q.s1_cap,
                              (uint64_t)P.L.key_bits, (uint64_t)P.L.tile_bits,
                              (uint64_t)P.any_need1};
    put(nums, sizeof(nums));
    put(&P.cfg, sizeof(P.cfg));
    put(P.td.data(), P.td.size() * sizeof(TileDev));
}

GraphCache::~GraphCache() { clear(); }

void GraphCache::clear()
{
    for (auto &s : slots)
        if (s.exec)
            (void)hipGraphExecDestroy(s.exec);
    slots.clear();
    last_key.clear();
}

// Enqueues (or replays) the device chain of a batch: guard reset, tree build,
// stages 1-3 and the merge.  With a graph, the chain is captured on first use
// and replayed while its key matches.
static owm_status enqueue_chain(owm_ctx *ctx, bool graph, const double *dx, const double *dy,
                                const double *dz, Prepared &P, BatchDev &B, TreeOut &T,
                                uint32_t *guard, QueryOut &qo, const RunMem *mem = nullptr)
{
    hipStream_t s = ctx->stream;
    Arena &A = mem ? *mem->arena : ctx->arena;
    auto body = [&]() -> owm_status {
        // build_tree zeroes the guard with its own scratch
        owm_status st = build_tree(ctx, s, P.L, B, dx, dy, dz, T, A, guard);
        if (st != OWM_OK)
            return st;
        TreeDev TV = tree_view(T);
        TV.X = dx;
        TV.Y = dy;
        st = enqueue_query(ctx, s, B, TV, P.Q, P.L.np, P.cfg, P.any_need1 ? 1u : 0u, A, qo, 1, 4,
                           nullptr, 0, mem);
        if (st != OWM_OK)
            return st;
        launch_add_u32(s, qo.cnt + 3, guard, qo.cnt + 3);
        return OWM_OK;
    };
    if (!graph)
        return body();
    GraphCache &G = ctx->graphs;
    std::vector<uint8_t> key;
    graph_key(ctx, dx, dy, dz, P, key);
    for (size_t i = 0; i < G.slots.size(); ++i)
        if (G.slots[i].key == key) {
            GraphCache::Slot hit = std::move(G.slots[i]);
            G.slots.erase(G.slots.begin() + i);
            qo = hit.qo;
            A.used = hit.used_after;
            hipError_t e = hipGraphLaunch(hit.exec, s);
            G.slots.insert(G.slots.begin(), std::move(hit));
            if (e != hipSuccess)
                return hip_fail(ctx, e, "graph launch");
            return OWM_OK;
        }
    // capture: phase timers are host events and stay out of the graph
    const bool prof_on = ctx->prof.on;
    ctx->prof.on = false;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        ctx->prof.on = prof_on;
        return hip_fail(ctx, e, "begin capture");
    }
    owm_status st = body();
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    ctx->prof.on = prof_on;
    if (st != OWM_OK) {
        if (g)
            (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess)
        return hip_fail(ctx, e, "end capture");
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess)
        return hip_fail(ctx, e, "graph instantiate");
    GraphCache::Slot slot;
    slot.key = std::move(key);
    slot.exec = exec;
    slot.qo = qo;
    slot.used_after = A.used;
    G.slots.insert(G.slots.begin(), std::move(slot));
    while (G.slots.size() > GraphCache::MAX_SLOTS) {
        (void)hipGraphExecDestroy(G.slots.back().exec);
        G.slots.pop_back();
    }
    e = hipGraphLaunch(exec, s);
    if (e != hipSuccess)
        return hip_fail(ctx, e, "graph launch");
    return OWM_OK;
}

// Runs the whole pipeline on a batch whose points are in device memory
// (concatenated SoA).  seeds/seed_off_host receive tile-local ids.  With
// OWM_B_GRAPH the device chain is captured once per batch layout and replayed.
owm_status run_batch(owm_ctx *ctx, const double *dx, const double *dy, const double *dz,
                     const owm_tile *tiles, uint32_t n_tiles, uint32_t flags, uint32_t *seeds,
                     uint64_t cap, uint64_t *seed_off_host)
{
    Prepared P;
    owm_status st = prepare_tree(ctx, dx, dy, tiles, n_tiles, P);
    if (st != OWM_OK)
        return st;
    st = prepare_query(ctx, &tiles[0].params, sizeof(owm_tile), P.box, P.L.np, P);
    if (st != OWM_OK)
        return st;
    stamp_classes(ctx, P);
    bool graph = (flags & OWM_B_GRAPH) != 0;
    const bool auto_graph = (flags & B_AUTO_GRAPH) != 0 && !graph;
    for (int attempt = 0;; ++attempt) {
        const uint64_t need = table_bytes(n_tiles) + tree_bytes(P.L) + build_scratch_bytes(P.L) +
                              query_bytes(P.Q, P.L.np, 0) + 8192;
        void *old_arena = ctx->arena.buf.ptr;
        st = ctx->arena.ensure(ctx, need);
        if (st != OWM_OK)
            return st;
        st = ensure_s1_list(ctx, P.Q.s1_cap);
        if (st != OWM_OK)
            return st;
        if (ctx->arena.buf.ptr != old_arena)
            ctx->graphs.clear();  // captured chains point into the old arena
        if (auto_graph) {
            // a call that repeats the previous one's layout exactly (same
            // inputs, arena, table and configuration) replays its chain from a
            // graph, captured on the first repeat; phase timing keeps it eager
            std::vector<uint8_t> key;
            graph_key(ctx, dx, dy, dz, P, key);
            graph = !ctx->prof.on && key == ctx->graphs.last_key;
            ctx->graphs.last_key.swap(key);
        }
        Arena &A = ctx->arena;
        uint32_t *guard = (uint32_t *)A.take(256);
        BatchDev B;
        st = upload_table(ctx, A, P, B);
        if (st != OWM_OK)
            return st;
        TreeOut T;
        carve_tree(A.take(tree_bytes(P.L)), P.L, T);
        T.pos = nullptr;  // the cloud stays available: stage 3 reads LLP coordinates from it
        QueryOut qo;
        st = enqueue_chain(ctx, graph, dx, dy, dz, P, B, T, guard, qo);
        if (st != OWM_OK)
            return st;
        uint64_t s1_need = 0;
        st = read_seeds(ctx, qo, P, seeds, cap, seed_off_host, &s1_need);
        if (st != OWM_OK || s1_need == 0)
            return st;
        if (attempt > 0) {
            set_err(ctx, "stage-1 join list overflow persisted");
            return OWM_E_INTERNAL;
        }
        // the join's exact size is known now: run again with room for it
        ctx->s1_need_max = std::max<uint64_t>(ctx->s1_need_max, s1_need);
        P.Q.s1_cap = s1_need + s1_need / 8 + 1024;
    }
}

uint64_t io_bytes(uint32_t n_tiles) { return IoLayout(n_tiles).bytes; }

// One tile inside a borrowed device region with caller-provided host staging:
// the steps of run_batch (bbox, plan, table, chain, readback) with every
// buffer carved from [base, base + bytes), so that nothing is allocated:
// no device or page-locked memory, no events (phase timing is paused), and the
// planning vectors are the context's, reserved when it was created.
owm_status run_tile_fixed(owm_ctx *ctx, const double *dx, const double *dy, const double *dz,
                          uint64_t n, const owm_params &params, uint32_t *seeds, uint64_t cap,
                          uint64_t *n_seeds, uint8_t *base, uint64_t bytes, uint8_t *h_io,
                          uint8_t *h_tab, uint64_t *need)
{
    *need = 0;
    Arena A;
    A.buf.ptr = base;
    A.buf.bytes = bytes;
    A.buf.borrowed = true;
    RunMem mem;
    mem.arena = &A;
    mem.h_io = h_io;
    mem.h_tab = h_tab;
    owm_tile tile;
    tile.offset = 0;
    tile.n = n;
    tile.params = params;
    struct Pause {  // phase timing would create events
        Profiler &p;
        bool on;
        explicit Pause(Profiler &q) : p(q), on(q.on) { p.on = false; }
        ~Pause() { p.on = on; }
    } pause(ctx->prof);
    Prepared &P = ctx->fixed_plan;
    owm_status st = prepare_tree(ctx, dx, dy, &tile, 1, P, &mem);
    if (st != OWM_OK)
        return st;
    st = prepare_query(ctx, &tile.params, sizeof(owm_tile), P.box, P.L.np, P);
    if (st != OWM_OK)
        return st;
    stamp_classes(ctx, P);
    for (int attempt = 0;; ++attempt) {
        const uint64_t list_bytes = align_up(std::max<uint64_t>(P.Q.s1_cap, 1) * 4, 256);
        *need = 256 + table_bytes(1) + tree_bytes(P.L) + build_scratch_bytes(P.L) +
                query_bytes(P.Q, P.L.np, 0) + list_bytes + 8192;
        if (*need > bytes) {
            set_err(ctx, "a tile of %llu points needs %llu bytes of scratch; this entry point "
                         "has %llu",
                    (unsigned long long)n, (unsigned long long)*need, (unsigned long long)bytes);
            return OWM_E_BUDGET;
        }
        A.used = 0;
        uint32_t *guard = (uint32_t *)A.take(256);
        mem.s1_list = (uint32_t *)A.take(list_bytes);
        mem.s1_words = list_bytes / 4;
        BatchDev B;
        st = upload_table(ctx, A, P, B, &mem);
        if (st != OWM_OK)
            return st;
        TreeOut T;
        carve_tree(A.take(tree_bytes(P.L)), P.L, T);
        T.pos = nullptr;
        QueryOut qo;
        st = enqueue_chain(ctx, false, dx, dy, dz, P, B, T, guard, qo, &mem);
        if (st != OWM_OK)
            return st;
        uint64_t s1_need = 0, off[2] = {0, 0};
        st = read_seeds(ctx, qo, P, seeds, cap, off, &s1_need, &mem);
        *n_seeds = off[1];
        if (st != OWM_OK || s1_need == 0)
            return st;
        if (attempt > 0) {
            set_err(ctx, "stage-1 join list overflow persisted");
            return OWM_E_INTERNAL;
        }
        // the join's exact size is known now: run again with room for it
        P.Q.s1_cap = s1_need + s1_need / 8 + 1024;
    }
}

}  // namespace owm

using owm::TreeLayout;
using owm::QueryLayout;

uint8_t *owm_ctx::pinned(uint8_t *&buf, uint64_t &cap, uint64_t need)
{
    if (need <= cap)
        return buf;
    // copies from or into the old buffer must have completed
    (void)hipStreamSynchronize(stream);
    if (buf)
        (void)hipHostFree(buf);
    buf = nullptr;
    cap = 0;
    const uint64_t want = need + need / 4 + 4096;
    if (hipHostMalloc((void **)&buf, want, 0) != hipSuccess) {
        buf = nullptr;
        return nullptr;
    }
    cap = want;
    return buf;
}

void owm::write_table(uint8_t *h, const std::vector<TileDev> &td, const TreeLayout &L,
                      const QueryLayout &Q)
{
    const uint32_t T = (uint32_t)td.size();
    memcpy(h, td.data(), T * sizeof(TileDev));
    const std::vector<uint32_t> *src[8] = {&L.pt_off,   &L.leaf_off, &L.node_off,
                                           &Q.win_off,  &Q.cell_off, &Q.blk_off,
                                           &L.grp_off,  &L.sort_units};
    uint64_t o = owm::align_up((uint64_t)T * sizeof(TileDev), 256);
    for (int k = 0; k < 8; ++k, o += owm::align_up((T + 1) * 4ull, 256))
        memcpy(h + o, src[k]->data(), (T + 1) * 4);
}

bool owm_ctx::stage_table(const std::vector<TileDev> &td, const TreeLayout &L, const QueryLayout &Q)
{
    const uint64_t bytes = owm::table_bytes((uint32_t)td.size());
    // the previous upload from this buffer must have completed
    (void)hipStreamSynchronize(stream);
    uint8_t *h = pinned(h_tab, h_tab_cap, bytes);
    if (!h)
        return false;
    owm::write_table(h, td, L, Q);
    return true;
}

uint64_t owm_ctx::room_for(const owm::DevBuf &b) const
{
    if (!budget)
        return ~0ull;
    uint64_t others = held_bytes() - b.bytes;
    return budget > others ? budget - others : 0;
}

uint64_t owm_ctx::held_bytes() const
{
    uint64_t b = arena.buf.bytes + small.buf.bytes + upload.bytes + s1_list.bytes;
    b += xyz_text.bytes + xyz_work.bytes + xyz_cloud.bytes + qs_work.bytes;
    b += trees_bytes.load();
    return b;
}


