// Size-class autotuner: times trial configurations of the whole batch chain
// and keeps the fastest per size class.  Knobs: leaf size (16..128, per tile
// class), stage-1 window block (bx x by windows, k points per lane per round,
// with or without leaf cell summaries) and the sort key width (8, 16, 24 or
// 32 Morton bits: one to four 8-bit radix passes), the last two per class of
// the batch's mean tile size.  This is the GPU counterpart of the reference's tuning of its tree
// granularity by sweeps over maxNumber and TBB chunk sizes (CPP-TBB-CPU
// scripts and Results/*.csv); there the sweep is a script, here it is a
// library call whose table the batch and serving paths use.
#include "owm_host.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using namespace owm;

namespace {

constexpr uint32_t LEAVES[] = {16, 32, 64, 128};
constexpr Stage1Cfg BLOCKS[] = {{8, 8, 4, 0}, {16, 8, 4, 0}, {8, 8, 2, 0},
                                {16, 8, 4, 1}, {16, 16, 4, 0}, {32, 16, 4, 0}};
constexpr int NBLOCKS = (int)(sizeof(BLOCKS) / sizeof(BLOCKS[0]));
constexpr uint32_t SORT_BITS[] = {32, 24, 16, 8};
constexpr uint64_t SUB_POINTS = 8ull << 20;  // points per leaf-size trial batch

void to_entry(const Tuner::Entry &e, owm_tune_entry &o)
{
    memset(&o, 0, sizeof(o));
    o.max_n = e.max_n;
    o.leaf_size = e.cfg.leaf;
    o.block_x = e.cfg.s1.bx;
    o.block_y = e.cfg.s1.by;
    o.block_k = e.cfg.s1.k;
    o.summaries = e.cfg.s1.lcs;
    o.sort_bits = e.cfg.sort_bits;
    o.ms = e.ms;
}

owm_status put_table(owm_ctx *ctx, owm_tune_entry *out, uint32_t cap, uint32_t *n_out)
{
    const uint32_t n = (uint32_t)ctx->tuner.table.size();
    if (n_out)
        *n_out = n;
    if (!out)
        return OWM_OK;
    if (cap < n) {
        set_err(ctx, "tuning table has %u entries, capacity %u", n, cap);
        return OWM_E_CAPACITY;
    }
    for (uint32_t k = 0; k < n; ++k)
        to_entry(ctx->tuner.table[k], out[k]);
    return OWM_OK;
}

// Best of `reps` timed runs (after one untimed run) of a batch under the
// context's current tuner state; wall time includes the final copy of seeds.
owm_status time_batch(owm_ctx *ctx, const double *dx, const double *dy, const double *dz,
                      const std::vector<owm_tile> &tiles, int reps, double *best_ms)
{
    const uint64_t np = tiles.back().offset + tiles.back().n;
    std::vector<uint32_t> seeds(np + 1);
    std::vector<uint64_t> off(tiles.size() + 1);
    *best_ms = 1e300;
    for (int r = 0; r <= reps; ++r) {
        auto t0 = std::chrono::steady_clock::now();
        owm_status st = run_batch(ctx, dx, dy, dz, tiles.data(), (uint32_t)tiles.size(), 0,
                                  seeds.data(), seeds.size(), off.data());
        if (st != OWM_OK)
            return st;
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (r > 0)
            *best_ms = std::min(*best_ms, ms);
    }
    return OWM_OK;
}

struct ForceGuard {
    Tuner &t;
    explicit ForceGuard(Tuner &tu, const TuneCfg &c) : t(tu)
    {
        t.forced = true;
        t.force = c;
    }
    ~ForceGuard() { t.forced = false; }
};

}  // namespace

owm_status owm_tune_get(owm_ctx *ctx, owm_tune_entry *out, uint32_t cap, uint32_t *n_out)
{
    if (!ctx)
        return OWM_E_ARG;
    return put_table(ctx, out, cap, n_out);
}

owm_status owm_tune_set(owm_ctx *ctx, const owm_tune_entry *in, uint32_t n)
{
    if (!ctx || (n && !in))
        return OWM_E_ARG;
    Tuner t;
    for (uint32_t k = 0; k < n; ++k) {
        const owm_tune_entry &e = in[k];
        TuneCfg c{e.leaf_size, {e.block_x, e.block_y, e.block_k, e.summaries}, e.sort_bits};
        if (c.leaf == 0 || c.leaf > 1024 || !stage1_shape_ok(c.s1) ||
            (c.sort_bits != 8 && c.sort_bits != 16 && c.sort_bits != 24 && c.sort_bits != 32)) {
            set_err(ctx, "tuning entry %u: unsupported configuration", k);
            return OWM_E_ARG;
        }
        const int cls = Tuner::class_of(e.max_n);
        if (Tuner::CLASS_MAX[cls] != e.max_n) {
            set_err(ctx, "tuning entry %u: max_n is not a class bound", k);
            return OWM_E_ARG;
        }
        t.table[cls].cfg = c;
        t.table[cls].ms = e.ms;
    }
    ctx->tuner.table = t.table;
    return OWM_OK;
}

owm_status owm_autotune(owm_ctx *ctx, const double *x, const double *y, const double *z,
                        int device_ptrs, const owm_tile *tiles, uint32_t n_tiles,
                        owm_tune_entry *out, uint32_t cap, uint32_t *n_out)
{
    if (!ctx || !tiles || n_tiles == 0 || !x || !y || !z || (cap && !out))
        return OWM_E_ARG;
    try {
        (void)hipSetDevice(ctx->device);
        if (ctx->tuner.forced) {
            set_err(ctx, "autotune is already running on this context");
            return OWM_E_ARG;
        }
        std::vector<owm_tile> all(tiles, tiles + n_tiles);
        for (auto &t : all)
            t.params.leaf_size = 0;  // the trial decides
        const uint64_t np = all.back().offset + all.back().n;
        const double *dx = x, *dy = y, *dz = z;
        if (!device_ptrs) {
            double *ux, *uy, *uz;
            owm_status st = upload_cloud(ctx, x, y, z, np, &ux, &uy, &uz);
            if (st != OWM_OK)
                return st;
            dx = ux;
            dy = uy;
            dz = uz;
        }
        Tuner fresh;
        std::vector<Tuner::Entry> table = fresh.table;

        // Leaf size per tile class, on a batch of (up to SUB_POINTS of) that
        // class's tiles copied together.
        DevBuf sub;
        for (int cls = 0; cls < 6; ++cls) {
            std::vector<uint32_t> pick;
            uint64_t pts = 0;
            for (uint32_t t = 0; t < n_tiles && pts < SUB_POINTS; ++t)
                if (all[t].n > 0 && Tuner::class_of(all[t].n) == cls) {
                    pick.push_back(t);
                    pts += all[t].n;
                }
            if (pick.empty())
                continue;
            const uint64_t b = (pts * 8 + 255) / 256 * 256;
            if (ctx->budget && ctx->held_bytes() + 3 * b > ctx->budget)
                continue;  // no room for a trial copy: keep the default
            HIP_TRY(hipStreamSynchronize(ctx->stream), "tune sync");
            HIP_TRY(sub.reserve(3 * b), "tune buffer");
            double *sx = (double *)sub.ptr, *sy = (double *)((uint8_t *)sub.ptr + b),
                   *sz = (double *)((uint8_t *)sub.ptr + 2 * b);
            std::vector<owm_tile> st_tiles;
            uint64_t o = 0;
            for (uint32_t t : pick) {
                const owm_tile &src = all[t];
                HIP_TRY(hipMemcpyAsync(sx + o, dx + src.offset, src.n * 8, hipMemcpyDeviceToDevice, ctx->stream), "tune copy");
                HIP_TRY(hipMemcpyAsync(sy + o, dy + src.offset, src.n * 8, hipMemcpyDeviceToDevice, ctx->stream), "tune copy");
                HIP_TRY(hipMemcpyAsync(sz + o, dz + src.offset, src.n * 8, hipMemcpyDeviceToDevice, ctx->stream), "tune copy");
                owm_tile nt = src;
                nt.offset = o;
                st_tiles.push_back(nt);
                o += src.n;
            }
            double best = 1e300;
            uint32_t best_leaf = table[cls].cfg.leaf;
            for (uint32_t leaf : LEAVES) {
                ForceGuard g(ctx->tuner, TuneCfg{leaf, table[cls].cfg.s1, table[cls].cfg.sort_bits});
                double ms;
                owm_status st = time_batch(ctx, sx, sy, sz, st_tiles, 2, &ms);
                if (st != OWM_OK)
                    return st;
                if (ms < best) {
                    best = ms;
                    best_leaf = leaf;
                }
            }
            table[cls].cfg.leaf = best_leaf;
            table[cls].ms = best;
        }
        HIP_TRY(hipStreamSynchronize(ctx->stream), "tune sync");
        sub.release();

        // Window block and sort width for the class of the whole batch, with
        // the tuned leaf sizes.
        ctx->tuner.table = table;
        const int bcls = Tuner::class_of(np / n_tiles);  // the launch's mean tile
        double best = 1e300;
        TuneCfg win = table[bcls].cfg;
        for (const Stage1Cfg &s1 : BLOCKS)
            for (uint32_t sb : SORT_BITS) {
                ForceGuard g(ctx->tuner, TuneCfg{0, s1, sb});
                double ms;
                owm_status st = time_batch(ctx, dx, dy, dz, all, 2, &ms);
                if (st != OWM_OK)
                    return st;
                if (ms < best) {
                    best = ms;
                    win.s1 = s1;
                    win.sort_bits = sb;
                }
            }
        table[bcls].cfg.s1 = win.s1;
        table[bcls].cfg.sort_bits = win.sort_bits;
        table[bcls].ms = best;
        ctx->tuner.table = table;
        return put_table(ctx, out, cap, n_out);
    } catch (const std::bad_alloc &) {
        ctx->tuner.forced = false;
        set_err(ctx, "host allocation failed");
        return OWM_E_NOMEM;
    } catch (...) {
        ctx->tuner.forced = false;
        set_err(ctx, "internal error");
        return OWM_E_INTERNAL;
    }
}

// ---------------------------------------------------------------------------
// Statistics-driven tuning.  The reference chose maxNumber and minRadius by
// hand from its statistics tool's leaf histograms (CPP-TBB-CPU/application/
// owm_stats.cpp, include/stats_func.hpp:6-56).  Here a cost model of the batch
// chain reads such histograms (owm_qtree_stats, maxNumber policy) of a few
// sample tiles per size class and picks the table without timing anything.
//
// Per point of a tile, in ns, for leaf size L, window block b and key width S:
//   M_LEAF / L + M_AMP * amp0 + M_COARSE * ampc + M_PASS * S / 8 + M_BLK[b]
//   + M_READ * rf
// plus, once per launch, (M_QUANT * q + M_UNDER * u) per point.  Over the
// histogram's leaves (density rho = points / side^2, weighted by their
// points): amp0 = mean (1 + s0/ex)(1 + s0/ey) - 1 with s0 = sqrt(L/rho) the
// side of an L-point leaf and ex, ey = bx*D + W - D, by*D + W - D the extent
// of a window block (W = Wsize, D = Displace), i.e. how many window blocks
// beyond the first read a point of a leaf that straddles their edges;
// ampc = the same with s = max(s0, tile side / 2^(S/2)) minus amp0, what a key
// too short to order the points of one Morton cell adds; rf = the points the
// tile's window blocks read over its points, from its window grid (blocks at
// the grid's right and top edges hold fewer windows but the same margins).
// With `slots` = workgroups of the shape the chip holds at once and `blocks` =
// window blocks in the launch: q = ceil(blocks / slots) * slots / blocks - 1
// (the idle share of the last round of workgroups), u = max(0, slots / blocks
// - 1) (a launch that does not fill the chip).  The constants are a
// least-squares fit (a free intercept per set) to the batch times of 67
// homogeneous tile sets x 96 configurations on MI355X
// (fastowm_amd/tune/fit_model.py on fastowm_amd/tune/trials_mi355x.json); on
// those sets the model's pick is 2.2 % slower than the fastest configuration
// on average (16 % at most; 2.5 % and 17 % with each set left out of the
// fit), against 64 % for leaf 32 / 8x8x4 / 32 bits.
namespace {

constexpr double M_LEAF = 0.581932;
constexpr double M_AMP = 0.0504456;
constexpr double M_COARSE = 0.00798269;
constexpr double M_PASS = 0.0149227;
constexpr double M_BLK[NBLOCKS] = {-0.00602212, -0.00863262, -0.00479202, 0.0158341, -0.00237524, 0.00598786};
constexpr double M_READ = 0.0196671;
constexpr double M_QUANT = 0.00418061;
constexpr double M_UNDER = 0.00662172;
constexpr double CUS = 256.0;                                     // compute units of the part
constexpr double WG_PER_CU[NBLOCKS] = {4.0, 4.0, 4.0, 3.0, 2.0, 1.0};  // workgroups of each shape

struct ModelTile {
    double n = 0, weight = 0;
    double W = 0, D = 0, side = 0;
    uint32_t ncols = 0, nrows = 0;
    std::vector<double> rho, w;
    double sw = 0;
    int cls = 0;
};

bool model_tile(const owm_tune_sample &s, ModelTile &m)
{
    if (s.n == 0 || !(s.weight >= (double)s.n) || !std::isfinite(s.weight) ||
        (s.n_bins && !s.bins) || !(s.summary.max_radius > 0.0f))
        return false;
    const double *box = (s.params.flags & OWM_F_HAVE_BBOX) ? s.params.bbox : s.summary.bbox;
    Derived d;
    std::string why;
    if (derive_params(s.params, s.n, box, d, why) != OWM_OK)
        return false;
    m.n = (double)s.n;
    m.weight = s.weight;
    m.W = (double)s.params.wsize;
    m.D = d.displace;
    m.side = std::max(box[1] - box[0], box[3] - box[2]);
    m.ncols = d.ncols;
    m.nrows = d.nrows;
    m.cls = Tuner::class_of(s.n);
    const double r0 = (double)s.summary.max_radius;
    for (uint64_t k = 0; k < s.n_bins; ++k) {
        const owm_leaf_bin &b = s.bins[k];
        if (b.points == 0 || b.leaves == 0 || b.level > 62)
            continue;
        const double side = 2.0 * r0 / std::ldexp(1.0, (int)b.level);
        m.rho.push_back((double)b.points / (side * side));
        m.w.push_back((double)b.points * (double)b.leaves);
        m.sw += m.w.back();
    }
    return m.sw > 0.0;
}

// ns per point of one tile, launch term left out
double tile_cost(const ModelTile &m, uint32_t leaf, int b, uint32_t sort_bits)
{
    const Stage1Cfg &B = BLOCKS[b];
    const double ex = B.bx * m.D + m.W - m.D, ey = B.by * m.D + m.W - m.D;
    const double cell = m.side / std::ldexp(1.0, (int)sort_bits / 2);
    double a0 = 0, a = 0;
    for (size_t i = 0; i < m.rho.size(); ++i) {
        const double s0 = std::sqrt((double)leaf / m.rho[i]), s = std::max(s0, cell);
        a0 += m.w[i] * ((1 + s0 / ex) * (1 + s0 / ey) - 1);
        a += m.w[i] * ((1 + s / ex) * (1 + s / ey) - 1);
    }
    a0 /= m.sw;
    a /= m.sw;
    const double nbx = (double)((m.ncols + B.bx - 1) / B.bx), nby = (double)((m.nrows + B.by - 1) / B.by);
    const double rf = (m.ncols * m.D + nbx * (m.W - m.D)) * (m.nrows * m.D + nby * (m.W - m.D)) /
                      ((m.ncols * m.D + m.W - m.D) * (m.nrows * m.D + m.W - m.D));
    return M_LEAF / leaf + M_AMP * a0 + M_COARSE * (a - a0) + M_PASS * sort_bits / 8.0 + M_BLK[b] +
           M_READ * rf;
}

double blocks_of(const ModelTile &m, int b)
{
    const Stage1Cfg &B = BLOCKS[b];
    return m.weight / m.n * (double)((m.ncols + B.bx - 1) / B.bx) *
           (double)((m.nrows + B.by - 1) / B.by);
}

struct Pick {
    int block = 0;
    uint32_t sort_bits = 32;
    uint32_t leaf[6] = {0, 0, 0, 0, 0, 0};
    double ns = 1e300;  // predicted ns for the tiles' points
};

// Window block and key width for one launch over `tiles`, and the leaf size of
// every class among them under that choice.
Pick pick_launch(const std::vector<const ModelTile *> &tiles)
{
    Pick best;
    double pts = 0;
    for (const ModelTile *m : tiles)
        pts += m->weight;
    for (int b = 0; b < NBLOCKS; ++b)
        for (uint32_t sb : SORT_BITS) {
            double blocks = 0;
            for (const ModelTile *m : tiles)
                blocks += blocks_of(*m, b);
            Pick p;
            p.block = b;
            p.sort_bits = sb;
            const double slots = CUS * WG_PER_CU[b];
            blocks = std::max(blocks, 1.0);
            const double q = std::ceil(blocks / slots) * slots / blocks - 1.0;
            p.ns = (M_QUANT * q + M_UNDER * std::max(0.0, slots / blocks - 1.0)) * pts;
            for (int c = 0; c < 6; ++c) {
                double cbest = 1e300;
                bool any = false;
                for (uint32_t leaf : LEAVES) {
                    double ns = 0;
                    for (const ModelTile *m : tiles)
                        if (m->cls == c) {
                            ns += m->weight * tile_cost(*m, leaf, b, sb);
                            any = true;
                        }
                    if (any && ns < cbest) {
                        cbest = ns;
                        p.leaf[c] = leaf;
                    }
                }
                if (any)
                    p.ns += cbest;
            }
            if (p.ns < best.ns)
                best = p;
        }
    return best;
}

owm_status model_table(const owm_tune_sample *samples, uint32_t n, std::vector<Tuner::Entry> &table)
{
    std::vector<ModelTile> tiles(n);
    for (uint32_t k = 0; k < n; ++k)
        if (!model_tile(samples[k], tiles[k]))
            return OWM_E_ARG;
    table = Tuner().table;
    if (n == 0)
        return OWM_OK;
    std::vector<const ModelTile *> all;
    double total = 0;
    for (const ModelTile &m : tiles) {
        all.push_back(&m);
        total += m.weight;
    }
    // the batch: leaf sizes of its classes, block and key width of its class
    const Pick whole = pick_launch(all);
    // a launch of one class's tiles alone: block and key width of that class
    for (int c = 0; c < 6; ++c) {
        std::vector<const ModelTile *> mine;
        for (const ModelTile &m : tiles)
            if (m.cls == c)
                mine.push_back(&m);
        if (mine.empty())
            continue;
        const Pick own = pick_launch(mine);
        table[c].cfg = TuneCfg{whole.leaf[c], BLOCKS[own.block], own.sort_bits};
        table[c].ms = own.ns * 1e-6;
    }
    double ntiles = 0;  // tiles the samples stand for
    for (const ModelTile &m : tiles)
        ntiles += m.weight / m.n;
    const int bc = Tuner::class_of((uint64_t)std::min(total / std::max(ntiles, 1.0), 1.8e19));
    table[bc].cfg.s1 = BLOCKS[whole.block];
    table[bc].cfg.sort_bits = whole.sort_bits;
    table[bc].ms = whole.ns * 1e-6;
    return OWM_OK;
}

}  // namespace

owm_status owm_tune_model(const owm_tune_sample *samples, uint32_t n_samples, owm_tune_entry *out,
                          uint32_t cap, uint32_t *n_out)
{
    if ((n_samples && !samples) || (cap && !out))
        return OWM_E_ARG;
    try {
        std::vector<Tuner::Entry> table;
        owm_status st = model_table(samples, n_samples, table);
        if (st != OWM_OK)
            return st;
        if (n_out)
            *n_out = (uint32_t)table.size();
        if (!out)
            return OWM_OK;
        if (cap < table.size())
            return OWM_E_CAPACITY;
        for (size_t k = 0; k < table.size(); ++k)
            to_entry(table[k], out[k]);
        return OWM_OK;
    } catch (...) {
        return OWM_E_NOMEM;
    }
}

owm_status owm_autotune_stats(owm_ctx *ctx, const double *x, const double *y, const double *z,
                              int device_ptrs, const owm_tile *tiles, uint32_t n_tiles,
                              owm_tune_entry *out, uint32_t cap, uint32_t *n_out)
{
    if (!ctx || !tiles || n_tiles == 0 || !x || !y || !z || (cap && !out))
        return OWM_E_ARG;
    try {
        (void)hipSetDevice(ctx->device);
        if (ctx->tuner.forced) {
            set_err(ctx, "autotune is running on this context");
            return OWM_E_ARG;
        }
        // up to OWM_TUNE_SAMPLES tiles of each class, spread over the batch
        std::vector<uint32_t> members[6];
        double cls_pts[6] = {0, 0, 0, 0, 0, 0};
        for (uint32_t t = 0; t < n_tiles; ++t)
            if (tiles[t].n > 0) {
                const int c = Tuner::class_of(tiles[t].n);
                members[c].push_back(t);
                cls_pts[c] += (double)tiles[t].n;
            }
        std::vector<owm_tune_sample> samples;
        std::vector<std::vector<owm_leaf_bin>> bins;
        samples.reserve(6 * OWM_TUNE_SAMPLES);
        bins.reserve(6 * OWM_TUNE_SAMPLES);
        for (int c = 0; c < 6; ++c) {
            const size_t cnt = members[c].size();
            const size_t k = std::min<size_t>(cnt, OWM_TUNE_SAMPLES);
            const size_t first = samples.size();
            double got = 0;
            for (size_t j = 0; j < k; ++j) {
                const owm_tile &tl = tiles[members[c][(2 * j + 1) * cnt / (2 * k)]];
                owm_qtree_params q;
                memset(&q, 0, sizeof q);
                q.policy = OWM_QT_MAXNUMBER;
                q.max_number = OWM_TUNE_STATS_MAXNUMBER;
                q.level = OWM_TUNE_STATS_LEVEL;
                if (tl.params.flags & OWM_F_HAVE_BBOX) {
                    q.flags = OWM_F_HAVE_BBOX;
                    memcpy(q.bbox, tl.params.bbox, sizeof q.bbox);
                }
                std::vector<owm_leaf_bin> hb(1024);
                owm_tune_sample s;
                memset(&s, 0, sizeof s);
                owm_status st = owm_qtree_stats(ctx, x + tl.offset, y + tl.offset, tl.n, device_ptrs,
                                                &q, hb.data(), hb.size(), &s.summary);
                if (st == OWM_E_CAPACITY) {
                    hb.resize(s.summary.n_bins);
                    st = owm_qtree_stats(ctx, x + tl.offset, y + tl.offset, tl.n, device_ptrs, &q,
                                         hb.data(), hb.size(), &s.summary);
                }
                if (st == OWM_E_ARG)
                    continue;  // more coincident points than a leaf may hold: no sample
                if (st != OWM_OK)
                    return st;
                hb.resize(s.summary.n_bins);
                bins.push_back(std::move(hb));
                s.n = tl.n;
                s.params = tl.params;
                s.bins = bins.back().data();
                s.n_bins = bins.back().size();
                samples.push_back(s);
                got += (double)tl.n;
            }
            for (size_t j = first; j < samples.size(); ++j)  // the class's share of the batch
                samples[j].weight = (double)samples[j].n * cls_pts[c] / got;
        }
        std::vector<Tuner::Entry> table;
        owm_status st = model_table(samples.data(), (uint32_t)samples.size(), table);
        if (st != OWM_OK) {
            set_err(ctx, "statistics of a sampled tile do not fit the tuning model");
            return st;
        }
        ctx->tuner.table = table;
        return put_table(ctx, out, cap, n_out);
    } catch (const std::bad_alloc &) {
        set_err(ctx, "host allocation failed");
        return OWM_E_NOMEM;
    } catch (...) {
        set_err(ctx, "internal error");
        return OWM_E_INTERNAL;
    }
}
