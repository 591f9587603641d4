// Size-class autotuner: times trial configurations of the whole batch chain
// and keeps the fastest per size class.  Knobs: leaf size (16..128, per tile
// class), stage-1 window block (bx x by windows, k points per lane per round,
// with or without leaf cell summaries) and the sort key width (16, 24 or 32
// Morton bits: two to four 8-bit radix passes), the last two per class of the
// batch total.  This is the GPU counterpart of the reference's tuning of its tree
// granularity by sweeps over maxNumber and TBB chunk sizes (CPP-TBB-CPU
// scripts and Results/*.csv); there the sweep is a script, here it is a
// library call whose table the batch and serving paths use.
#include "owm_host.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

using namespace owm;

namespace {

constexpr uint32_t LEAVES[] = {16, 32, 64, 128};
constexpr Stage1Cfg BLOCKS[] = {{8, 8, 4, 0}, {16, 8, 4, 0}, {8, 8, 2, 0}, {16, 8, 4, 1}};
constexpr uint32_t SORT_BITS[] = {32, 24, 16};
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
            (c.sort_bits != 16 && c.sort_bits != 24 && c.sort_bits != 32)) {
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
        const int bcls = Tuner::class_of(np);
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
