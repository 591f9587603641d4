// C ABI entry points (include/owm.h).  Each catches every C++ exception and
// returns a status; none exits the process.
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "owm_host.h"

using namespace owm;

#define GUARD_BEGIN try {
#define GUARD_END(ctx)                                              \
    }                                                               \
    catch (const std::bad_alloc &) {                                \
        set_err(ctx, "host allocation failed");                     \
        return OWM_E_NOMEM;                                         \
    }                                                               \
    catch (const std::exception &e) {                               \
        set_err(ctx, "internal error: %s", e.what());               \
        return OWM_E_INTERNAL;                                      \
    }

static inline uint64_t al256(uint64_t v) { return (v + 255) / 256 * 256; }

// Device scratch of owm_run_tile_device_noalloc: a file-scope array of the
// library's code object, of which every device holds its own copy from the
// time the library's code is loaded on it until the process ends.  Its
// host-side name is not a device address: the address hipGetSymbolAddress
// gives on each device (resolved once per device, with that device current,
// when the first context is created on it) is what host code passes.
static constexpr uint64_t NOALLOC_SCRATCH_BYTES = 256ull << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];
static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static uint8_t *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static hipError_t g_scratch_err[NOALLOC_MAX_DEVICES];
static bool g_scratch_done[NOALLOC_MAX_DEVICES];

// With `device` current.
static void resolve_scratch(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device])
        return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? (uint8_t *)p : nullptr;
    if (e != hipSuccess)
        (void)hipGetLastError();
    g_scratch_done[device] = true;
}

extern "C" {

const char *owm_status_str(owm_status s)
{
    switch (s) {
    case OWM_OK: return "ok";
    case OWM_E_ARG: return "invalid argument";
    case OWM_E_NOMEM: return "out of memory";
    case OWM_E_BUDGET: return "over the HBM budget";
    case OWM_E_DEVICE: return "device error";
    case OWM_E_CAPACITY: return "output buffer too small";
    case OWM_E_CANCELLED: return "cancelled";
    case OWM_E_NOTREADY: return "not ready";
    case OWM_E_NOTFOUND: return "not found";
    case OWM_E_INTERNAL: return "internal error";
    }
    return "unknown status";
}

const char *owm_last_error(const owm_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

owm_status owm_ctx_create(int device, uint64_t hbm_budget_bytes, owm_ctx **out)
{
    if (!out)
        return OWM_E_ARG;
    *out = nullptr;
    GUARD_BEGIN
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return OWM_E_DEVICE;
    if (hipSetDevice(device) != hipSuccess)
        return OWM_E_DEVICE;
    owm_ctx *c = new owm_ctx();
    c->device = device;
    resolve_scratch(device);
    // the planning state of owm_run_tile_device_noalloc, sized for one tile
    // now so that those calls allocate nothing (not even host memory)
    {
        owm::Prepared &F = c->fixed_plan;
        F.td.reserve(1);
        F.n.reserve(1);
        F.leaf.reserve(1);
        F.box.reserve(1);
        for (auto *v : {&F.L.pt_off, &F.L.leaf_off, &F.L.node_off, &F.L.grp_off, &F.L.sort_units,
                        &F.Q.win_off, &F.Q.cell_off, &F.Q.blk_off})
            v->reserve(2);
        c->err.reserve(1024);
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return OWM_E_DEVICE;
    }
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) {
        delete c;
        return OWM_E_DEVICE;
    }
    c->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(freeb * 0.9);
    *out = c;
    return OWM_OK;
    GUARD_END(nullptr)
}

void owm_ctx_destroy(owm_ctx *ctx)
{
    if (!ctx)
        return;
    owm_serve_stop(ctx);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
}

uint64_t owm_ctx_device_bytes(const owm_ctx *ctx) { return ctx ? ctx->held_bytes() : 0; }

owm_status owm_host_alloc(uint64_t bytes, void **out)
{
    if (!out)
        return OWM_E_ARG;
    *out = nullptr;
    if (bytes == 0)
        return OWM_OK;
    return hipHostMalloc(out, bytes, 0) == hipSuccess ? OWM_OK : OWM_E_NOMEM;
}

void owm_host_free(void *ptr)
{
    if (ptr)
        (void)hipHostFree(ptr);
}

owm_status owm_derive(const owm_params *params, uint64_t n, const double *bbox_or_null,
                      owm_derived *out)
{
    if (!params || !out)
        return OWM_E_ARG;
    const double *b = (params->flags & OWM_F_HAVE_BBOX) ? params->bbox : bbox_or_null;
    if (!b)
        return OWM_E_ARG;
    Derived D;
    std::string why;
    owm_status st = derive_params(*params, n, b, D, why);
    if (st != OWM_OK)
        return st;
    out->width = D.width;
    out->height = D.height;
    out->density = D.density;
    out->displace = D.displace;
    out->min_num_points = D.min_num_points;
    out->ncols = D.ncols;
    out->nrows = D.nrows;
    out->ncolsg = D.ncolsg;
    out->nrowsg = D.nrowsg;
    for (int k = 0; k < 4; ++k)
        out->bbox[k] = b[k];
    return OWM_OK;
}

owm_status owm_query_seed_count(const owm_params *params, uint64_t n, const double *bbox_or_null,
                                uint64_t *max_seeds)
{
    if (!params || !max_seeds)
        return OWM_E_ARG;
    const double *b = (params->flags & OWM_F_HAVE_BBOX) ? params->bbox : bbox_or_null;
    if (!b) {
        *max_seeds = n;
        return OWM_OK;
    }
    Derived D;
    std::string why;
    owm_status st = derive_params(*params, n, b, D, why);
    if (st != OWM_OK)
        return st;
    const uint64_t w = (uint64_t)D.ncols * D.nrows + (uint64_t)D.ncolsg * D.nrowsg;
    *max_seeds = w < n ? w : n;
    return OWM_OK;
}

}  // extern "C"

namespace owm {
owm_status upload_cloud(owm_ctx *ctx, const double *x, const double *y, const double *z,
                               uint64_t n, double **dx, double **dy, double **dz)
{
    uint64_t b = al256(n * 8ull);
    if (ctx->upload.bytes < 3 * b) {
        if (ctx->budget && ctx->held_bytes() - ctx->upload.bytes + 3 * b > ctx->budget) {
            set_err(ctx, "input upload needs %llu bytes over the budget", (unsigned long long)(3 * b));
            return OWM_E_BUDGET;
        }
        (void)hipStreamSynchronize(ctx->stream);
        hipError_t e = ctx->upload.reserve(3 * b, ctx->room_for(ctx->upload));
        if (e != hipSuccess)
            return hip_fail(ctx, e, "upload buffer");
    }
    uint8_t *base = (uint8_t *)ctx->upload.ptr;
    *dx = (double *)base;
    *dy = (double *)(base + b);
    *dz = (double *)(base + 2 * b);
    hipError_t e;
    if (n) {
        if ((e = hipMemcpyAsync(*dx, x, n * 8, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
            (e = hipMemcpyAsync(*dy, y, n * 8, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
            (e = hipMemcpyAsync(*dz, z, n * 8, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
            return hip_fail(ctx, e, "cloud H2D");
    }
    return OWM_OK;
}
}  // namespace owm

extern "C" {

static owm_status run_one(owm_ctx *ctx, const double *dx, const double *dy, const double *dz,
                          uint64_t n, const owm_params *params, uint32_t *seeds, uint64_t cap,
                          uint64_t *n_seeds)
{
    owm_tile t;
    t.offset = 0;
    t.n = n;
    t.params = *params;
    uint64_t off[2] = {0, 0};
    owm_status st = run_batch(ctx, dx, dy, dz, &t, 1, B_AUTO_GRAPH, seeds, cap, off);
    if (n_seeds)
        *n_seeds = off[1];
    return st;
}

owm_status owm_run_tile(owm_ctx *ctx, const double *x, const double *y, const double *z, uint64_t n,
                        const owm_params *params, uint32_t *seeds, uint64_t cap, uint64_t *n_seeds)
{
    if (!ctx || !params || !n_seeds || (n && (!x || !y || !z)) || (cap && !seeds))
        return OWM_E_ARG;
    GUARD_BEGIN
    (void)hipSetDevice(ctx->device);
    double *dx, *dy, *dz;
    owm_status st = upload_cloud(ctx, x, y, z, n, &dx, &dy, &dz);
    if (st != OWM_OK)
        return st;
    return run_one(ctx, dx, dy, dz, n, params, seeds, cap, n_seeds);
    GUARD_END(ctx)
}

owm_status owm_run_tile_device(owm_ctx *ctx, const double *d_x, const double *d_y,
                               const double *d_z, uint64_t n, const owm_params *params,
                               uint32_t *seeds, uint64_t cap, uint64_t *n_seeds)
{
    if (!ctx || !params || !n_seeds || (n && (!d_x || !d_y || !d_z)) || (cap && !seeds))
        return OWM_E_ARG;
    GUARD_BEGIN
    (void)hipSetDevice(ctx->device);
    return run_one(ctx, d_x, d_y, d_z, n, params, seeds, cap, n_seeds);
    GUARD_END(ctx)
}

owm_status owm_run_tile_device_noalloc(owm_ctx *ctx, const double *d_x, const double *d_y,
                                       const double *d_z, uint64_t n, const owm_params *params,
                                       uint32_t *seeds, uint64_t cap, uint64_t *n_seeds)
{
    if (!ctx || !params || !n_seeds || (n && (!d_x || !d_y || !d_z)) || (cap && !seeds))
        return OWM_E_ARG;
    GUARD_BEGIN
    (void)hipSetDevice(ctx->device);
    const int dev = ctx->device;
    uint8_t *base = dev >= 0 && dev < NOALLOC_MAX_DEVICES ? g_scratch_ptr[dev] : nullptr;
    if (!base) {
        set_err(ctx, "no static scratch on device %d: %s", dev,
                dev >= 0 && dev < NOALLOC_MAX_DEVICES ? hipGetErrorString(g_scratch_err[dev])
                                                      : "device index too large");
        return OWM_E_DEVICE;
    }
    // host staging of one tile's readbacks and table, on this call's stack
    alignas(256) uint8_t h_io[2048];
    alignas(256) uint8_t h_tab[4096];
    if (io_bytes(1) > sizeof h_io || table_bytes(1) > sizeof h_tab) {
        set_err(ctx, "host staging of one tile does not fit");
        return OWM_E_INTERNAL;
    }
    uint64_t need = 0;
    return run_tile_fixed(ctx, d_x, d_y, d_z, n, *params, seeds, cap, n_seeds, base,
                          NOALLOC_SCRATCH_BYTES, h_io, h_tab, &need);
    GUARD_END(ctx)
}

owm_status owm_batch_run_device(owm_ctx *ctx, const double *d_x, const double *d_y,
                                const double *d_z, const owm_tile *tiles, uint32_t n_tiles,
                                uint32_t flags, uint32_t *seeds, uint64_t cap, uint64_t *seed_off)
{
    if (!ctx || !tiles || !seed_off || (cap && !seeds))
        return OWM_E_ARG;
    GUARD_BEGIN
    (void)hipSetDevice(ctx->device);
    return run_batch(ctx, d_x, d_y, d_z, tiles, n_tiles, flags, seeds, cap, seed_off);
    GUARD_END(ctx)
}

owm_status owm_batch_run(owm_ctx *ctx, const double *x, const double *y, const double *z,
                         const owm_tile *tiles, uint32_t n_tiles, uint32_t flags, uint32_t *seeds,
                         uint64_t cap, uint64_t *seed_off)
{
    if (!ctx || !tiles || !seed_off || (cap && !seeds) || n_tiles == 0)
        return OWM_E_ARG;
    GUARD_BEGIN
    (void)hipSetDevice(ctx->device);
    uint64_t n = tiles[n_tiles - 1].offset + tiles[n_tiles - 1].n;
    double *dx, *dy, *dz;
    owm_status st = upload_cloud(ctx, x, y, z, n, &dx, &dy, &dz);
    if (st != OWM_OK)
        return st;
    return run_batch(ctx, dx, dy, dz, tiles, n_tiles, flags, seeds, cap, seed_off);
    GUARD_END(ctx)
}

// ------------------------------------------------------------------ trees --

owm_status owm_tree_build(owm_ctx *ctx, const double *x, const double *y, const double *z,
                          uint64_t n, int device_ptrs, uint32_t leaf_size, owm_tree **out)
{
    if (!ctx || !out || n == 0 || !x || !y || !z || leaf_size > 1024)
        return OWM_E_ARG;
    *out = nullptr;
    GUARD_BEGIN
    (void)hipSetDevice(ctx->device);
    const double *dx = x, *dy = y, *dz = z;
    if (!device_ptrs) {
        double *ux, *uy, *uz;
        owm_status st = upload_cloud(ctx, x, y, z, n, &ux, &uy, &uz);
        if (st != OWM_OK)
            return st;
        dx = ux;
        dy = uy;
        dz = uz;
    }
    owm_tile tile{};
    tile.offset = 0;
    tile.n = n;
    tile.params.leaf_size = leaf_size;
    std::unique_ptr<owm_tree> tr(new owm_tree());
    tr->ctx = ctx;
    tr->n = n;
    owm_status st = prepare_tree(ctx, dx, dy, &tile, 1, tr->P);
    if (st != OWM_OK)
        return st;
    tr->P.Q.ntiles = 1;
    tr->P.Q.win_off.assign(2, 0);
    tr->P.Q.cell_off.assign(2, 0);
    tr->P.Q.blk_off.assign(2, 0);
    const uint64_t tb = tree_bytes(tr->P.L);
    if (ctx->budget && ctx->held_bytes() + tb > ctx->budget) {
        set_err(ctx, "tree needs %llu device bytes over the budget", (unsigned long long)tb);
        return OWM_E_BUDGET;
    }
    hipError_t e = tr->mem.reserve(tb, tb);
    if (e != hipSuccess)
        return hip_fail(ctx, e, "tree allocation");
    ctx->trees_bytes += tr->mem.bytes;
    carve_tree((uint8_t *)tr->mem.ptr, tr->P.L, tr->T);
    st = ctx->arena.ensure(ctx, table_bytes(1) + build_scratch_bytes(tr->P.L) + 8192);
    if (st != OWM_OK) {
        ctx->trees_bytes -= tr->mem.bytes;
        return st;
    }
    uint32_t *guard = (uint32_t *)ctx->arena.take(256);
    launch_fill_u32(ctx->stream, guard, 64, 0u);
    BatchDev B;
    st = upload_table(ctx, ctx->arena, tr->P, B);
    if (st == OWM_OK)
        st = build_tree(ctx, ctx->stream, tr->P.L, B, dx, dy, dz, tr->T, ctx->arena, guard);
    uint32_t hg = 0;
    if (st == OWM_OK) {
        e = hipMemcpyAsync(&hg, guard, 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx->stream);
        ctx->prof.collect();
        if (e != hipSuccess)
            st = hip_fail(ctx, e, "tree build");
        else if (hg) {
            set_err(ctx, "tree build guard tripped (flags %u)", hg);
            st = OWM_E_INTERNAL;
        }
    }
    if (st != OWM_OK) {
        ctx->trees_bytes -= tr->mem.bytes;
        return st;
    }
    *out = tr.release();
    return OWM_OK;
    GUARD_END(ctx)
}

void owm_tree_free(owm_tree *tree)
{
    if (!tree)
        return;
    if (tree->ctx) {
        (void)hipStreamSynchronize(tree->ctx->stream);
        tree->ctx->trees_bytes -= tree->mem.bytes;
    }
    delete tree;
}

uint64_t owm_tree_points(const owm_tree *tree) { return tree ? tree->n : 0; }

owm_status owm_tree_info(const owm_tree *tree, uint64_t *n_leaves, uint64_t *n_internal,
                         uint64_t *root_count, double *root_bbox)
{
    if (!tree)
        return OWM_E_ARG;
    owm_ctx *ctx = tree->ctx;
    GUARD_BEGIN
    if (n_leaves)
        *n_leaves = tree->P.L.nleaf;
    if (n_internal)
        *n_internal = tree->P.L.nnode;
    uint32_t root = tree->P.td[0].root;
    double4 box = make_double4(0, 0, 0, 0);
    uint32_t cnt = 0;
    if (root != OWM_NONE) {
        const void *bsrc, *csrc;
        if (root & OWM_LEAF_BIT) {
            bsrc = tree->T.lf_box + (root & ~OWM_LEAF_BIT);
            csrc = tree->T.lf_cnt + (root & ~OWM_LEAF_BIT);
        } else {
            bsrc = tree->T.nd_box + root;
            csrc = tree->T.nd_cnt + root;
        }
        hipError_t e = hipMemcpy(&box, bsrc, sizeof box, hipMemcpyDeviceToHost);
        if (e == hipSuccess)
            e = hipMemcpy(&cnt, csrc, 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess)
            return hip_fail(ctx, e, "tree info");
    }
    if (root_count)
        *root_count = cnt;
    if (root_bbox) {
        root_bbox[0] = box.x;
        root_bbox[1] = box.z;
        root_bbox[2] = box.y;
        root_bbox[3] = box.w;
    }
    return OWM_OK;
    GUARD_END(ctx)
}

owm_status owm_tree_read(const owm_tree *tree, owm_tree_view *v)
{
    if (!tree || !v || !tree->ctx)
        return OWM_E_ARG;
    owm_ctx *ctx = tree->ctx;
    GUARD_BEGIN
    (void)hipSetDevice(ctx->device);
    const TileDev &TD = tree->P.td[0];
    const uint64_t n = tree->P.L.np, nl = tree->P.L.nleaf, nn = tree->P.L.nnode;
    v->n = n;
    v->n_leaves = nl;
    v->n_internal = nn;
    v->leaf_size = TD.leaf_size;
    v->key_bits = (uint32_t)tree->P.L.key_bits;
    v->root = TD.root;
    v->qmin[0] = TD.qminx;
    v->qmin[1] = TD.qminy;
    v->qscale = TD.qscale;
    HIP_TRY(hipStreamSynchronize(ctx->stream), "tree read");
    auto get = [&](void *dst, const void *src, uint64_t bytes) -> owm_status {
        if (!dst || !bytes)
            return OWM_OK;
        HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "tree read");
        return OWM_OK;
    };
    // boxes are stored (minx, miny, maxx, maxy); the view uses header order
    auto boxes = [&](double *dst, const double4 *src, uint64_t cnt) -> owm_status {
        if (!dst || !cnt)
            return OWM_OK;
        std::vector<double4> tmp(cnt);
        owm_status st = get(tmp.data(), src, cnt * sizeof(double4));
        for (uint64_t i = 0; st == OWM_OK && i < cnt; ++i) {
            dst[4 * i + 0] = tmp[i].x;
            dst[4 * i + 1] = tmp[i].z;
            dst[4 * i + 2] = tmp[i].y;
            dst[4 * i + 3] = tmp[i].w;
        }
        return st;
    };
    const TreeOut &T = tree->T;
    owm_status st = OWM_OK;
    const struct {
        void *dst;
        const void *src;
        uint64_t bytes;
    } copies[] = {
        {v->leaf_key, T.lf_key0, nl * 4}, {v->leaf_start, T.lf_start, nl * 4},
        {v->leaf_points, T.lf_npts, nl * 4}, {v->leaf_count, T.lf_cnt, nl * 4},
        {v->leaf_min, T.lf_zid, nl * 4}, {v->node_count, T.nd_cnt, nn * 4},
        {v->node_min, T.nd_zid, nn * 4}, {v->node_child, T.nd_child, nn * 8},
    };
    for (const auto &c : copies)
        if (st == OWM_OK)
            st = get(c.dst, c.src, c.bytes);
    // the points are (x, y) pairs and (z, id, key) triples on the device
    if (st == OWM_OK && n && (v->xs || v->ys)) {
        std::vector<double2> xy(n);
        st = get(xy.data(), T.pts.xy, n * 16);
        for (uint64_t i = 0; st == OWM_OK && i < n; ++i) {
            if (v->xs)
                v->xs[i] = xy[i].x;
            if (v->ys)
                v->ys[i] = xy[i].y;
        }
    }
    if (st == OWM_OK && n && (v->ids || v->zs)) {
        std::vector<PointZ> zi(n);
        st = get(zi.data(), T.pts.zi, n * 16);
        for (uint64_t i = 0; st == OWM_OK && i < n; ++i) {
            if (v->ids)
                v->ids[i] = zi[i].id;
            if (v->zs)
                v->zs[i] = zi[i].z;
        }
    }
    if (st == OWM_OK)
        st = boxes(v->leaf_bbox, T.lf_box, nl);
    if (st == OWM_OK)
        st = boxes(v->node_bbox, T.nd_box, nn);
    return st;
    GUARD_END(ctx)
}

// Stages from..to on a resident tree.
static owm_status tree_query(owm_ctx *ctx, owm_tree *tree, const owm_params *params, int from,
                             int to, const uint32_t *h_in, uint64_t n_in, Prepared &P, QueryOut &qo,
                             uint64_t s1_cap = 0)
{
    P = tree->P;
    owm_status st = prepare_query(ctx, params, 0, P.box, P.L.np, P);
    if (st != OWM_OK)
        return st;
    if (s1_cap)
        P.Q.s1_cap = s1_cap;
    st = ensure_s1_list(ctx, P.Q.s1_cap);
    if (st != OWM_OK)
        return st;
    st = ctx->arena.ensure(ctx, table_bytes(1) + query_bytes(P.Q, P.L.np, n_in) + 4096);
    if (st != OWM_OK)
        return st;
    BatchDev B;
    st = upload_table(ctx, ctx->arena, P, B);
    if (st != OWM_OK)
        return st;
    return enqueue_query(ctx, ctx->stream, B, tree_view(tree->T), P.Q, P.L.np, P.cfg,
                         P.any_need1 ? 1u : 0u, ctx->arena, qo, from, to, h_in, n_in);
}

static owm_status fetch_counts(owm_ctx *ctx, const QueryOut &qo, uint32_t *hc, uint64_t *s1_needed)
{
    hipError_t e = hipMemcpyAsync(hc, qo.cnt, 16 * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    ctx->prof.collect();
    if (e != hipSuccess)
        return hip_fail(ctx, e, "counts D2H");
    return check_guards(ctx, hc, s1_needed);
}

// tree_query + counts, re-running once with the exact stage-1 join size when
// the default list was too small.
static owm_status tree_query_sync(owm_ctx *ctx, owm_tree *tree, const owm_params *params, int from,
                                  int to, const uint32_t *h_in, uint64_t n_in, Prepared &P,
                                  QueryOut &qo, uint32_t *hc)
{
    uint64_t need = 0;
    owm_status st = tree_query(ctx, tree, params, from, to, h_in, n_in, P, qo);
    if (st == OWM_OK)
        st = fetch_counts(ctx, qo, hc, &need);
    if (st == OWM_OK && need) {
        st = tree_query(ctx, tree, params, from, to, h_in, n_in, P, qo, need + need / 8 + 1024);
        if (st == OWM_OK)
            st = fetch_counts(ctx, qo, hc, &need);
        if (st == OWM_OK && need) {
            set_err(ctx, "stage-1 join list overflow persisted");
            st = OWM_E_INTERNAL;
        }
    }
    return st;
}

owm_status owm_stage1(owm_ctx *ctx, owm_tree *tree, const owm_params *params, uint32_t *min_ids,
                      uint64_t cap, owm_derived *derived)
{
    if (!ctx || !tree || !params || (cap && !min_ids))
        return OWM_E_ARG;
    GUARD_BEGIN
    (void)hipSetDevice(ctx->device);
    Prepared P;
    QueryOut qo;
    uint32_t hc[16];
    owm_status st = tree_query_sync(ctx, tree, params, 1, 1, nullptr, 0, P, qo, hc);
    if (derived && (st == OWM_OK || st == OWM_E_CAPACITY))
        owm_derive(params, tree->n, P.box[0].data(), derived);
    if (st != OWM_OK)
        return st;
    if (P.Q.nwin > cap) {
        set_err(ctx, "min_ids holds %llu, %llu windows", (unsigned long long)cap,
                (unsigned long long)P.Q.nwin);
        return OWM_E_CAPACITY;
    }
    hipError_t e = hipMemcpy(min_ids, qo.min_ids, P.Q.nwin * 4ull, hipMemcpyDeviceToHost);
    if (e != hipSuccess)
        return hip_fail(ctx, e, "min_ids D2H");
    return OWM_OK;
    GUARD_END(ctx)
}

owm_status owm_stage2(owm_ctx *ctx, owm_tree *tree, const owm_params *params,
                      const uint32_t *min_ids, uint64_t n_ids, uint32_t *llp, uint64_t cap,
                      uint64_t *n_llp)
{
    if (!ctx || !tree || !params || !n_llp || (n_ids && !min_ids) || (cap && !llp) ||
        n_ids >= (1ull << 31))
        return OWM_E_ARG;
    GUARD_BEGIN
    (void)hipSetDevice(ctx->device);
    for (uint64_t i = 0; i < n_ids; ++i)
        if (min_ids[i] != OWM_NONE && min_ids[i] >= tree->n) {
            set_err(ctx, "min_ids[%llu] = %u is not a point index", (unsigned long long)i, min_ids[i]);
            return OWM_E_ARG;
        }
    Prepared P;
    QueryOut qo;
    uint32_t hc[16];
    owm_status st = tree_query_sync(ctx, tree, params, 2, 2, min_ids, n_ids, P, qo, hc);
    if (st != OWM_OK)
        return st;
    *n_llp = hc[0];
    if (hc[0] > cap) {
        set_err(ctx, "llp holds %llu, %u needed", (unsigned long long)cap, hc[0]);
        return OWM_E_CAPACITY;
    }
    if (hc[0]) {
        hipError_t e = hipMemcpy(llp, qo.llp, hc[0] * 4ull, hipMemcpyDeviceToHost);
        if (e != hipSuccess)
            return hip_fail(ctx, e, "llp D2H");
    }
    return OWM_OK;
    GUARD_END(ctx)
}

owm_status owm_stage3(owm_ctx *ctx, owm_tree *tree, const owm_params *params, const uint32_t *llp,
                      uint64_t n_llp, uint32_t *adds, uint64_t cap, uint64_t *n_adds)
{
    if (!ctx || !tree || !params || !n_adds || (n_llp && !llp) || (cap && !adds) ||
        n_llp >= (1ull << 31))
        return OWM_E_ARG;
    GUARD_BEGIN
    (void)hipSetDevice(ctx->device);
    for (uint64_t i = 0; i < n_llp; ++i)
        if (llp[i] >= tree->n) {
            set_err(ctx, "llp[%llu] = %u is not a point index", (unsigned long long)i, llp[i]);
            return OWM_E_ARG;
        }
    Prepared P;
    QueryOut qo;
    uint32_t hc[16];
    owm_status st = tree_query_sync(ctx, tree, params, 3, 3, llp, n_llp, P, qo, hc);
    if (st != OWM_OK)
        return st;
    *n_adds = hc[1];
    if (hc[1] > cap) {
        set_err(ctx, "adds holds %llu, %u needed", (unsigned long long)cap, hc[1]);
        return OWM_E_CAPACITY;
    }
    if (hc[1]) {
        hipError_t e = hipMemcpy(adds, qo.llp + hc[0], hc[1] * 4ull, hipMemcpyDeviceToHost);
        if (e != hipSuccess)
            return hip_fail(ctx, e, "adds D2H");
    }
    return OWM_OK;
    GUARD_END(ctx)
}

owm_status owm_tree_run(owm_ctx *ctx, owm_tree *tree, const owm_params *params, uint32_t *seeds,
                        uint64_t cap, uint64_t *n_seeds)
{
    if (!ctx || !tree || !params || !n_seeds || (cap && !seeds))
        return OWM_E_ARG;
    GUARD_BEGIN
    (void)hipSetDevice(ctx->device);
    Prepared P;
    QueryOut qo;
    uint64_t off[2] = {0, 0};
    uint64_t need = 0;
    owm_status st = tree_query(ctx, tree, params, 1, 4, nullptr, 0, P, qo);
    if (st == OWM_OK)
        st = read_seeds(ctx, qo, P, seeds, cap, off, &need);
    if (st == OWM_OK && need) {
        st = tree_query(ctx, tree, params, 1, 4, nullptr, 0, P, qo, need + need / 8 + 1024);
        if (st == OWM_OK)
            st = read_seeds(ctx, qo, P, seeds, cap, off, &need);
        if (st == OWM_OK && need) {
            set_err(ctx, "stage-1 join list overflow persisted");
            st = OWM_E_INTERNAL;
        }
    }
    *n_seeds = off[1];
    return st;
    GUARD_END(ctx)
}

owm_status owm_profile_enable(owm_ctx *ctx, int enable)
{
    if (!ctx)
        return OWM_E_ARG;
    GUARD_BEGIN
    (void)hipStreamSynchronize(ctx->stream);
    ctx->prof.collect();
    ctx->prof.reset();
    ctx->prof.on = enable != 0;
    return OWM_OK;
    GUARD_END(ctx)
}

owm_status owm_profile_phases(owm_ctx *ctx, uint32_t mask)
{
    if (!ctx)
        return OWM_E_ARG;
    GUARD_BEGIN
    (void)hipStreamSynchronize(ctx->stream);
    ctx->prof.collect();
    ctx->prof.mask = mask ? mask : ~0u;
    return OWM_OK;
    GUARD_END(ctx)
}

owm_status owm_profile_read(owm_ctx *ctx, owm_phase_stat *out, uint32_t cap, uint32_t *n)
{
    if (!ctx || !n || (cap && !out))
        return OWM_E_ARG;
    GUARD_BEGIN
    (void)hipStreamSynchronize(ctx->stream);
    ctx->prof.collect();
    *n = PH_COUNT;
    if (cap < PH_COUNT)
        return OWM_E_CAPACITY;
    for (int k = 0; k < PH_COUNT; ++k) {
        memset(out[k].name, 0, sizeof out[k].name);
        strncpy(out[k].name, phase_name(k), sizeof out[k].name - 1);
        out[k].ms = ctx->prof.ms[k];
        out[k].calls = ctx->prof.calls[k];
        out[k].bytes = ctx->prof.bytes[k];
        out[k].items = ctx->prof.items[k];
    }
    return OWM_OK;
    GUARD_END(ctx)
}

}  // extern "C"
