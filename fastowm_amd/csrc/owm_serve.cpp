// Serving loop: continuous batching of OWM requests with HBM-budget admission.
//
// owm_serve_start splits the context's device budget into a resident-tree
// pool (a quarter) and `max_inflight` slots.  Each slot is an internal context
// (own stream, own arenas) driven by its own worker thread, so the groups of
// different slots overlap on the device.  A free worker takes the oldest
// pending requests from one shared queue: a keyed request at the head alone
// (its tree is kept resident under its key, so later requests with the same
// key and other parameters only run stages 1-3), else the oldest unkeyed
// requests (past keyed ones waiting behind the head) as one batch of up to
// `max_batch_tiles` tiles whose estimated footprint fits the slot.  A group
// that does not fit its slot is split; a single request that does not fit
// fails with OWM_E_BUDGET.  Every result is computed exactly as owm_run_tile
// would compute it alone (tiles never interact inside a batch).
//
// No counterpart in the reference, which processes one cloud per process
// (application/owm_optim3.cpp:10-272).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <unordered_map>

#include "owm_host.h"

namespace owm {

using clk = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    const double *x = nullptr, *y = nullptr, *z = nullptr;
    uint64_t n = 0;
    owm_params p{};
    uint64_t key = 0;
    bool device = false;  // x, y, z are device pointers (HBM)
    clk::time_point t_submit, t_done;
    int state = 0;  // 0 pending, 1 running, 2 done
    bool cancelled = false;
    owm_status status = OWM_OK;
    std::string err;
    std::vector<uint32_t> seeds;
};

struct Resident {
    owm_tree *tree = nullptr;
    uint64_t bytes = 0, n = 0, last_use = 0;
    int in_use = 0;
};

struct ServeState {
    owm_ctx *owner = nullptr;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> reqs;
    std::vector<owm_ctx *> slots;
    std::vector<std::thread> workers;
    std::map<uint64_t, Resident> trees;
    uint64_t tree_bytes = 0, tree_budget = 0, slot_budget = 0, use_clock = 0, next_id = 1;
    uint32_t max_batch = 64;
    bool stop = false;
};

// Rough device footprint of one tile in a batch (tree, sort scratch, packed
// cloud, upload, query scratch): used only to form groups.
static uint64_t footprint(uint64_t n) { return 420ull * n + (8ull << 20); }

// Host-input bytes a group may carry: a group's chain starts once its whole
// upload is in, so groups of host tiles stay small enough for one slot's
// PCIe upload to overlap the other slots' device work (device-input groups
// grow to max_batch_tiles).
static constexpr uint64_t HOST_GROUP_BYTES = 128ull << 20;

static void finish(ServeState *S, const std::shared_ptr<Request> &r, owm_status st,
                   const char *err)
{
    std::lock_guard<std::mutex> g(S->mu);
    r->status = r->cancelled ? OWM_E_CANCELLED : st;
    if (err && st != OWM_OK)
        r->err = err;
    r->t_done = clk::now();
    r->state = 2;
    S->cv_done.notify_all();
}

// Runs unkeyed requests as one batch on a slot; splits on OWM_E_BUDGET and
// OWM_E_ARG.
static void run_group(ServeState *S, owm_ctx *slot, std::vector<std::shared_ptr<Request>> &g)
{
    uint64_t total = 0;
    for (auto &r : g)
        total += r->n;
    const uint64_t b = (total * 8 + 255) / 256 * 256;
    // device requests that are consecutive slices of one cloud (tiles of a
    // resident survey submitted in order) already form the batch's input:
    // the chain reads them in place
    const double *x0 = nullptr, *y0 = nullptr, *z0 = nullptr;  // the first slice
    const double *ix = nullptr, *iy = nullptr, *iz = nullptr;  // where the next must start
    bool inplace = true;
    for (auto &r : g) {
        if (r->n == 0)
            continue;
        if (!r->device || (ix && (r->x != ix || r->y != iy || r->z != iz))) {
            inplace = false;
            break;
        }
        if (!x0) {
            x0 = r->x;
            y0 = r->y;
            z0 = r->z;
        }
        ix = r->x + r->n;
        iy = r->y + r->n;
        iz = r->z + r->n;
    }
    inplace = inplace && x0 != nullptr;
    owm_status st = OWM_OK;
    if (!inplace && slot->upload.bytes < 3 * b) {
        if (slot->held_bytes() - slot->upload.bytes + 3 * b > slot->budget) {
            set_err(slot, "group input needs %llu bytes over the slot budget",
                    (unsigned long long)(3 * b));
            st = OWM_E_BUDGET;
        } else {
            (void)hipStreamSynchronize(slot->stream);
            hipError_t e = slot->upload.reserve(3 * b, slot->room_for(slot->upload));
            if (e != hipSuccess)
                st = hip_fail(slot, e, "serving upload buffer");
        }
    }
    std::vector<owm_tile> tiles(g.size());
    uint64_t off = 0;
    double *ux = (double *)slot->upload.ptr, *uy = ux + b / 8, *uz = uy + b / 8;
    const double *dx = inplace ? x0 : ux, *dy = inplace ? y0 : uy, *dz = inplace ? z0 : uz;
    for (size_t i = 0; i < g.size() && st == OWM_OK; ++i) {
        tiles[i].offset = off;
        tiles[i].n = g[i]->n;
        tiles[i].params = g[i]->p;
        if (g[i]->n && !inplace) {
            // host inputs cross PCIe; device inputs are copied within HBM
            const hipMemcpyKind kind = g[i]->device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
            hipError_t e = hipMemcpyAsync(ux + off, g[i]->x, g[i]->n * 8, kind, slot->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(uy + off, g[i]->y, g[i]->n * 8, kind, slot->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(uz + off, g[i]->z, g[i]->n * 8, kind, slot->stream);
            if (e != hipSuccess)
                st = hip_fail(slot, e, "serving upload");
        }
        off += g[i]->n;
    }
    // seeds land in the slot's page-locked buffer (reused, never cleared:
    // a fresh zeroed array of one slot per point would cost the host more
    // than the device spends on the group)
    const uint64_t cap = total + g.size() + 1;
    uint32_t *seeds = nullptr;
    if (st == OWM_OK) {
        seeds = (uint32_t *)slot->pinned(slot->h_seeds, slot->h_seeds_cap, cap * 4);
        if (!seeds) {
            set_err(slot, "pinned host allocation failed");
            st = OWM_E_NOMEM;
        }
    }
    std::vector<uint64_t> so(g.size() + 1);
    if (st == OWM_OK)
        st = owm_batch_run_device(slot, dx, dy, dz, tiles.data(), (uint32_t)g.size(), 0, seeds,
                                  cap, so.data());
    // a group over the slot's budget, or holding a request the library
    // rejects (an empty or all-NaN tile, bad parameters), runs its requests
    // one at a time, so that each gets the result it would get alone
    if ((st == OWM_E_BUDGET || st == OWM_E_ARG) && g.size() > 1) {
        for (auto &r : g) {
            std::vector<std::shared_ptr<Request>> one{r};
            run_group(S, slot, one);
        }
        return;
    }
    for (size_t i = 0; i < g.size(); ++i) {
        if (st == OWM_OK)
            g[i]->seeds.assign(seeds + so[i], seeds + so[i + 1]);
        finish(S, g[i], st, slot->err.c_str());
    }
}

// A keyed request: query the resident tree of its key, building and caching
// it first when needed (trees are evicted least-recently-used when the pool
// is full; a tree that does not fit the pool is built, used and dropped).
static void run_keyed(ServeState *S, owm_ctx *slot, const std::shared_ptr<Request> &r)
{
    owm_tree *tree = nullptr;
    bool cached = false;
    {
        std::lock_guard<std::mutex> g(S->mu);
        auto it = S->trees.find(r->key);
        if (it != S->trees.end() && it->second.n == r->n) {
            tree = it->second.tree;
            it->second.in_use++;
            it->second.last_use = ++S->use_clock;
            cached = true;
        }
    }
    owm_status st = OWM_OK;
    if (!tree) {
        st = owm_tree_build(slot, r->x, r->y, r->z, r->n, r->device ? 1 : 0, r->p.leaf_size, &tree);
        if (st == OWM_OK) {
            // hand the tree to the pool (or keep it private when it cannot fit)
            slot->trees_bytes -= tree->mem.bytes;
            tree->ctx = nullptr;
            std::lock_guard<std::mutex> g(S->mu);
            uint64_t need = tree->mem.bytes;
            while (S->tree_bytes + need > S->tree_budget) {
                auto victim = S->trees.end();
                for (auto it = S->trees.begin(); it != S->trees.end(); ++it)
                    if (it->second.in_use == 0 &&
                        (victim == S->trees.end() || it->second.last_use < victim->second.last_use))
                        victim = it;
                if (victim == S->trees.end())
                    break;
                S->tree_bytes -= victim->second.bytes;
                owm_tree_free(victim->second.tree);
                S->trees.erase(victim);
            }
            auto it = S->trees.find(r->key);
            if (S->tree_bytes + need <= S->tree_budget && it == S->trees.end()) {
                Resident R;
                R.tree = tree;
                R.bytes = need;
                R.n = r->n;
                R.in_use = 1;
                R.last_use = ++S->use_clock;
                S->trees[r->key] = R;
                S->tree_bytes += need;
                cached = true;
            }
        }
    }
    uint64_t ns = 0;
    uint32_t *seeds = nullptr;
    if (st == OWM_OK) {
        seeds = (uint32_t *)slot->pinned(slot->h_seeds, slot->h_seeds_cap, (r->n + 1) * 4);
        if (!seeds) {
            set_err(slot, "pinned host allocation failed");
            st = OWM_E_NOMEM;
        }
    }
    if (st == OWM_OK)
        st = owm_tree_run(slot, tree, &r->p, seeds, r->n + 1, &ns);
    if (st == OWM_OK)
        r->seeds.assign(seeds, seeds + ns);
    if (tree) {
        if (cached) {
            std::lock_guard<std::mutex> g(S->mu);
            auto it = S->trees.find(r->key);
            if (it != S->trees.end() && it->second.tree == tree)
                it->second.in_use--;
        } else {
            owm_tree_free(tree);
        }
    }
    finish(S, r, st, slot->err.c_str());
}

static void worker(ServeState *S, owm_ctx *slot)
{
    (void)hipSetDevice(slot->device);
    for (;;) {
        std::vector<std::shared_ptr<Request>> group;
        std::shared_ptr<Request> keyed;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty(); });
            if (S->stop)
                return;
            // the oldest request first: a keyed one runs alone; otherwise the
            // oldest unkeyed requests form a group, taken past any keyed ones
            // that wait behind the head (those keep their place in the queue)
            uint64_t bytes = 0, host_bytes = 0;
            for (auto it = S->queue.begin(); it != S->queue.end();) {
                auto r = *it;
                if (r->cancelled) {
                    it = S->queue.erase(it);
                    r->status = OWM_E_CANCELLED;
                    r->t_done = clk::now();
                    r->state = 2;
                    S->cv_done.notify_all();
                    continue;
                }
                if (r->key) {
                    if (group.empty()) {
                        keyed = r;
                        r->state = 1;
                        S->queue.erase(it);
                        break;
                    }
                    ++it;
                    continue;
                }
                uint64_t f = footprint(r->n);
                const uint64_t h = r->device ? 0 : 24 * r->n;
                if (!group.empty() && (group.size() >= S->max_batch || bytes + f > S->slot_budget ||
                                       host_bytes + h > HOST_GROUP_BYTES))
                    break;
                group.push_back(r);
                bytes += f;
                host_bytes += h;
                r->state = 1;
                it = S->queue.erase(it);
            }
        }
        if (keyed)
            run_keyed(S, slot, keyed);
        else if (!group.empty())
            run_group(S, slot, group);
    }
}

static owm_status copy_result(owm_ctx *ctx, Request &r, uint32_t *seeds, uint64_t cap,
                              uint64_t *n_seeds, double *latency_s)
{
    if (latency_s)
        *latency_s = std::chrono::duration<double>(r.t_done - r.t_submit).count();
    if (n_seeds)
        *n_seeds = r.seeds.size();
    if (r.status != OWM_OK) {
        set_err(ctx, "request %llu: %s", (unsigned long long)r.id, r.err.c_str());
        return r.status;
    }
    if (r.seeds.size() > cap) {
        set_err(ctx, "request %llu: %zu seeds, capacity %llu", (unsigned long long)r.id,
                r.seeds.size(), (unsigned long long)cap);
        return OWM_E_CAPACITY;
    }
    if (!r.seeds.empty())
        memcpy(seeds, r.seeds.data(), r.seeds.size() * 4);
    return OWM_OK;
}

}  // namespace owm

owm_ctx::~owm_ctx()
{
    if (h_tab)
        (void)hipHostFree(h_tab);
    if (h_io)
        (void)hipHostFree(h_io);
    if (h_seeds)
        (void)hipHostFree(h_seeds);
    if (stream)
        (void)hipStreamDestroy(stream);
}

using namespace owm;

extern "C" {

owm_status owm_serve_start(owm_ctx *ctx, uint32_t max_inflight, uint32_t max_batch_tiles)
{
    if (!ctx || max_inflight == 0 || max_inflight > 16)
        return OWM_E_ARG;
    if (ctx->serve) {
        set_err(ctx, "serving already started");
        return OWM_E_ARG;
    }
    try {
        ServeState *S = new ServeState();
        S->owner = ctx;
        S->max_batch = max_batch_tiles ? max_batch_tiles : 64;
        uint64_t held = ctx->held_bytes();
        uint64_t avail = ctx->budget > held ? ctx->budget - held : 0;
        S->tree_budget = avail / 4;
        S->slot_budget = (avail - S->tree_budget) / max_inflight;
        for (uint32_t k = 0; k < max_inflight; ++k) {
            owm_ctx *slot = nullptr;
            owm_status st = owm_ctx_create(ctx->device, S->slot_budget, &slot);
            if (st != OWM_OK) {
                for (auto s : S->slots)
                    owm_ctx_destroy(s);
                delete S;
                set_err(ctx, "cannot create serving slot %u", k);
                return st;
            }
            slot->tuner = ctx->tuner;
            S->slots.push_back(slot);
        }
        ctx->serve = S;
        for (auto slot : S->slots)
            S->workers.emplace_back(worker, S, slot);
        return OWM_OK;
    } catch (const std::exception &e) {
        set_err(ctx, "serve start: %s", e.what());
        return OWM_E_INTERNAL;
    }
}

static owm_status submit_impl(owm_ctx *ctx, const double *x, const double *y, const double *z,
                              uint64_t n, const owm_params *params, uint64_t tile_key,
                              uint64_t *req_id, bool device)
{
    if (!ctx || !params || !req_id || (n && (!x || !y || !z)))
        return OWM_E_ARG;
    ServeState *S = ctx->serve;
    if (!S) {
        set_err(ctx, "serving not started");
        return OWM_E_ARG;
    }
    try {
        auto r = std::make_shared<Request>();
        r->x = x;
        r->y = y;
        r->z = z;
        r->n = n;
        r->p = *params;
        r->key = tile_key;
        r->device = device;
        r->t_submit = clk::now();
        {
            std::lock_guard<std::mutex> g(S->mu);
            r->id = S->next_id++;
            S->reqs[r->id] = r;
            S->queue.push_back(r);
        }
        S->cv_work.notify_one();
        *req_id = r->id;
        return OWM_OK;
    } catch (const std::exception &e) {
        set_err(ctx, "serve submit: %s", e.what());
        return OWM_E_INTERNAL;
    }
}

owm_status owm_serve_submit(owm_ctx *ctx, const double *x, const double *y, const double *z,
                            uint64_t n, const owm_params *params, uint64_t tile_key,
                            uint64_t *req_id)
{
    return submit_impl(ctx, x, y, z, n, params, tile_key, req_id, false);
}

owm_status owm_serve_submit_device(owm_ctx *ctx, const double *d_x, const double *d_y,
                                   const double *d_z, uint64_t n, const owm_params *params,
                                   uint64_t tile_key, uint64_t *req_id)
{
    return submit_impl(ctx, d_x, d_y, d_z, n, params, tile_key, req_id, true);
}

static owm_status poll_impl(owm_ctx *ctx, uint64_t req_id, uint32_t *seeds, uint64_t cap,
                            uint64_t *n_seeds, double *latency_s, bool wait)
{
    if (!ctx)
        return OWM_E_ARG;
    ServeState *S = ctx->serve;
    if (!S)
        return OWM_E_NOTFOUND;
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->reqs.find(req_id);
        if (it == S->reqs.end())
            return OWM_E_NOTFOUND;
        r = it->second;
        if (wait)
            S->cv_done.wait(lk, [&] { return r->state == 2; });
        if (r->state != 2)
            return OWM_E_NOTREADY;
    }
    owm_status st = copy_result(ctx, *r, seeds, cap, n_seeds, latency_s);
    if (st != OWM_E_CAPACITY) {
        std::lock_guard<std::mutex> g(S->mu);
        S->reqs.erase(req_id);
    }
    return st;
}

owm_status owm_serve_poll(owm_ctx *ctx, uint64_t req_id, uint32_t *seeds, uint64_t cap,
                          uint64_t *n_seeds, double *latency_s)
{
    return poll_impl(ctx, req_id, seeds, cap, n_seeds, latency_s, false);
}

owm_status owm_serve_wait(owm_ctx *ctx, uint64_t req_id, uint32_t *seeds, uint64_t cap,
                          uint64_t *n_seeds, double *latency_s)
{
    return poll_impl(ctx, req_id, seeds, cap, n_seeds, latency_s, true);
}

owm_status owm_serve_cancel(owm_ctx *ctx, uint64_t req_id)
{
    if (!ctx)
        return OWM_E_ARG;
    ServeState *S = ctx->serve;
    if (!S)
        return OWM_E_NOTFOUND;
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->reqs.find(req_id);
    if (it == S->reqs.end())
        return OWM_E_NOTFOUND;
    it->second->cancelled = true;
    return OWM_OK;
}

owm_status owm_serve_evict(owm_ctx *ctx, uint64_t tile_key)
{
    if (!ctx)
        return OWM_E_ARG;
    ServeState *S = ctx->serve;
    if (!S)
        return OWM_E_NOTFOUND;
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->trees.find(tile_key);
    if (it == S->trees.end())
        return OWM_E_NOTFOUND;
    if (it->second.in_use) {
        set_err(ctx, "tree of key %llu is in use", (unsigned long long)tile_key);
        return OWM_E_NOTREADY;
    }
    S->tree_bytes -= it->second.bytes;
    owm_tree_free(it->second.tree);
    S->trees.erase(it);
    return OWM_OK;
}

owm_status owm_serve_stop(owm_ctx *ctx)
{
    if (!ctx || !ctx->serve)
        return OWM_OK;
    ServeState *S = ctx->serve;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
        for (auto &r : S->queue) {
            r->status = OWM_E_CANCELLED;
            r->t_done = clk::now();
            r->state = 2;
        }
        S->queue.clear();
    }
    S->cv_work.notify_all();
    for (auto &t : S->workers)
        t.join();
    S->cv_done.notify_all();
    for (auto slot : S->slots)
        (void)hipStreamSynchronize(slot->stream);
    for (auto &kv : S->trees)
        owm_tree_free(kv.second.tree);
    for (auto slot : S->slots)
        owm_ctx_destroy(slot);
    ctx->serve = nullptr;
    delete S;
    return OWM_OK;
}

}  // extern "C"
