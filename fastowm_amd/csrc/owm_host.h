// Host-side state of the OWM library (C++; never crosses the C ABI).
#pragma once
#include <array>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/owm.h"
#include "owm_kernels.h"

struct owm_ctx;

namespace owm {

struct DevBuf {
    void *ptr = nullptr;
    uint64_t bytes = 0;
    bool borrowed = false;  // memory owned elsewhere: never freed or regrown here
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf();
    void release();
    // Grows to at least `need` bytes (with 25% headroom, never above `cap`).
    hipError_t reserve(uint64_t need, uint64_t cap = ~0ull);
};

// Bump allocator over one device buffer, reset at the start of every run.
struct Arena {
    DevBuf buf;
    uint64_t used = 0;  // bump offset; callers may save and restore it to re-run a phase
    owm_status ensure(owm_ctx *ctx, uint64_t bytes);
    uint8_t *take(uint64_t bytes);
};

struct Derived {
    double bbox[4];
    double width, height, density, displace;
    uint32_t min_num_points;
    uint32_t ncols, nrows, ncolsg, nrowsg;
    int32_t h, hb;
    double initx, inity, initxg, inityg;
};

struct TreeLayout {
    uint32_t ntiles = 0;
    int tile_bits = 0, key_bits = 32;
    uint64_t np = 0, nleaf = 0, nnode = 0, ngrp = 0;
    std::vector<uint32_t> pt_off, leaf_off, node_off, grp_off;
    std::vector<uint32_t> sort_units;  // first sort tile of every tile's segment
    // key digit counts already taken in the bbox pass (device, the context's
    // small arena), or null: the build takes them itself
    const uint32_t *counts = nullptr;
    // the last pass stamps window classes (1: some tile's gcls is set) or
    // window columns (2: every tile's pairs is set)
    int stamp = 0;
};

struct QueryLayout {
    uint32_t ntiles = 0;
    uint64_t nwin = 0, ncell = 0, nblk = 0;
    uint64_t s1_cap = 0;  // capacity of the stage-1 leaf/block join list
    uint64_t lcs_leaves = 0;  // leaves with summary slots (0: no tile uses summaries)
    uint64_t s1_alg_bytes = 0;  // sum over tiles of 24 n Wsize / Displace
    std::vector<uint32_t> win_off, cell_off, blk_off;
};

struct QueryOut {
    const uint32_t *min_ids = nullptr;         // stage-1 minima (unsorted, per window)
    const uint32_t *llp = nullptr;             // LLPs (then stage-3 adds) before the merge sort
    const uint32_t *cell_res = nullptr;        // stage-3 result per cell
    const uint32_t *seeds = nullptr;           // merged, sorted
    uint32_t *cnt = nullptr;  // [0] n_llp [1] n_add [2] n_seeds [3] error flags [6] s1 list need
    uint32_t *seed_off = nullptr;
};

// Size-class tuning table.  The leaf size is looked up by a tile's point
// count; the stage-1 window block and the sort key width (Morton bits per
// tile, 8, 16, 24 or 32: one to four 8-bit passes) by a launch's mean tile
// size (points / tiles), so that a batch of many small tiles and one large
// tile of the same total keep apart.
struct TuneCfg {
    uint32_t leaf;
    Stage1Cfg s1;
    uint32_t sort_bits;
};

struct Tuner {
    struct Entry {
        uint64_t max_n;
        TuneCfg cfg;
        double ms;  // time of the winning trial (0 = default, never tuned)
    };
    static constexpr uint64_t CLASS_MAX[6] = {1ull << 16, 1ull << 18, 1ull << 20,
                                              1ull << 22, 1ull << 24, ~0ull};
    std::vector<Entry> table;
    bool forced = false;  // a tuning trial overrides the table
    TuneCfg force{};
    Tuner() { reset(); }
    // defaults: tiles of up to 2^20 points leaf 128 and 8-bit keys (one sort
    // pass), larger ones leaf 64 and 16-bit keys; 16x16 window blocks
    void reset()
    {
        table.clear();
        for (uint64_t m : CLASS_MAX)
            table.push_back({m, m <= (1ull << 20) ? TuneCfg{128, {16, 16, 4, 0}, 8}
                                                  : TuneCfg{64, {16, 16, 4, 0}, 16}, 0.0});
    }
    static int class_of(uint64_t n)
    {
        int c = 0;
        while (c < 5 && n > CLASS_MAX[c])
            ++c;
        return c;
    }
    const TuneCfg &cfg_for(uint64_t n) const
    {
        if (forced)
            return force;
        return table[class_of(n)].cfg;
    }
    // a trial with force.leaf == 0 keeps the table's leaf sizes
    uint32_t leaf_for(uint64_t n) const
    {
        return forced && force.leaf ? force.leaf : table[class_of(n)].cfg.leaf;
    }
    Stage1Cfg stage1_for(uint64_t n) const { return cfg_for(n).s1; }
    uint32_t sort_bits_for(uint64_t n) const { return cfg_for(n).sort_bits; }
};

// Where a run takes its memory when not from the context's own arenas and
// pinned buffers (which grow on demand): one borrowed device region that
// holds everything (bbox pass, table, tree, build and query scratch, stage-1
// join list) and caller-provided host staging.  Nothing is allocated.
struct RunMem {
    Arena *arena = nullptr;      // the borrowed device region
    uint32_t *s1_list = nullptr; // stage-1 join list (carved from the region)
    uint64_t s1_words = 0;
    uint8_t *h_io = nullptr;     // host readbacks (tile offsets, bbox keys, counts, ranges)
    uint8_t *h_tab = nullptr;    // host staging of the tile table
};

struct Prepared {
    std::vector<TileDev> td;
    std::vector<uint64_t> n;
    std::vector<uint32_t> leaf;
    std::vector<std::array<double, 4>> box;
    TreeLayout L;
    QueryLayout Q;
    Stage1Cfg cfg{8, 8, 4, 0};
    bool any_need1 = false;
};

owm_status derive_params(const owm_params &P, uint64_t n, const double *bbox, Derived &D,
                         std::string &why);
void fill_query(TileDev &T, const owm_params &P, const Derived &D);
int bit_length(uint64_t v);
// Quantisation of tile t's keys: frames[t] when given, else the square over
// qbox[t].
void plan_tree(int sort_bits, std::vector<TileDev> &tiles, const std::vector<uint64_t> &n,
               const std::vector<uint32_t> &leaf, const std::vector<std::array<double, 4>> &qbox,
               TreeLayout &L, const std::vector<KeyFrame> *frames = nullptr);
void plan_query(std::vector<TileDev> &tiles, const Stage1Cfg &cfg, QueryLayout &Q);
void carve_tree(uint8_t *base, const TreeLayout &L, TreeOut &T);
uint64_t sort_scratch_bytes(uint64_t n, int passes, uint32_t nseg, uint64_t lb_tiles);
// Key/value LSD radix sort (sort.hip): keys in k0 (values in v0, or implicit
// 0..n-1 with iota); hist, counters and lookback, carved in that order after
// v1, must start at zero.  run_sort returns which pair holds the result
// (0: k0/v0, 1: k1/v1) and times the passes as `phase`.
struct SortBufs {
    uint32_t *k0, *v0, *k1, *v1, *hist, *counters, *lookback;
};
SortBufs carve_sort(uint8_t *base, uint64_t n, int passes, uint32_t nseg = 1, uint64_t lb_tiles = 0);
int run_sort(owm_ctx *ctx, int phase, hipStream_t s, SortBufs &S, uint32_t n_max,
             const uint32_t *d_n, int passes, bool keys_only, bool iota);
uint64_t build_scratch_bytes(const TreeLayout &L);
uint64_t table_bytes(uint32_t n_tiles);
owm_status build_tree(owm_ctx *ctx, hipStream_t s, const TreeLayout &L, const BatchDev &B,
                      const double *dx, const double *dy, const double *dz, const TreeOut &T,
                      Arena &scratch, uint32_t *guard);
owm_status enqueue_query(owm_ctx *ctx, hipStream_t s, const BatchDev &B, const TreeDev &TV,
                         const QueryLayout &Q, uint64_t np, const Stage1Cfg &cfg,
                         uint32_t any_need1, Arena &scratch, QueryOut &out, int from, int to,
                         const uint32_t *h_in, uint64_t n_in, const RunMem *mem = nullptr);
owm_status prepare_tree(owm_ctx *ctx, const double *dx, const double *dy, const owm_tile *tiles,
                        uint32_t n_tiles, Prepared &P, const RunMem *mem = nullptr);
owm_status prepare_query(owm_ctx *ctx, const owm_params *params, uint32_t stride_bytes,
                         const std::vector<std::array<double, 4>> &box, uint64_t np_total,
                         Prepared &P);
owm_status upload_table(owm_ctx *ctx, Arena &A, const Prepared &P, BatchDev &B,
                        const RunMem *mem = nullptr);
owm_status read_seeds(owm_ctx *ctx, const QueryOut &qo, const Prepared &P, uint32_t *seeds,
                      uint64_t cap, uint64_t *seed_off_host, uint64_t *s1_needed,
                      const RunMem *mem = nullptr);
// Host bytes of the readback and table staging of an n_tiles batch.
uint64_t io_bytes(uint32_t n_tiles);
// Writes the tile table (tiles, then the eight offset arrays, in the device
// layout) into host memory h of table_bytes(td.size()) bytes.
void write_table(uint8_t *h, const std::vector<::TileDev> &td, const TreeLayout &L,
                 const QueryLayout &Q);
// One tile entirely inside the device region [base, base + bytes) with the
// given host staging; allocates nothing.  OWM_E_BUDGET (and *need set) when
// the tile's work does not fit the region.
owm_status run_tile_fixed(owm_ctx *ctx, const double *dx, const double *dy, const double *dz,
                          uint64_t n, const owm_params &params, uint32_t *seeds, uint64_t cap,
                          uint64_t *n_seeds, uint8_t *base, uint64_t bytes, uint8_t *h_io,
                          uint8_t *h_tab, uint64_t *need);
// Makes ctx->s1_list hold at least cap entries.
owm_status ensure_s1_list(owm_ctx *ctx, uint64_t cap);
// Device-guard flags of a finished query: OWM_OK, or OWM_E_INTERNAL with a
// message; a stage-1 list overflow alone sets *s1_needed and returns OWM_OK.
owm_status check_guards(owm_ctx *ctx, const uint32_t *hc, uint64_t *s1_needed);
uint32_t pick_leaf(owm_ctx *ctx, uint64_t n, uint32_t requested);
Stage1Cfg pick_stage1(owm_ctx *ctx, uint64_t mean_tile);
void set_err(owm_ctx *ctx, const char *fmt, ...);
owm_status hip_fail(owm_ctx *ctx, hipError_t e, const char *what);
// Returns the failure status (with `ctx` in scope) when a HIP call fails.
#define HIP_TRY(call, what)                        \
    do {                                           \
        hipError_t e_ = (call);                    \
        if (e_ != hipSuccess)                      \
            return hip_fail(ctx, e_, what);        \
    } while (0)
// Copies a host cloud into the context's upload buffer (stream-ordered).
owm_status upload_cloud(owm_ctx *ctx, const double *x, const double *y, const double *z,
                        uint64_t n, double **dx, double **dy, double **dz);
// run_batch flag (internal, beside OWM_B_GRAPH): replay from a graph when the
// call repeats the previous auto-graph call's chain exactly.
constexpr uint32_t B_AUTO_GRAPH = 1u << 31;
owm_status run_batch(owm_ctx *ctx, const double *dx, const double *dy, const double *dz,
                     const owm_tile *tiles, uint32_t n_tiles, uint32_t flags, uint32_t *seeds,
                     uint64_t cap, uint64_t *seed_off_host);
uint64_t tree_bytes(const TreeLayout &L);
uint64_t query_bytes(const QueryLayout &Q, uint64_t np, uint64_t n_in);
TreeDev tree_view(const TreeOut &T);

struct ServeState;

enum Phase {
    PH_BBOX, PH_MORTON, PH_SORT_TREE, PH_LEAVES, PH_KARRAS, PH_LEVELS, PH_STAGE1,
    PH_S2_MARK, PH_STAGE2, PH_STAGE3, PH_MERGE, PH_S1_JOIN, PH_LCS, PH_XYZ, PH_QSTATS, PH_COUNT
};

const char *phase_name(int ph);

// HIP-event timing of the kernel families on the context's stream.
struct Profiler {
    bool on = false;
    uint32_t mask = ~0u;  // phases that record events (bit k = phase k)
    struct Rec {
        int ph;
        hipEvent_t a, b;
        uint64_t bytes, items;
        uint32_t launches;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    double ms[PH_COUNT] = {};
    uint64_t calls[PH_COUNT] = {}, bytes[PH_COUNT] = {}, items[PH_COUNT] = {};
    int begin(hipStream_t s, int ph);
    // launches: kernel launches of the phase between begin and end
    void end(hipStream_t s, int idx, uint64_t bytes, uint64_t items, uint32_t launches = 1);
    void collect();  // after the stream has been synchronised
    // bytes and items of a phase known only after its results are read back
    void add_work(int ph, uint64_t bytes, uint64_t items);
    void reset();
    ~Profiler();
};

// Captured batch chains, most recently used first.
struct GraphCache {
    static constexpr size_t MAX_SLOTS = 4;
    struct Slot {
        std::vector<uint8_t> key;
        hipGraphExec_t exec = nullptr;
        QueryOut qo;
        uint64_t used_after = 0;  // arena offset after the chain's allocations
    };
    std::vector<Slot> slots;
    std::vector<uint8_t> last_key;  // key of the last auto-graph call (repeat detection)
    void clear();
    ~GraphCache();
};

}  // namespace owm

struct owm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    std::string err;
    owm::Arena arena;   // tree + scratch of the current run
    owm::Arena small;   // bbox pass
    owm::DevBuf upload; // host inputs staged to the device
    owm::DevBuf s1_list; // stage-1 leaf/block join list (grows on demand)
    // .xyz reader: the text and index, parse work (and the cloud when no line
    // was blank), the compacted cloud
    owm::DevBuf xyz_text, xyz_work, xyz_cloud;
    owm::DevBuf qs_work;  // quadtree statistics: point state, child slots, histograms
    uint64_t s1_need_max = 0;  // largest join that overflowed its first estimate
    owm::Tuner tuner;
    std::atomic<uint64_t> trees_bytes{0};
    // pinned host staging: the tile table (tiles, then eight offset arrays, in
    // the device layout, uploaded by one copy) and small readbacks (tile point
    // offsets, bbox keys, counts, seed ranges)
    uint8_t *h_tab = nullptr;
    uint64_t h_tab_cap = 0;
    uint8_t *h_io = nullptr;
    uint64_t h_io_cap = 0;
    uint8_t *h_seeds = nullptr;  // serving slots: the seeds of a group, reused
    uint64_t h_seeds_cap = 0;
    uint8_t *pinned(uint8_t *&buf, uint64_t &cap, uint64_t need);
    owm::ServeState *serve = nullptr;
    owm::Profiler prof;  // owned; released by owm_serve_stop
    owm::GraphCache graphs;  // captured batch chains (OWM_B_GRAPH)
    owm::Prepared fixed_plan;  // planning state of run_tile_fixed, reserved at creation
    // copies the tile table into h_tab (false: no pinned memory)
    bool stage_table(const std::vector<::TileDev> &td, const owm::TreeLayout &L,
                     const owm::QueryLayout &Q);
    uint64_t held_bytes() const;
    // Largest size buffer `b` may grow to without exceeding the budget.
    uint64_t room_for(const owm::DevBuf &b) const;
    ~owm_ctx();
};

// A device-resident tree (owm_tree_build): its own device buffer, the host copy
// of its layout, and the context whose stream built it (null once handed to
// the serving cache, which then owns it).
struct owm_tree {
    owm_ctx *ctx = nullptr;
    owm::DevBuf mem;
    owm::Prepared P;
    owm::TreeOut T{};
    uint64_t n = 0;
};
