// Host-side launchers of the OWM kernels (one translation unit per kernel
// family; every launcher enqueues on the given stream and never synchronises).
#pragma once
#include "owm_dev.h"

namespace owm {

// ---- radix sort (sort.hip) ----
// Segments sorted independently: [pt[t], pt[t+1]) for t < nseg, sort tiles
// units[t] .. units[t+1]-1 (units[t] = sum of ceil(n_s / record_tile_keys(n, nseg))
// over s < t), digit offsets of segment t at digit_off + t * digit_stride.
// pt == null: one segment [0, n).
struct SortSegs {
    const uint32_t *pt = nullptr;
    const uint32_t *units = nullptr;
    const uint32_t *tile_seg = nullptr;  // segment of every sort tile (else searched in units)
    uint32_t nseg = 0;
    uint32_t digit_stride = 0;
};
uint32_t sort_tiles(uint32_t n);
// points per tile of an n-point record sort over nseg segments
uint32_t record_tile_keys(uint32_t n, uint32_t nseg);
constexpr uint32_t SORT_SEG_HIST = 4 * 256;  // histogram words per segment (all passes)
// counts[t * SORT_SEG_HIST + p * 256 + d] -> hist[same] = segment base +
// exclusive offsets (counts may equal hist); tile_seg[u] = t for the sort
// tiles u of segment t (units[t] <= u < units[t+1])
void launch_sort_scan_seg(hipStream_t s, const uint32_t *counts, uint32_t *hist, const uint32_t *pt,
                          const uint32_t *units, uint32_t nseg, int passes, uint32_t *counters,
                          uint32_t *tile_seg);
void launch_fill_u32(hipStream_t s, uint32_t *p, uint64_t n, uint32_t v);
// Zeroes up to FILL_MAX ranges (u32 words) in one launch.
constexpr uint32_t FILL_MAX = 16;
struct FillRanges {
    uint32_t count = 0;
    uint32_t *ptr[FILL_MAX] = {};
    uint64_t len[FILL_MAX] = {};
};
void launch_fill_ranges(hipStream_t s, const FillRanges &r);
void launch_sort_hist(hipStream_t s, const uint32_t *keys, const uint32_t *d_n, uint32_t n_max,
                      int passes, uint32_t *hist);
void launch_sort_scan(hipStream_t s, uint32_t *hist, int passes, uint32_t *counters);
// seg_tiles: sort tiles over all segments (units[nseg]) when seg.pt is set
void launch_onesweep(hipStream_t s, const uint32_t *keys_in, const uint32_t *vals_in,
                     uint32_t *keys_out, uint32_t *vals_out, const uint32_t *d_n, uint32_t n_max,
                     int shift, const uint32_t *digit_off, uint32_t *lookback,
                     uint32_t *tile_counter, bool keys_only, bool iota, bool write_keys,
                     uint32_t grid_cap, const SortSegs &seg = SortSegs(), uint32_t seg_tiles = 0);

// Input of a pass of the tree sort: the first pass reads the caller's
// coordinate arrays (x, y, z; it numbers the points and computes their keys
// from the tiles' quantisation), later passes the previous pass's records.
struct PointsIn {
    const double *x, *y, *z;
    PointsDev rec;
};

// One pass of the segmented tree sort over n points (segments seg,
// seg_tiles sort tiles in all), by the 8-bit digit at `shift` of keys of
// key_bits bits; whole 32-byte records move.  pos (may be null): pos[id] =
// output position, written by a last pass.
void launch_sort_records(hipStream_t s, const BatchDev &B, const PointsIn &in, const PointsOut &out,
                         uint32_t *pos, uint32_t n, int shift, int key_bits,
                         const uint32_t *digit_off, uint32_t *lookback, uint32_t *tile_counter,
                         bool first, int stamp, const SortSegs &seg, uint32_t seg_tiles);

// ---- tree (tree.hip) ----
struct TreeOut {
    PointsOut pts;  // the cloud in Morton order
    uint32_t *pos;
    double4 *lf_box;
    uint32_t *lf_cnt, *lf_start, *lf_npts, *lf_zid, *lf_key0;
    unsigned long long *lf_zkey;
    double4 *nd_box;
    uint32_t *nd_cnt, *nd_zid;
    unsigned long long *nd_zkey;
    uint2 *nd_child;  // children (OWM_LEAF_BIT | leaf, or node)
};

// bbox per tile as order-preserving u64 keys [minx, miny, maxx(inv), maxy(inv)]
void launch_bbox(hipStream_t s, const BatchDev &B, const double *x, const double *y,
                 unsigned long long *bbox_keys);
// bbox_keys holds bbox_key_words(ntiles) words; the result is in the first 4*ntiles
// (min x, min y, max x, max y as order keys per tile).
uint64_t bbox_key_words(uint32_t ntiles);
void launch_bbox_init(hipStream_t s, unsigned long long *bbox_keys, uint32_t ntiles);
// Per tile SORT_SEG_HIST digit counters of every pass of the points' sort keys
// (the keys themselves are not stored: the first sort pass computes them again).
// exact bbox words (as launch_bbox) and the tiles' key histograms (as
// launch_morton) in one pass over x and y, the keys quantised over frames
// sampled from each tile (written to frames, ntiles entries); bk initialised
// by launch_bbox_init, hist zeroed.
void launch_bbox_keys(hipStream_t s, const BatchDev &B, const double *x, const double *y,
                      KeyFrame *frames, unsigned long long *bk, uint32_t *hist, int passes,
                      uint32_t key_bits);
void launch_morton(hipStream_t s, const BatchDev &B, const double *x, const double *y,
                   uint32_t *hist, int passes);
// Leaf aggregates over the points in Morton order (T.pts); each leaf's first
// key is its first point's.
void launch_leaf_aggs(hipStream_t s, const BatchDev &B, const TreeOut &T);
// Aggregate levels: level k (1..nlev) holds the merged aggregate of every
// aligned block of 2^k leaves of the batch, size[k] entries from entry off[k].
constexpr uint32_t AGG_MAX_LEV = 32;
struct AggLevels {
    double4 *box;
    uint32_t *cnt, *zid;
    unsigned long long *zk;
    uint32_t nlev;
    uint32_t off[AGG_MAX_LEV + 1], size[AGG_MAX_LEV + 1];
};
// Level sizes for nleaf leaves (pointers unset); *entries = entries of all levels.
AggLevels agg_levels_layout(uint64_t nleaf, uint64_t *entries);
// Levels from the leaf aggregates (one launch per 11 levels).
void launch_agg_levels(hipStream_t s, const TreeOut &T, const AggLevels &V);
// Karras tree over each tile's leaves; every node's aggregate is merged from
// the levels' blocks that tile its leaf range (levels built first).
void launch_karras(hipStream_t s, const BatchDev &B, const TreeOut &T, const AggLevels &V);

// ---- .xyz reader (xyz.hip) ----
// The text buffer is zero-padded to a whole number of xyz_chunk_bytes() plus
// 16 bytes.  index: newline counts per chunk, their exclusive scan (chunk_off
// has nchunks+1 entries, the last the total) and every newline's position.
uint64_t xyz_chunk_bytes();
void launch_xyz_count(hipStream_t s, const void *text, uint64_t padded_bytes, uint32_t *chunk_cnt,
                      uint64_t *chunk_off);
void launch_xyz_newlines(hipStream_t s, const void *text, uint64_t padded_bytes,
                         const uint64_t *chunk_off, uint64_t *nlpos);
void launch_xyz_parse(hipStream_t s, const void *text, uint64_t bytes, const uint64_t *nlpos,
                      uint64_t nnl, uint64_t nlines, double *x, double *y, double *z,
                      uint64_t *blank, uint64_t *host, uint64_t *host_start, uint32_t cap,
                      uint32_t *counts);
void launch_xyz_scatter(hipStream_t s, const uint64_t *line, const double *v, uint64_t n, double *x,
                        double *y, double *z);
void launch_xyz_compact(hipStream_t s, const double *x, const double *y, const double *z,
                        uint64_t nlines, const uint64_t *blank, uint64_t nb, double *ox, double *oy,
                        double *oz);

// ---- quadtree leaf statistics (qstats.hip) ----
// The reference's parallel_qtree (maxNumber or minRadius policy) resolved one
// depth at a time: open nodes (split nodes) of depth d-1 own child slots
// 4 * open + quadrant at depth d.
constexpr int QS_MAX_DEPTH = 64;
constexpr uint32_t QS_HB = 1024;  // exact bins of leaf point counts per depth
// A point's 32-bit path key holds its quadrants at depths 1..16: the sorted
// depths 1..S (S = QsDev.sort_levels, 12 or 16) in bits 0..2S-1 (depth 1
// highest), depths S+1..16 in the bits above (depth S+1 highest).
constexpr int QS_KEY_LEVELS = 16;
// Root centre and, per depth k < QS_KEY_LEVELS, the offset from a depth-k
// centre to its children's (createQuadrants: radius * 0.5f, as a double).
struct QsPath {
    double cx, cy;
    double off[QS_KEY_LEVELS];
};
struct QsDev {
    const double *x, *y;
    uint64_t n;
    // points in path-key order: key (null: none) and the index of the point
    // at each position (null: identity)
    const uint32_t *key, *perm;
    uint2 *rng[2];    // open-node key ranges [first, end) down to depth sort_levels, by parity
    uint32_t *st;     // below sort_levels, per point: child slot at the last routed depth (OWM_NONE: in a leaf)
    uint32_t *cnt;    // per child slot: points routed into it (zero outside its classify)
    uint32_t *remap;  // per child slot: open-node index at its depth, OWM_NONE for a leaf
    double2 *cen[2];  // open-node centres, by depth parity
    uint32_t *nopen;  // [QS_MAX_DEPTH + 1] open nodes per depth (capped by open_cap)
    unsigned long long *hist;  // [(QS_MAX_DEPTH + 1) * QS_HB] leaves per (depth, points)
    uint2 *ovf;       // (depth, points) of every leaf holding >= QS_HB points
    uint32_t *novf;   // entries written to ovf
    uint32_t *flags;  // bit 0: ovf full, bit 1: more open nodes than open_cap
    uint32_t ovf_cap, open_cap;
    uint32_t policy;  // 0 maxNumber, 1 minRadius
    int sort_levels;  // depths sorted in the keys (12 or 16)
    uint32_t max_number, level;
    float min_radius;
};
// Root at depth 0: open node 0 with every point's key range.  cnt, nopen,
// hist, novf and flags must be zero.
void launch_qs_root(hipStream_t s, const QsDev &D, double cx, double cy);
// Depth d <= sort_levels from the open depth-(d-1) nodes' key ranges (keys
// sorted by their sorted depths); rd, off as for launch_qs_level.
void launch_qs_split(hipStream_t s, const QsDev &D, int d, float rd, double off);
// Hands the points of the nodes open at depth sort_levels to launch_qs_level.
void launch_qs_deep_start(hipStream_t s, const QsDev &D);
// key[i]: the quadrants point i takes at depths 1..16 (quadrantIdx against
// the createQuadrants centres), laid out as above.
void launch_qs_path(hipStream_t s, const double *x, const double *y, uint64_t n, const QsPath &P,
                    int sort_levels, uint32_t *key);
// Depth d > sort_levels: route the points of the open depth-(d-1) nodes to
// their child slots (by their key down to depth 16, by their coordinates
// below), then classify the slots (rd = the depth-d radius, off = the centre
// offset of createQuadrants at depth d-1).
void launch_qs_level(hipStream_t s, const QsDev &D, int d, float rd, double off);

// ---- stages (stages.hip) ----
struct Stage1Cfg {
    uint32_t bx, by;  // window block shape
    uint32_t k;       // points per lane per binning round
    uint32_t lcs;     // 1: leaf cell summaries where a tile's leaves are small enough
    uint32_t pairs;   // 1: window columns (every tile's edges pair up; records stamped so)
};
// Largest window side in window steps for stage 1 on window columns, and the
// fewest 32x32-window blocks for which it uses them (one per workgroup slot
// of the device, two per CU; else 16x16).
constexpr uint32_t S1_PAIRS_MAX = 7;
constexpr uint64_t S1_PAIRS_BIG_BLOCKS = 512;
bool stage1_shape_ok(const Stage1Cfg &c);
uint64_t stage1_work_words(uint64_t nblk);
// Stage 1 in three launches.  The join lists, per window block, the leaves
// whose bbox meets it (work: stage1_work_words(nblk) u32; list capacity cap;
// *total_out receives the entries needed, and bit 16 of err is set when they
// exceed cap, after which the call must be repeated with a larger list).  The
// leaf cell summaries are built for the tiles whose TileDev.lcs is set.  The
// bins write one minimum index per window (lc null: no tile uses summaries;
// else only shapes with stage1_shape_ok(lcs=1)).
void launch_stage1_join(hipStream_t s, const BatchDev &B, const TreeDev &T, const Stage1Cfg &cfg,
                        uint32_t *work, uint32_t *list, uint32_t cap, uint32_t *err,
                        uint32_t *total_out);
void launch_leaf_cells(hipStream_t s, const BatchDev &B, const TreeDev &T, const LeafCells &lc);
void launch_stage1_bins(hipStream_t s, const BatchDev &B, const TreeDev &T, const Stage1Cfg &cfg,
                        const LeafCells *lc, const uint32_t *work, const uint32_t *list,
                        uint32_t cap, uint32_t *min_ids);
uint64_t leaf_cells_bytes(uint64_t nleaf);
LeafCells carve_leaf_cells(uint8_t *base, uint64_t nleaf);
// window minima -> point bitmaps: `seen` (one window) and `llp` (two, or one on
// tiles with stage2_need == 1); both np/32 words, zeroed beforehand
void launch_stage2_mark(hipStream_t s, const BatchDev &B, const uint32_t *min_ids, uint32_t nwin,
                        uint32_t np, uint32_t any_need1, uint32_t *seen, uint32_t *llp);
// ids[*d_off .. *d_off + *d_n) (n_max bounds *d_n) marked in `bits`
void launch_bits_set(hipStream_t s, const uint32_t *ids, const uint32_t *d_off, const uint32_t *d_n,
                     uint32_t n_max, uint32_t np, uint32_t *bits);
// set bits of words[0..nwords) -> ascending indices + count
void launch_bits_compact(hipStream_t s, const uint32_t *words, uint32_t nwords, uint32_t *out,
                         uint32_t *d_count, uint32_t *lookback, uint32_t *tile_counter);
void launch_stage3(hipStream_t s, const BatchDev &B, const TreeDev &T, const uint32_t *llp,
                   const uint32_t *n_llp, uint32_t n_llp_max, uint32_t *cell_flag,
                   uint32_t *cell_res, uint32_t *err);
// cell_res (NONE = no add) -> compacted list appended after the LLPs
void launch_compact_valid(hipStream_t s, const uint32_t *in, uint32_t n, uint32_t *out,
                          const uint32_t *d_out_base, uint32_t *d_count, uint32_t *lookback,
                          uint32_t *tile_counter);
void launch_seed_ranges(hipStream_t s, const BatchDev &B, const uint32_t *seeds,
                        const uint32_t *d_nseeds, uint32_t *seed_off);
// seeds (sorted batch indices, count at d_nseeds <= n_max) -> tile-local indices
void launch_seeds_local(hipStream_t s, const BatchDev &B, uint32_t *seeds, const uint32_t *d_nseeds,
                        uint32_t n_max);
void launch_add_u32(hipStream_t s, const uint32_t *a, const uint32_t *b, uint32_t *out);

}  // namespace owm
