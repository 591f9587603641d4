/*
 * owm.h -- C ABI of the MI355X-native OWM (Overlapping Window Minimum)
 * ground-seed pipeline.
 *
 * A LiDAR cloud (fp64 x, y, z arrays) and the OWM parameters go in; the
 * indices of the ground seed points come out, sorted ascending, 0-based into
 * the caller's arrays (the reference's 1-based id minus one; the reference's
 * sentinel slot 0 does not exist here).
 *
 * Plain C types only: no C++ or torch types cross this boundary, no call ever
 * exits the process, no exception crosses it.  Every entry point returns an
 * owm_status; owm_last_error() gives the message of the last failure on a
 * context.  One context may be used by one host thread at a time (the serving
 * loop owns its own thread and locks its own queue).
 */
#ifndef OWM_H
#define OWM_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum owm_status {
    OWM_OK = 0,
    OWM_E_ARG = 1,       /* invalid parameters or inputs */
    OWM_E_NOMEM = 2,     /* device or host allocation failed */
    OWM_E_BUDGET = 3,    /* the work does not fit the context's HBM budget */
    OWM_E_DEVICE = 4,    /* HIP runtime error */
    OWM_E_CAPACITY = 5,  /* output buffer too small; the size needed is returned */
    OWM_E_CANCELLED = 6, /* served request was cancelled */
    OWM_E_NOTREADY = 7,  /* served request not finished yet */
    OWM_E_NOTFOUND = 8,  /* unknown request id or tree handle */
    OWM_E_INTERNAL = 9   /* a device-side capacity guard tripped (see message) */
} owm_status;

/* owm_params.flags */
#define OWM_F_HAVE_BBOX 1u /* bbox[] holds the header bbox; else computed from the data */

/* Parameters of one OWM run.  Field meanings follow the reference CLI
 * (CPP-TBB-CPU/application/owm_optim3.cpp:14-40): -W Wsize, -B Bsize,
 * -O Overlap, -s maxNumber.  bbox is the header bbox of readXYZfile
 * (CPP-TBB-CPU/src/optim3_func.cpp:681-700), in header order. */
typedef struct owm_params {
    uint32_t wsize;     /* sliding-window side in metres (u32, as the reference) */
    uint32_t bsize;     /* stage-3 grid cell side in metres; 0 skips stage 3 */
    double overlap;     /* window overlap ratio, 0 <= overlap < 1 */
    uint32_t leaf_size; /* maxNumber points per tree leaf; 0 = autotuned */
    uint32_t flags;     /* OWM_F_* */
    double bbox[4];     /* minx, maxx, miny, maxy when OWM_F_HAVE_BBOX */
} owm_params;

/* Derived parameters (application/owm_optim3.cpp:92-132), for callers that
 * size their buffers: windows = ncols*nrows, grid cells = ncolsg*nrowsg. */
typedef struct owm_derived {
    double width, height, density, displace;
    uint32_t min_num_points; /* after the reference's ushort truncation */
    uint32_t ncols, nrows, ncolsg, nrowsg;
    double bbox[4];
} owm_derived;

typedef struct owm_ctx owm_ctx;
typedef struct owm_tree owm_tree;

/* ---- context ------------------------------------------------------------ */

/* Creates a context on HIP device `device`.  hbm_budget_bytes caps the device
 * memory the context may hold (0 = 90% of the device's free memory). */
owm_status owm_ctx_create(int device, uint64_t hbm_budget_bytes, owm_ctx **out);
void owm_ctx_destroy(owm_ctx *ctx);
const char *owm_last_error(const owm_ctx *ctx);
const char *owm_status_str(owm_status s);
/* Device bytes currently held by the context's arenas. */
uint64_t owm_ctx_device_bytes(const owm_ctx *ctx);

/* Page-locked host memory for seed buffers and host clouds: copies between it
 * and the device run as one DMA instead of staged pageable chunks.  No
 * reference counterpart (the reference's buffers are plain vectors).
 * owm_host_free(NULL) is a no-op. */
owm_status owm_host_alloc(uint64_t bytes, void **out);
void owm_host_free(void *ptr);

/* Per-phase device timing, measured with HIP events recorded on the
 * context's own stream around each kernel family.  `bytes` is the phase's
 * algorithmic HBM traffic (what the work must read and write once), so
 * bytes / ms is the achieved bandwidth against the roofline.
 * No counterpart in the reference (its -DDEBUG timers, ordered_builder.inl:433-540,
 * print per-kernel times to stdout). */
typedef struct owm_phase_stat {
    char name[32];
    double ms;       /* summed device time */
    uint64_t calls;  /* launches of the phase */
    uint64_t bytes;  /* summed algorithmic bytes */
    uint64_t items;  /* summed work items (points, windows, keys ...) */
} owm_phase_stat;
/* enable != 0 starts recording and zeroes the counters; 0 stops. */
owm_status owm_profile_enable(owm_ctx *ctx, int enable);
/* Phases that record events while profiling is on: bit k = the k-th phase of
 * owm_profile_read (0 = all, the default).  Each recorded phase costs two
 * events on the stream per occurrence, so timing one kernel family leaves the
 * others undisturbed. */
owm_status owm_profile_phases(owm_ctx *ctx, uint32_t mask);
owm_status owm_profile_read(owm_ctx *ctx, owm_phase_stat *out, uint32_t cap, uint32_t *n);

/* Parameter derivation only (no device work): owm_optim3.cpp:92-132. */
owm_status owm_derive(const owm_params *params, uint64_t n, const double *bbox_or_null,
                      owm_derived *out);
/* Seed-buffer sizing: an upper bound on the seeds of an n-point tile,
 * min(n, ncols*nrows + ncolsg*nrowsg) (every seed is a window minimum or a
 * stage-3 addition); n when neither params nor bbox_or_null give a bbox.  The
 * reference sizes its minIDs buffer from nCols*nRows (owm_optim3.cpp:134). */
owm_status owm_query_seed_count(const owm_params *params, uint64_t n, const double *bbox_or_null,
                                uint64_t *max_seeds);

/* ---- whole pipeline, one tile -------------------------------------------
 * Replaces the timed loop of the reference driver: tree construction
 * (parallel_qtree + storeMinAndNumPoints, owm_optim3.cpp:77-89), stage 1,
 * sort + stage 2, stage 3 and the final merge (owm_optim3.cpp:143-230).
 * seeds: caller-owned, capacity `cap`; on OWM_E_CAPACITY *n_seeds holds the
 * size needed and nothing is written. */
owm_status owm_run_tile(owm_ctx *ctx, const double *x, const double *y, const double *z,
                        uint64_t n, const owm_params *params, uint32_t *seeds, uint64_t cap,
                        uint64_t *n_seeds);
/* Same, with x/y/z already in device memory (HBM); seeds go to host memory. */
owm_status owm_run_tile_device(owm_ctx *ctx, const double *d_x, const double *d_y,
                               const double *d_z, uint64_t n, const owm_params *params,
                               uint32_t *seeds, uint64_t cap, uint64_t *n_seeds);

/* owm_run_tile_device without any allocation: no counterpart in the reference.
 * For callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as owm_run_tile_device; every device buffer of the run is carved
 * from a 256 MiB static scratch array that each device holds for the life of
 * the process, host staging lives on the call's stack, phase timing is paused,
 * and nothing is allocated (no hipMalloc/hipFree, no page-locked or host heap
 * memory, no events).  A tile whose work needs more scratch than that (the
 * work takes about 76 bytes per point under the default table, so up to
 * about 3.5 million points; longer sort keys take more) is
 * refused with OWM_E_BUDGET and a message giving the bytes needed.  The scratch is shared by every call on a device and
 * keeps the last call's contents: calls that overlap on one device (from two
 * threads or two contexts) race and must be serialised by the caller. */
owm_status owm_run_tile_device_noalloc(owm_ctx *ctx, const double *d_x, const double *d_y,
                                       const double *d_z, uint64_t n, const owm_params *params,
                                       uint32_t *seeds, uint64_t cap, uint64_t *n_seeds);

/* ---- reader: LiDAR .xyz text -> device-resident cloud --------------------
 * Replaces readXYZfile (CPP-TBB-CPU/src/optim3_func.cpp:665-722): text holds
 * the whole file (bytes long, host memory).  A first line holding a single
 * number starts the reference's 5-number header "N minx maxx miny maxy" (one
 * per line, as in its *H.xyz files); N then bounds the points read and the
 * box is returned.  Every other non-blank line is one point: three numbers
 * separated by blanks, the rest of the line ignored (the reference's
 * fscanf("%lf %lf %lf") + skip to '\n'); the doubles equal strtod's.  Blank
 * lines are skipped; a line with fewer than three numbers is OWM_E_ARG.  The
 * text is parsed on the device; lines outside the exact device path (more
 * than 19 significant digits, |decimal exponent| > 22 with the mantissa,
 * inf/nan/hex) are parsed by the host (host_lines counts them).  The cloud
 * lives in device memory owned by the context, valid until the next
 * owm_xyz_parse or owm_ctx_destroy; pass it to owm_run_tile_device or the
 * batch calls (bbox with OWM_F_HAVE_BBOX when flags says so). */
typedef struct owm_cloud {
    const double *x, *y, *z; /* device pointers */
    uint64_t n;
    double bbox[4];          /* header box: minx, maxx, miny, maxy */
    uint32_t flags;          /* OWM_F_HAVE_BBOX when the file had a header */
    uint64_t host_lines;     /* lines the host parsed */
} owm_cloud;
owm_status owm_xyz_parse(owm_ctx *ctx, const char *text, uint64_t bytes, owm_cloud *out);
/* Copies a parsed cloud to caller-owned host arrays of cloud->n doubles
 * (no reference counterpart: readXYZfile fills host memory directly). */
owm_status owm_cloud_read(owm_ctx *ctx, const owm_cloud *cloud, double *x, double *y, double *z);
/* readXYZfile's table of the reference's header-less clouds
 * (CPP-TBB-CPU/src/optim3_func.cpp:667-680): when `path` contains one of their
 * file names (INAER_2011_Alcoy.xyz, INAER_2011_Alcoy_Core.xyz,
 * BABCOCK_2017_Arzua_3B.xyz, V21_/V19_group1_densified_point_cloud.xyz,
 * sample24.xyz), *n_points and bbox (minx, maxx, miny, maxy) get the values
 * the reference hard-codes for it; else OWM_E_NOTFOUND.  Host only. */
owm_status owm_xyz_known_cloud(const char *path, uint64_t *n_points, double *bbox);

/* ---- GOLD comparison (host only) -----------------------------------------
 * The reference's check_results (CPP-TBB-CPU/src/optim3_func.cpp:725-794): a
 * seed (sx, sy, sz) is correct when a GOLD point lies strictly inside the
 * square of half side (float)Displace * 0.5 around it with |z - gz| < 0.01;
 * *percent = correct seeds / n_gold * 100 (the driver's Correctness column). */
owm_status owm_gold_check(const double *sx, const double *sy, const double *sz, uint64_t n_seeds,
                          const double *gx, const double *gy, const double *gz, uint64_t n_gold,
                          double displace, uint64_t *n_correct, double *percent);

/* The driver's results file (save_time, CPP-TBB-CPU/src/optim3_func.cpp:797-828,
 * same argument order): appends "exec input threads minRadius maxNumber level
 * treeTime owmTime correctness" to results_path (minRadius and correctness as
 * %g, the times as %.6f), after the header line "ExecName InputFile NumThreads
 * MinRadius MaxNumber Level TreeTime OwmTime Correctness" when the file does
 * not exist.  OWM_E_ARG when the file cannot be written.  Host only. */
owm_status owm_save_time(const char *exec_name, const char *results_path, const char *input_file,
                         int num_threads, float min_radius, int max_number, int level,
                         double tree_time, double owm_time, double correctness);

/* ---- quadtree leaf statistics (owm_stats) ---------------------------------
 * The leaves of the reference's parallel_qtree (CPP-TBB-CPU/src/optim3_func.cpp:
 * 218-253) as the statistics tool walks them (CPP-TBB-CPU/application/
 * owm_stats.cpp:60-136, include/stats_func.hpp:6-56): `statsmaxnum` builds it
 * with the maxNumber policy, `stats` with minRadius.  Computed on the device
 * without building the tree; the result is the distinct (depth, points)
 * pairs of the leaves with how many leaves have each, from which every
 * histogram of the tool follows (points per leaf, radius, density, level). */
#define OWM_QT_MAXNUMBER 0u /* split a node holding more than max_number points */
#define OWM_QT_MINRADIUS 1u /* split a node holding a point while radius * 0.5 > min_radius */
#define OWM_QT_MAX_LEVEL 10u
#define OWM_QT_MAX_DEPTH 64u
typedef struct owm_qtree_params {
    uint32_t policy;     /* OWM_QT_MAXNUMBER or OWM_QT_MINRADIUS */
    uint32_t max_number; /* -s maxNumber (>= 1 under OWM_QT_MAXNUMBER) */
    float min_radius;    /* -r minRadius */
    uint32_t level;      /* -L: depths 0..level are split whatever they hold (<= OWM_QT_MAX_LEVEL) */
    uint32_t flags;      /* OWM_F_HAVE_BBOX: bbox is the header box; else the data's */
    double bbox[4];      /* minx, maxx, miny, maxy */
} owm_qtree_params;
typedef struct owm_leaf_bin {
    uint32_t level;  /* depth of the leaves (root = 0) */
    uint32_t points; /* points each holds */
    uint64_t leaves; /* leaves with this depth and count */
} owm_leaf_bin;
typedef struct owm_qtree_summary {
    double bbox[4];     /* the box the root was centred on */
    float max_radius;   /* the root's radius (getRadius, a float) */
    uint32_t depth;     /* deepest leaf */
    uint64_t leaves;    /* leaf nodes */
    uint64_t n_bins; /* distinct (level, points) pairs */
} owm_qtree_summary;
/* bins[0 .. min(cap, n_bins)) in ascending (level, points) order;
 * OWM_E_CAPACITY when cap < out->n_bins (out is filled).  OWM_E_ARG when a
 * node would split below OWM_QT_MAX_DEPTH (more than max_number points at one
 * spot: the reference recursion never ends there) or n >= 2^30 under
 * OWM_QT_MINRADIUS.  x, y: device pointers when device_ptrs != 0. */
owm_status owm_qtree_stats(owm_ctx *ctx, const double *x, const double *y, uint64_t n,
                           int device_ptrs, const owm_qtree_params *params,
                           owm_leaf_bin *bins, uint64_t cap, owm_qtree_summary *out);
/* Appends the tool's four reports to the text file `path` as save_histogram,
 * save_areas, save_density and save_levels do (stats_func.hpp:120-366; the
 * average density is n / (round2d(width) * round2d(height))). */
owm_status owm_qtree_stats_write(const char *path, const owm_qtree_params *params,
                                 const owm_qtree_summary *summary, const owm_leaf_bin *bins,
                                 uint64_t n_bins, uint64_t n_points);

/* ---- batch: many independent tiles in one launch chain ------------------
 * The tiles are slices [offset, offset+n) of one concatenated SoA cloud.
 * Seeds of tile t are written to seeds[seed_off[t] .. seed_off[t+1]) with
 * tile-local indices; seed_off has n_tiles+1 entries. */
typedef struct owm_tile {
    uint64_t offset;
    uint64_t n;
    owm_params params;
} owm_tile;

/* flags of owm_batch_run*: */
#define OWM_B_GRAPH 1u /* replay the device chain from a captured hipGraph */

owm_status owm_batch_run(owm_ctx *ctx, const double *x, const double *y, const double *z,
                         const owm_tile *tiles, uint32_t n_tiles, uint32_t flags,
                         uint32_t *seeds, uint64_t cap, uint64_t *seed_off);
owm_status owm_batch_run_device(owm_ctx *ctx, const double *d_x, const double *d_y,
                                const double *d_z, const owm_tile *tiles, uint32_t n_tiles,
                                uint32_t flags, uint32_t *seeds, uint64_t cap,
                                uint64_t *seed_off);

/* ---- size-class autotuner -------------------------------------------------
 * No counterpart in the reference library (there, tree granularity was
 * tuned by sweep scripts over maxNumber).  Point-count classes: up to 2^16,
 * 2^18, 2^20, 2^22, 2^24 and beyond.  A tile whose params.leaf_size is 0
 * takes the leaf size of its own class; the stage-1 window block and the sort
 * key width come from the class of the launch's mean tile size (its points
 * over its tiles). */
#define OWM_TUNE_CLASSES 6u
typedef struct owm_tune_entry {
    uint64_t max_n;      /* class bound: point counts up to max_n */
    uint32_t leaf_size;  /* points per leaf (maxNumber) */
    /* stage-1 window block, points per lane per round (whole runs whose every
     * tile has a window side of at most 7 whole window steps use 32x32 or
     * 16x16 blocks of window columns instead) */
    uint32_t block_x, block_y, block_k;
    uint32_t summaries;  /* 1: stage 1 reads per-leaf summaries of equal window-edge classes */
    uint32_t sort_bits;  /* Morton bits of the per-tile sort key: 8, 16, 24 or 32 (1-4 passes) */
    double ms;           /* time of the winning trial (owm_autotune) or the model's
                          * predicted time (owm_autotune_stats); 0 = default */
} owm_tune_entry;
/* Times trial configurations on the given batch (tiles as in owm_batch_run,
 * their leaf_size ignored): leaf sizes 16..128 per tile class on the tiles of
 * that class, then window blocks and sort widths for the whole batch.  Stores
 * the winners in the context's table (used by every later call, and copied
 * into serving slots at owm_serve_start) and returns the table in out
 * (cap >= OWM_TUNE_CLASSES). */
owm_status owm_autotune(owm_ctx *ctx, const double *x, const double *y, const double *z,
                        int device_ptrs, const owm_tile *tiles, uint32_t n_tiles,
                        owm_tune_entry *out, uint32_t cap, uint32_t *n_out);
owm_status owm_tune_get(owm_ctx *ctx, owm_tune_entry *out, uint32_t cap, uint32_t *n_out);
/* Loads a table (e.g. one saved from owm_tune_get); classes not listed keep
 * the defaults: block 16x16 with 4 points per lane, no summaries; up to 2^20
 * points leaf 128 and 8-bit keys, beyond leaf 64 and 16-bit keys (over the
 * tile sets the statistics model below was fitted to, 4 % slower than each
 * set's fastest configuration on average, against 15 % for the best single
 * configuration for all classes). */
owm_status owm_tune_set(owm_ctx *ctx, const owm_tune_entry *in, uint32_t n);

/* ---- statistics-driven tuning ----------------------------------------------
 * The reference picked maxNumber and minRadius by hand from its statistics
 * tool's leaf histograms (CPP-TBB-CPU/application/owm_stats.cpp,
 * include/stats_func.hpp:6-56).  Here a cost model of the batch chain, fitted
 * to trial times on MI355X, reads such histograms (owm_qtree_stats under the
 * maxNumber policy with the settings below) of a few tiles per size class and
 * picks the table -- leaf size per tile class, window block and key width per
 * class of a launch's total -- without timing anything. */
#define OWM_TUNE_STATS_MAXNUMBER 256u /* histogram settings the model expects */
#define OWM_TUNE_STATS_LEVEL 0u
#define OWM_TUNE_SAMPLES 16u          /* tiles sampled per class by owm_autotune_stats */
typedef struct owm_tune_sample {
    uint64_t n;                /* points of the sampled tile */
    double weight;             /* points of the batch it stands for (>= n) */
    owm_params params;         /* its OWM parameters (wsize, overlap; bbox if flagged) */
    const owm_leaf_bin *bins;  /* its owm_qtree_stats histogram (settings above) */
    uint64_t n_bins;
    owm_qtree_summary summary; /* the same call's summary (box, root radius) */
} owm_tune_sample;
/* The model alone, host code: the table it picks for a batch the samples
 * describe.  Classes without samples keep the defaults; out[k].ms is the
 * predicted time of a launch of that class's tiles (for the class of the total
 * weight: of the whole batch).  OWM_E_ARG for a sample whose parameters or
 * histogram the model cannot use; cap >= OWM_TUNE_CLASSES. */
owm_status owm_tune_model(const owm_tune_sample *samples, uint32_t n_samples,
                          owm_tune_entry *out, uint32_t cap, uint32_t *n_out);
/* owm_autotune's counterpart without trials: histograms of up to
 * OWM_TUNE_SAMPLES tiles per class (spread over the batch; a tile with more
 * coincident points than a histogram leaf may hold is skipped), then the
 * model's table, stored in the context like owm_autotune's winners. */
owm_status owm_autotune_stats(owm_ctx *ctx, const double *x, const double *y, const double *z,
                              int device_ptrs, const owm_tile *tiles, uint32_t n_tiles,
                              owm_tune_entry *out, uint32_t cap, uint32_t *n_out);

/* ---- stage-level API on a device-resident tree ---------------------------
 * Mirrors the reference's internal boundary (include/optim3_func.hpp:85-93):
 * parallel_qtree + storeMinAndNumPoints -> owm_tree_build, stage1 ->
 * owm_stage1, sort + stage2 -> owm_stage2, stage3 -> owm_stage3.  The tree
 * depends only on the cloud and the leaf size, so one tree serves any number
 * of parameter re-queries. */
owm_status owm_tree_build(owm_ctx *ctx, const double *x, const double *y, const double *z,
                          uint64_t n, int device_ptrs, uint32_t leaf_size, owm_tree **out);
void owm_tree_free(owm_tree *tree);
uint64_t owm_tree_points(const owm_tree *tree);
/* Tree statistics for checks: counts of leaves and internal nodes, the root's
 * point count and bbox (minx, maxx, miny, maxy). */
owm_status owm_tree_info(const owm_tree *tree, uint64_t *n_leaves, uint64_t *n_internal,
                         uint64_t *root_count, double *root_bbox);

/* The whole structure of a resident tree, for checks in the spirit of the
 * reference's checkNumPoints (CPP-TBB-CPU/src/optim3_func.cpp:270-283, called
 * at application/owm_optim3.cpp:90), which walks its quadtree to compare
 * counts.  owm_tree_read fills the scalars and copies every array whose
 * pointer is non-null (caller-owned; sizes n, n_leaves, n_internal).  Points
 * sit in Morton order: a point's sort key is the top key_bits bits of the
 * 64-bit interleave of qx (even bits) and qy (odd bits), where
 * q = (uint32)((v - qmin) * qscale) in IEEE double, 0 when not > 0 (or NaN)
 * and 0xFFFFFFFF from 4294967295 up.  (qmin, qscale) is the square over a
 * sample of the cloud when the library took the bbox itself (128 runs of 16
 * consecutive points starting at r*n/128; every point up to 2048 of them),
 * else over the given bbox; qscale maps its side to just under 2^32.  Leaf j
 * spans tree positions
 * [leaf_start, leaf_start + leaf_points), leaf_size of them but the last. */
#define OWM_TREE_LEAF 0x80000000u /* child reference bit: a leaf, not an internal node */
typedef struct owm_tree_view {
    uint64_t n, n_leaves, n_internal;
    uint32_t leaf_size, key_bits;
    uint32_t root;          /* internal node index, or OWM_TREE_LEAF | leaf index */
    double qmin[2], qscale; /* Morton quantisation of x and y */
    uint32_t *ids;          /* [n] input index of the point at each tree position */
    double *xs, *ys, *zs;   /* [n] its coordinates */
    uint32_t *leaf_key;     /* [n_leaves] sort key of the leaf's first point */
    uint32_t *leaf_start;   /* [n_leaves] first tree position */
    uint32_t *leaf_points;  /* [n_leaves] tree positions the leaf spans */
    uint32_t *leaf_count;   /* [n_leaves] its points with non-NaN x and y */
    double *leaf_bbox;      /* [4 n_leaves] minx, maxx, miny, maxy of those */
    uint32_t *leaf_min;     /* [n_leaves] input index of the lowest point with z < DBL_MAX
                               (lowest z, then lowest index); 0xFFFFFFFF if none */
    uint32_t *node_child;   /* [2 n_internal] left and right child references */
    uint32_t *node_count;   /* [n_internal] as leaf_count, over the node's leaves */
    double *node_bbox;      /* [4 n_internal] */
    uint32_t *node_min;     /* [n_internal] */
} owm_tree_view;
owm_status owm_tree_read(const owm_tree *tree, owm_tree_view *view);

/* Stage 1: min_ids[jj*ncols + ii] = index of the lowest point of window
 * (ii, jj) if it holds at least minNumPoints points, else 0xFFFFFFFF
 * (the reference's -1).  cap >= ncols*nrows. */
owm_status owm_stage1(owm_ctx *ctx, owm_tree *tree, const owm_params *params,
                      uint32_t *min_ids, uint64_t cap, owm_derived *derived);
/* Stage 2 (sort + run length): ids that are the minimum of >= 2 windows
 * (all valid minima when overlap == 0), ascending. */
owm_status owm_stage2(owm_ctx *ctx, owm_tree *tree, const owm_params *params,
                      const uint32_t *min_ids, uint64_t n_ids, uint32_t *llp, uint64_t cap,
                      uint64_t *n_llp);
/* Stage 3: for every Bsize cell with no LLP strictly inside, the lowest
 * point strictly inside it, if any (cell order). */
owm_status owm_stage3(owm_ctx *ctx, owm_tree *tree, const owm_params *params,
                      const uint32_t *llp, uint64_t n_llp, uint32_t *adds, uint64_t cap,
                      uint64_t *n_adds);
/* Stages 1-3 on a resident tree; seeds as owm_run_tile. */
owm_status owm_tree_run(owm_ctx *ctx, owm_tree *tree, const owm_params *params,
                        uint32_t *seeds, uint64_t cap, uint64_t *n_seeds);

/* ---- serving: continuous batching with HBM-budget admission --------------
 * A background thread admits submitted tiles while device memory and
 * in-flight slots allow, runs each admitted group as one batch on its own
 * stream (groups overlap on the device), and keeps trees of tiles submitted
 * with a nonzero tile_key resident, so re-queries of the same key with other
 * parameters skip the upload and the tree build.  Input arrays must stay
 * valid until the request completes. */
owm_status owm_serve_start(owm_ctx *ctx, uint32_t max_inflight, uint32_t max_batch_tiles);
owm_status owm_serve_submit(owm_ctx *ctx, const double *x, const double *y, const double *z,
                            uint64_t n, const owm_params *params, uint64_t tile_key,
                            uint64_t *req_id);
/* The same with x/y/z already in device memory (HBM) of the context's device:
 * a served tile then never crosses PCIe (host inputs are bounded by it at 24
 * bytes per point).  Requests that are consecutive slices of one cloud
 * (tiles of a resident survey submitted in order) are batched in place;
 * others are copied within HBM into their group's buffer.  The arrays must
 * stay valid and unchanged until the request completes. */
owm_status owm_serve_submit_device(owm_ctx *ctx, const double *d_x, const double *d_y,
                                   const double *d_z, uint64_t n, const owm_params *params,
                                   uint64_t tile_key, uint64_t *req_id);
/* OWM_E_NOTREADY while running; on completion copies the seeds and returns
 * the request's own status.  latency_s: submit -> seeds ready. */
owm_status owm_serve_poll(owm_ctx *ctx, uint64_t req_id, uint32_t *seeds, uint64_t cap,
                          uint64_t *n_seeds, double *latency_s);
owm_status owm_serve_wait(owm_ctx *ctx, uint64_t req_id, uint32_t *seeds, uint64_t cap,
                          uint64_t *n_seeds, double *latency_s);
owm_status owm_serve_cancel(owm_ctx *ctx, uint64_t req_id);
/* Drops the resident tree of tile_key (if any). */
owm_status owm_serve_evict(owm_ctx *ctx, uint64_t tile_key);
owm_status owm_serve_stop(owm_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* OWM_H */
