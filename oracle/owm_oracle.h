/*
 * CPU oracle for the OWM ground-seed pipeline -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load this library, and only as the checker or the timed CPU baseline.
 * The product path (fastowm_amd) never links or calls it.
 *
 * Semantics (restated from CPP-TBB-CPU, the exact `o3memoB` build:
 * application/owm_optim3.cpp:92-132,143-230 and src/optim3_func.cpp:541-662)
 * with two deliberate choices that make the result a function of the input
 * alone:
 *   - the window/cell minimum is the point with the lowest z, ties broken by
 *     the lowest point index (the reference breaks ties by tree insertion or
 *     traversal order);
 *   - counts are exact counts of points strictly inside the open box (the
 *     reference adds whole-node aggregates when a node's closed geometric box
 *     lies inside the window, src/optim3_func.cpp:426-436, which can differ for
 *     points lying exactly on a node edge).
 * Point ids are 0-based indices into the caller's arrays (reference id - 1).
 */
#ifndef OWM_ORACLE_H
#define OWM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_NONE 0xFFFFFFFFu

typedef struct {
    /* inputs */
    uint32_t wsize;     /* -W */
    uint32_t bsize;     /* -B (0 = no stage 3) */
    double overlap;     /* -O */
    uint32_t leaf_size; /* -s maxNumber */
    int32_t level;      /* -L parallel tree level */
    double minx, maxx, miny, maxy;
    /* derived (oracle_derive) */
    double width, height, density, displace;
    uint32_t min_num_points;
    uint32_t ncols, nrows, ncolsg, nrowsg;
    int32_t h, hb;
    double initx, inity, initxg, inityg;
} oracle_params;

int oracle_derive(oracle_params *p, uint64_t n);

void *oracle_tree_build(const double *x, const double *y, const double *z, uint64_t n,
                        const oracle_params *p);
void oracle_tree_free(void *tree);
uint64_t oracle_tree_nodes(const void *tree);
/* checkNumPoints-style structural check: returns 0 when every node's count,
 * bbox and min agree with its children / points. */
int oracle_tree_check(const void *tree);

void oracle_stage1(const void *tree, const oracle_params *p, uint32_t *min_ids);
/* stage1 over window columns [col0, col1) only, same indices as oracle_stage1
 * (the CPU share of a column split of the windows). */
void oracle_stage1_cols(const void *tree, const oracle_params *p, uint32_t col0, uint32_t col1,
                        uint32_t *min_ids);
uint32_t oracle_stage2(const oracle_params *p, uint32_t *min_ids, uint32_t nwin, uint32_t *llp);
uint32_t oracle_stage3(const void *tree, const oracle_params *p, const uint32_t *llp,
                       uint32_t nllp, uint32_t *adds);
int oracle_run(const double *x, const double *y, const double *z, uint64_t n,
               oracle_params *p, uint32_t *seeds, uint64_t cap, uint64_t *nseeds,
               double *times4);

/* Brute force (no tree), O(N * windows): pins the tree-based functions. */
void oracle_stage1_brute(const double *x, const double *y, const double *z, uint64_t n,
                         const oracle_params *p, uint32_t *min_ids);
uint32_t oracle_stage3_brute(const double *x, const double *y, const double *z, uint64_t n,
                             const oracle_params *p, const uint32_t *llp, uint32_t nllp,
                             uint32_t *adds);

void oracle_set_threads(int n);
int oracle_get_threads(void);

#ifdef __cplusplus
}
#endif
/* The reference reader's point loop on a text held in memory (strtod per
 * number, rest of line skipped, blank lines skipped); -1 on a short line. */
int64_t oracle_parse_xyz(const char *text, uint64_t bytes, double *x, double *y, double *z,
                         uint64_t cap);

/* Leaf statistics of the reference's parallel_qtree (owm_stats): policy 0 =
 * maxNumber (statsmaxnum), 1 = minRadius (stats); the box is the cloud's
 * (minx, maxx, miny, maxy) as readXYZfile gives it.  Writes the distinct
 * (depth, points) pairs of the leaves, ascending, with their leaf counts, up to
 * cap entries, and returns how many there are; -1 for bad arguments, -2 when a
 * split would go below depth max_depth. */
int64_t oracle_qtree_stats(const double *x, const double *y, uint64_t n, double minx, double maxx,
                           double miny, double maxy, int policy, uint32_t max_number,
                           float min_radius, int level, int max_depth, uint32_t *cls_level,
                           uint32_t *cls_count, uint64_t *cls_leaves, uint64_t cap);

#endif
