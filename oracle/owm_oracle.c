/*
 * CPU oracle for the OWM ground-seed pipeline -- TEST INFRASTRUCTURE ONLY
 * (see owm_oracle.h for who may use it and for the exact semantics).
 *
 * Built with -ffp-contract=off: every window edge is computed with the same
 * sequence of IEEE double operations as the reference, without fused
 * multiply-add.
 */
#include "owm_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* round2d: src/optim3_func.cpp:1-3 */
static double round2d(double z) { return round(z * 100.0) / 100.0; }

/* Parameter derivation: application/owm_optim3.cpp:92-132 (types of the
 * reference: u32 Wsize/Bsize, ushort nCols/nRows/nColsg/nRowsg, u32 then
 * ushort minNumPoints, stage1's int Wsize/2 and stage3's int Bsize/2).
 * Values the reference would wrap or divide by zero on are rejected (1). */
int oracle_derive(oracle_params *p, uint64_t n)
{
    if (p->wsize == 0 || p->wsize > 65535u || p->bsize > 65535u)
        return 1;
    if (!(p->overlap >= 0.0 && p->overlap < 1.0))
        return 1;
    double width = round2d(p->maxx - p->minx);
    double height = round2d(p->maxy - p->miny);
    if (!(width > 0.0) || !(height > 0.0) || !isfinite(width) || !isfinite(height))
        return 1;
    p->width = width;
    p->height = height;
    p->density = (double)n / (width * height);
    double mnp = 0.5 * p->density * p->wsize * p->wsize;
    if (!(mnp < 4294967296.0))
        return 1;
    p->min_num_points = (uint16_t)(uint32_t)mnp; /* u32, then ushort argument */
    p->displace = round2d(p->wsize * (1 - p->overlap));
    if (!(p->displace > 0.0))
        return 1;
    double c, r;
    if (p->overlap > 0.0) {
        c = round((width + 2 * p->wsize * p->overlap) / p->displace) - 1;
        r = round((height + 2 * p->wsize * p->overlap) / p->displace) - 1;
    } else {
        c = floor(width / p->wsize) + 1;
        r = floor(height / p->wsize) + 1;
    }
    if (!(c >= 1.0 && c <= 65535.0 && r >= 1.0 && r <= 65535.0))
        return 1;
    p->ncols = (uint32_t)c;
    p->nrows = (uint32_t)r;
    if (p->bsize > 0) {
        double cg = floor(width / p->bsize) + 1, rg = floor(height / p->bsize) + 1;
        if (!(cg >= 1.0 && cg <= 65535.0 && rg >= 1.0 && rg <= 65535.0))
            return 1;
        p->ncolsg = (uint32_t)cg;
        p->nrowsg = (uint32_t)rg;
    } else {
        p->ncolsg = p->nrowsg = 0;
    }
    p->h = (int32_t)(p->wsize / 2);
    p->hb = (int32_t)(p->bsize / 2);
    /* stage1: initX = min.x - Wsize/2 + Displace (src/optim3_func.cpp:546-547) */
    p->initx = p->minx - p->h + p->displace;
    p->inity = p->miny - p->h + p->displace;
    /* stage3: initX = min.x + Bsize/2 (src/optim3_func.cpp:630-631) */
    p->initxg = p->minx + p->hb;
    p->inityg = p->miny + p->hb;
    if (p->leaf_size == 0)
        p->leaf_size = 32;
    if (p->level < 0)
        p->level = 0;
    return 0;
}

/* ---------------------------------------------------------------- tree ---- */

/* Quadtree node: Qtree_t (include/optim3_func.hpp:44-57) plus a tight bbox of
 * the node's points, which is what makes the whole-node shortcut exact. */
typedef struct {
    int32_t child; /* first of 4 consecutive children, -1 for a leaf */
    double cx, cy;
    float r;
    uint32_t *pts;
    uint32_t npts, cap;
    uint32_t cnt; /* points with non-NaN x and y */
    uint32_t min; /* lowest (z, index) among eligible points, ORACLE_NONE if none */
    double bx0, by0, bx1, by1;
} onode;

typedef struct {
    onode *nodes;
    int32_t n, cap;
} pool;

typedef struct {
    onode *nodes;
    int64_t nnodes;
    const double *x, *y, *z;
    uint64_t npts;
} otree;

#define MIN_SPLIT_R 1e-6f

static int32_t pool_new(pool *P, double cx, double cy, float r)
{
    if (P->n == P->cap) {
        P->cap = P->cap ? P->cap * 2 : 64;
        P->nodes = (onode *)realloc(P->nodes, sizeof(onode) * (size_t)P->cap);
    }
    onode *nd = &P->nodes[P->n];
    memset(nd, 0, sizeof(*nd));
    nd->child = -1;
    nd->cx = cx;
    nd->cy = cy;
    nd->r = r;
    nd->min = ORACLE_NONE;
    return P->n++;
}

/* createQuadrants: src/optim3_func.cpp:59-73 (float radius arithmetic). */
static void create_quadrants(pool *P, int32_t node)
{
    double cx = P->nodes[node].cx, cy = P->nodes[node].cy;
    float r = P->nodes[node].r;
    float nr = (float)(r * 0.5);
    int32_t first = -1;
    for (int i = 0; i < 4; ++i) {
        double ncx = cx + (double)(r * ((i & 2) ? 0.5f : -0.5f));
        double ncy = cy + (double)(r * ((i & 1) ? 0.5f : -0.5f));
        int32_t c = pool_new(P, ncx, ncy, nr);
        if (i == 0)
            first = c;
    }
    P->nodes[node].child = first;
}

/* quadrantIdx: src/optim3_func.cpp:91-100 */
static inline int quadrant(const onode *nd, double px, double py)
{
    return (px >= nd->cx ? 2 : 0) | (py >= nd->cy ? 1 : 0);
}

static void push_pt(onode *nd, uint32_t p)
{
    if (nd->npts == nd->cap) {
        nd->cap = nd->cap ? nd->cap * 2 : 8;
        nd->pts = (uint32_t *)realloc(nd->pts, sizeof(uint32_t) * nd->cap);
    }
    nd->pts[nd->npts++] = p;
}

/* insertPointMaxNum: src/optim3_func.cpp:160-184.  A full leaf splits and
 * re-inserts its points; leaves below MIN_SPLIT_R stop splitting so that more
 * than maxNumber coincident points cannot recurse forever. */
static void insert_maxnum(pool *P, int32_t node, uint32_t p, const double *x, const double *y,
                          uint32_t maxnum)
{
    for (;;) {
        onode *nd = &P->nodes[node];
        if (nd->child < 0) {
            if (nd->npts < maxnum || nd->r < MIN_SPLIT_R) {
                push_pt(nd, p);
                return;
            }
            uint32_t *old = nd->pts;
            uint32_t nold = nd->npts;
            nd->pts = NULL;
            nd->npts = nd->cap = 0;
            create_quadrants(P, node);
            for (uint32_t k = 0; k < nold; ++k) {
                onode *cur = &P->nodes[node];
                int32_t c = cur->child + quadrant(cur, x[old[k]], y[old[k]]);
                insert_maxnum(P, c, old[k], x, y, maxnum);
            }
            free(old);
            continue;
        }
        node = nd->child + quadrant(nd, x[p], y[p]);
    }
}

static inline int eligible(double zv) { return zv < DBL_MAX; }

static inline uint32_t better(uint32_t a, uint32_t b, const double *z)
{
    if (a == ORACLE_NONE)
        return b;
    if (b == ORACLE_NONE)
        return a;
    if (z[b] < z[a] || (z[b] == z[a] && b < a))
        return b;
    return a;
}

/* storeMinAndNumPoints: src/optim3_func.cpp:255-307, extended with the tight
 * bbox and the (z, index) tie rule. */
static void aggregate(onode *nodes, int32_t node, const double *x, const double *y, const double *z)
{
    onode *nd = &nodes[node];
    nd->cnt = 0;
    nd->min = ORACLE_NONE;
    nd->bx0 = nd->by0 = INFINITY;
    nd->bx1 = nd->by1 = -INFINITY;
    if (nd->child < 0) {
        for (uint32_t k = 0; k < nd->npts; ++k) {
            uint32_t p = nd->pts[k];
            if (isnan(x[p]) || isnan(y[p]))
                continue;
            nd->cnt++;
            if (x[p] < nd->bx0) nd->bx0 = x[p];
            if (x[p] > nd->bx1) nd->bx1 = x[p];
            if (y[p] < nd->by0) nd->by0 = y[p];
            if (y[p] > nd->by1) nd->by1 = y[p];
            if (eligible(z[p]))
                nd->min = better(nd->min, p, z);
        }
        return;
    }
    for (int i = 0; i < 4; ++i) {
        aggregate(nodes, nd->child + i, x, y, z);
        const onode *c = &nodes[nd->child + i];
        if (c->cnt == 0)
            continue;
        nd->cnt += c->cnt;
        if (c->bx0 < nd->bx0) nd->bx0 = c->bx0;
        if (c->bx1 > nd->bx1) nd->bx1 = c->bx1;
        if (c->by0 < nd->by0) nd->by0 = c->by0;
        if (c->by1 > nd->by1) nd->by1 = c->by1;
        nd->min = better(nd->min, c->min, z);
    }
}

/* Aggregates of the eagerly built top levels; nodes at depth `level` are
 * subtree roots that were aggregated while their subtree was filled. */
static void agg_top(onode *all, int32_t node, int depth, int level, const double *z)
{
    onode *a = &all[node];
    if (depth >= level || a->child < 0)
        return;
    a->cnt = 0;
    a->min = ORACLE_NONE;
    a->bx0 = a->by0 = INFINITY;
    a->bx1 = a->by1 = -INFINITY;
    for (int i = 0; i < 4; ++i) {
        agg_top(all, a->child + i, depth + 1, level, z);
        const onode *c = &all[a->child + i];
        if (c->cnt == 0)
            continue;
        a->cnt += c->cnt;
        if (c->bx0 < a->bx0) a->bx0 = c->bx0;
        if (c->bx1 > a->bx1) a->bx1 = c->bx1;
        if (c->by0 < a->by0) a->by0 = c->by0;
        if (c->by1 > a->by1) a->by1 = c->by1;
        a->min = better(a->min, c->min, z);
    }
}

static void build_top(pool *P, int32_t node, int lev, int maxlev, int32_t *work, int32_t *nwork)
{
    if (lev < maxlev) {
        create_quadrants(P, node);
        int32_t c = P->nodes[node].child;
        for (int i = 0; i < 4; ++i)
            build_top(P, c + i, lev + 1, maxlev, work, nwork);
    } else {
        work[(*nwork)++] = node;
    }
}

/* parallel_qtree: src/optim3_func.cpp:218-253.  The top `level` levels are
 * created eagerly, points are binned to those subtrees in index order, and the
 * subtrees are filled in parallel with the maxNumber policy.  Insertion order
 * within a subtree is ascending index, so the tree is deterministic. */
void *oracle_tree_build(const double *x, const double *y, const double *z, uint64_t n,
                        const oracle_params *p)
{
    int level = p->level > 8 ? 8 : p->level;
    /* getRadius / getCenter: src/optim3_func.cpp:6-26 (float maxRadius) */
    double rx = (p->maxx - p->minx) / 2.0, ry = (p->maxy - p->miny) / 2.0;
    float maxr = (float)(rx >= ry ? rx : ry);
    pool top = {0};
    int32_t root = pool_new(&top, p->minx + rx, p->miny + ry, maxr);
    int32_t nsub_max = 1;
    for (int l = 0; l < level; ++l)
        nsub_max *= 4;
    int32_t *work = (int32_t *)malloc(sizeof(int32_t) * nsub_max);
    int32_t nwork = 0;
    build_top(&top, root, 0, level, work, &nwork);

    /* route every point to its top-level subtree */
    uint32_t *sub = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    int32_t *leafslot = (int32_t *)malloc(sizeof(int32_t) * top.n);
    for (int32_t i = 0; i < top.n; ++i)
        leafslot[i] = -1;
    for (int32_t w = 0; w < nwork; ++w)
        leafslot[work[w]] = w;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        int32_t nd = root;
        while (top.nodes[nd].child >= 0)
            nd = top.nodes[nd].child + quadrant(&top.nodes[nd], x[i], y[i]);
        sub[i] = (uint32_t)leafslot[nd];
    }
    /* counting sort of point ids by subtree (stable, so ascending ids) */
    uint64_t *off = (uint64_t *)calloc((size_t)nwork + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i)
        off[sub[i] + 1]++;
    for (int32_t w = 0; w < nwork; ++w)
        off[w + 1] += off[w];
    uint32_t *ord = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)nwork + 1));
    memcpy(cur, off, sizeof(uint64_t) * ((size_t)nwork + 1));
    for (uint64_t i = 0; i < n; ++i)
        ord[cur[sub[i]]++] = (uint32_t)i;
    free(cur);
    free(sub);

    /* fill the subtrees in parallel, each in its own pool */
    pool *pools = (pool *)calloc((size_t)nwork, sizeof(pool));
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t w = 0; w < nwork; ++w) {
        const onode *t = &top.nodes[work[w]];
        pool *P = &pools[w];
        int32_t r0 = pool_new(P, t->cx, t->cy, t->r);
        for (uint64_t k = off[w]; k < off[w + 1]; ++k)
            insert_maxnum(P, r0, ord[k], x, y, p->leaf_size);
        aggregate(P->nodes, r0, x, y, z);
    }
    free(ord);
    free(off);

    /* concatenate: top nodes first, then each subtree pool */
    int64_t total = top.n;
    int64_t *base = (int64_t *)malloc(sizeof(int64_t) * (nwork ? nwork : 1));
    for (int32_t w = 0; w < nwork; ++w) {
        base[w] = total - 1; /* pool node 0 replaces the top work node */
        total += pools[w].n - 1;
    }
    onode *all = (onode *)malloc(sizeof(onode) * (size_t)total);
    memcpy(all, top.nodes, sizeof(onode) * (size_t)top.n);
    for (int32_t w = 0; w < nwork; ++w) {
        pool *P = &pools[w];
        for (int32_t k = 0; k < P->n; ++k) {
            onode nd = P->nodes[k];
            if (nd.child >= 0)
                nd.child = (int32_t)(base[w] + nd.child);
            int64_t dst = k == 0 ? work[w] : base[w] + k;
            all[dst] = nd;
        }
        free(P->nodes);
    }
    free(pools);
    free(base);
    free(work);
    free(leafslot);
    free(top.nodes);

    /* aggregates of the top levels (subtree roots are already done) */
    otree *T = (otree *)malloc(sizeof(otree));
    T->nodes = all;
    T->nnodes = total;
    T->x = x;
    T->y = y;
    T->z = z;
    T->npts = n;
    agg_top(all, root, 0, level, z);
    return T;
}

void oracle_tree_free(void *tree)
{
    otree *T = (otree *)tree;
    if (!T)
        return;
    for (int64_t i = 0; i < T->nnodes; ++i)
        free(T->nodes[i].pts);
    free(T->nodes);
    free(T);
}

uint64_t oracle_tree_nodes(const void *tree) { return (uint64_t)((const otree *)tree)->nnodes; }

/* checkNumPoints (src/optim3_func.cpp:270-283), extended to bbox and min. */
int oracle_tree_check(const void *tree)
{
    const otree *T = (const otree *)tree;
    uint64_t seen = 0, finite = 0;
    for (uint64_t i = 0; i < T->npts; ++i)
        finite += !(isnan(T->x[i]) || isnan(T->y[i]));
    for (int64_t i = 0; i < T->nnodes; ++i) {
        const onode *nd = &T->nodes[i];
        uint32_t cnt = 0, mn = ORACLE_NONE;
        if (nd->child < 0) {
            seen += nd->npts;
            for (uint32_t k = 0; k < nd->npts; ++k) {
                uint32_t p = nd->pts[k];
                if (isnan(T->x[p]) || isnan(T->y[p]))
                    continue;
                cnt++;
                if (eligible(T->z[p]))
                    mn = better(mn, p, T->z);
            }
        } else {
            for (int c = 0; c < 4; ++c) {
                cnt += T->nodes[nd->child + c].cnt;
                mn = better(mn, T->nodes[nd->child + c].min, T->z);
            }
        }
        if (cnt != nd->cnt || mn != nd->min)
            return 1;
    }
    if (seen != T->npts || T->nodes[0].cnt != finite)
        return 2;
    return 0;
}

typedef struct {
    uint32_t cnt;
    uint32_t best;
} qres;

/* findValidMin / findValidMinMemo (src/optim3_func.cpp:390-478) with exact
 * whole-node tests on the tight bbox: a node whose points all lie strictly
 * inside the open box contributes its aggregate, a node whose points all lie
 * outside is skipped, anything else is descended point by point. */
static void query(const otree *T, int32_t node, double bx0, double by0, double bx1, double by1,
                  qres *r)
{
    const onode *nd = &T->nodes[node];
    if (nd->cnt == 0)
        return;
    if (nd->bx1 <= bx0 || nd->bx0 >= bx1 || nd->by1 <= by0 || nd->by0 >= by1)
        return;
    if (nd->bx0 > bx0 && nd->bx1 < bx1 && nd->by0 > by0 && nd->by1 < by1) {
        r->cnt += nd->cnt;
        r->best = better(r->best, nd->min, T->z);
        return;
    }
    if (nd->child < 0) {
        const double *x = T->x, *y = T->y, *z = T->z;
        for (uint32_t k = 0; k < nd->npts; ++k) {
            uint32_t p = nd->pts[k];
            /* insideBox2D: src/optim3_func.cpp:326-332 (strict) */
            if (x[p] > bx0 && y[p] > by0 && x[p] < bx1 && y[p] < by1) {
                r->cnt++;
                if (eligible(z[p]))
                    r->best = better(r->best, p, z);
            }
        }
        return;
    }
    for (int i = 0; i < 4; ++i)
        query(T, nd->child + i, bx0, by0, bx1, by1, r);
}

/* Window (ii, jj) of stage1 (src/optim3_func.cpp:541-561): centre
 * (initX + ii*Displace, initY + jj*Displace), half side Wsize/2, boxes from
 * makeBox (src/optim3_func.cpp:317-324). */
static inline void window_box(const oracle_params *p, uint32_t ii, uint32_t jj, double *b)
{
    double cx = p->initx + (int)ii * p->displace;
    double cy = p->inity + (int)jj * p->displace;
    double hx = (double)(float)p->h;
    b[0] = cx - hx;
    b[1] = cy - hx;
    b[2] = cx + hx;
    b[3] = cy + hx;
}

/* Stage-3 cell (ii, jj) (src/optim3_func.cpp:627-662): centre
 * (initX + ii*Bsize, initY + jj*Bsize) with the integer product, half side
 * Bsize/2. */
static inline void cell_box(const oracle_params *p, uint32_t ii, uint32_t jj, double *b)
{
    double cx = p->initxg + (int)(ii * p->bsize);
    double cy = p->inityg + (int)(jj * p->bsize);
    double hb = (double)(float)p->hb;
    b[0] = cx - hb;
    b[1] = cy - hb;
    b[2] = cx + hb;
    b[3] = cy + hb;
}

/* stage1 (exact, -DNOMEMOA): min_ids[jj*ncols + ii] = min id when the window
 * holds at least minNumPoints points and an eligible minimum, else NONE.
 * (The reference stores the sentinel id 0 for a valid window without any
 * eligible point; that id is not a point, so it is never emitted here.) */
void oracle_stage1(const void *tree, const oracle_params *p, uint32_t *min_ids)
{
    const otree *T = (const otree *)tree;
    int64_t W = (int64_t)p->ncols * p->nrows;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t s = 0; s < W; ++s) {
        uint32_t ii = (uint32_t)(s % p->ncols), jj = (uint32_t)(s / p->ncols);
        double b[4];
        window_box(p, ii, jj, b);
        qres r = {0, ORACLE_NONE};
        query(T, 0, b[0], b[1], b[2], b[3], &r);
        min_ids[s] = (r.cnt >= p->min_num_points && r.best != ORACLE_NONE) ? r.best : ORACLE_NONE;
    }
}

/* stage1 over the window columns [col0, col1) only (every row), written at
 * the same indices as oracle_stage1: the CPU share of a column split of the
 * windows between devices (SYCL-CUDA-GPU/octree/octree_traverse.inl:119-180
 * gives wCols = nCols*factor columns to one side, basic/traverse.h:916-1037
 * the rest to the TBB query). */
void oracle_stage1_cols(const void *tree, const oracle_params *p, uint32_t col0, uint32_t col1,
                        uint32_t *min_ids)
{
    const otree *T = (const otree *)tree;
    if (col1 > p->ncols)
        col1 = p->ncols;
    if (col0 >= col1)
        return;
    const int64_t w = col1 - col0, W = w * (int64_t)p->nrows;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t s = 0; s < W; ++s) {
        uint32_t ii = col0 + (uint32_t)(s % w), jj = (uint32_t)(s / w);
        double b[4];
        window_box(p, ii, jj, b);
        qres r = {0, ORACLE_NONE};
        query(T, 0, b[0], b[1], b[2], b[3], &r);
        min_ids[(int64_t)jj * p->ncols + ii] =
            (r.cnt >= p->min_num_points && r.best != ORACLE_NONE) ? r.best : ORACLE_NONE;
    }
}

static void radix_sort_u32(uint32_t *a, uint32_t n)
{
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (int shift = 0; shift < 32; shift += 8) {
        uint32_t cnt[257] = {0};
        for (uint32_t i = 0; i < n; ++i)
            cnt[((a[i] >> shift) & 255u) + 1]++;
        for (int d = 0; d < 256; ++d)
            cnt[d + 1] += cnt[d];
        for (uint32_t i = 0; i < n; ++i)
            tmp[cnt[(a[i] >> shift) & 255u]++] = a[i];
        memcpy(a, tmp, sizeof(uint32_t) * n);
    }
    free(tmp);
}

/* sort + stage2 (application/owm_optim3.cpp:158-188, src/optim3_func.cpp:610-625):
 * ids that are the minimum of at least two valid windows, ascending.  With
 * Overlap == 0 every valid minimum is kept (the reference leaves minIDs
 * uncompacted there; here the valid ids are compacted and de-duplicated). */
uint32_t oracle_stage2(const oracle_params *p, uint32_t *min_ids, uint32_t nwin, uint32_t *llp)
{
    radix_sort_u32(min_ids, nwin);
    uint32_t nl = 0;
    uint32_t need = p->overlap != 0.0 ? 2u : 1u;
    for (uint32_t i = 0; i < nwin;) {
        uint32_t id = min_ids[i], j = i + 1;
        while (j < nwin && min_ids[j] == id)
            ++j;
        if (id != ORACLE_NONE && j - i >= need)
            llp[nl++] = id;
        i = j;
    }
    return nl;
}

/* Index of the stage-3 cell that strictly contains (px, py), or -1. */
static int64_t cell_of(const oracle_params *p, double px, double py)
{
    if (!(px == px) || !(py == py))
        return -1;
    double fx = floor((px - p->minx) / p->bsize), fy = floor((py - p->miny) / p->bsize);
    if (!(fx > -4.0 && fx < 70000.0 && fy > -4.0 && fy < 70000.0))
        return -1;
    for (int dj = -1; dj <= 1; ++dj) {
        int64_t jj = (int64_t)fy + dj;
        if (jj < 0 || jj >= p->nrowsg)
            continue;
        for (int di = -1; di <= 1; ++di) {
            int64_t ii = (int64_t)fx + di;
            if (ii < 0 || ii >= p->ncolsg)
                continue;
            double b[4];
            cell_box(p, (uint32_t)ii, (uint32_t)jj, b);
            if (px > b[0] && py > b[1] && px < b[2] && py < b[3])
                return jj * p->ncolsg + ii;
        }
    }
    return -1;
}

/* stage3 (src/optim3_func.cpp:627-662, grid tree built at
 * application/owm_optim3.cpp:192-199): a cell with no LLP strictly inside
 * contributes the cloud's minimum strictly inside it, if any.  Output is in
 * cell order (row-major); the reference's parallel_reduce order varies. */
uint32_t oracle_stage3(const void *tree, const oracle_params *p, const uint32_t *llp,
                       uint32_t nllp, uint32_t *adds)
{
    const otree *T = (const otree *)tree;
    if (p->bsize == 0)
        return 0;
    int64_t G = (int64_t)p->ncolsg * p->nrowsg;
    uint8_t *has = (uint8_t *)calloc((size_t)G, 1);
    for (uint32_t k = 0; k < nllp; ++k) {
        int64_t c = cell_of(p, T->x[llp[k]], T->y[llp[k]]);
        if (c >= 0)
            has[c] = 1;
    }
    uint32_t *res = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)G);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t c = 0; c < G; ++c) {
        res[c] = ORACLE_NONE;
        if (has[c])
            continue;
        double b[4];
        cell_box(p, (uint32_t)(c % p->ncolsg), (uint32_t)(c / p->ncolsg), b);
        qres r = {0, ORACLE_NONE};
        query(T, 0, b[0], b[1], b[2], b[3], &r);
        if (r.cnt > 0)
            res[c] = r.best;
    }
    uint32_t na = 0;
    for (int64_t c = 0; c < G; ++c)
        if (res[c] != ORACLE_NONE)
            adds[na++] = res[c];
    free(res);
    free(has);
    return na;
}

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

/* The timed driver loop of application/owm_optim3.cpp:77-230: tree, stage 1,
 * sort + stage 2, stage 3, then the seed list (stage-3 adds ++ LLPs), returned
 * here sorted ascending.  times4 = {tree, s1, s2, s3} seconds. */
int oracle_run(const double *x, const double *y, const double *z, uint64_t n, oracle_params *p,
               uint32_t *seeds, uint64_t cap, uint64_t *nseeds, double *times4)
{
    if (oracle_derive(p, n))
        return 1;
    double t0 = omp_get_wtime();
    void *T = oracle_tree_build(x, y, z, n, p);
    double t1 = omp_get_wtime();
    uint32_t W = p->ncols * p->nrows;
    uint32_t *mins = (uint32_t *)malloc(sizeof(uint32_t) * W);
    oracle_stage1(T, p, mins);
    double t2 = omp_get_wtime();
    uint32_t *llp = (uint32_t *)malloc(sizeof(uint32_t) * W);
    uint32_t nl = oracle_stage2(p, mins, W, llp);
    double t3 = omp_get_wtime();
    uint32_t G = p->ncolsg * p->nrowsg;
    uint32_t *adds = (uint32_t *)malloc(sizeof(uint32_t) * (G ? G : 1));
    uint32_t na = oracle_stage3(T, p, llp, nl, adds);
    double t4 = omp_get_wtime();
    if (times4) {
        times4[0] = t1 - t0;
        times4[1] = t2 - t1;
        times4[2] = t3 - t2;
        times4[3] = t4 - t3;
    }
    *nseeds = (uint64_t)nl + na;
    int rc = 0;
    if (*nseeds > cap) {
        rc = 2;
    } else {
        memcpy(seeds, adds, sizeof(uint32_t) * na);
        memcpy(seeds + na, llp, sizeof(uint32_t) * nl);
        qsort(seeds, (size_t)(*nseeds), sizeof(uint32_t), cmp_u32);
    }
    free(adds);
    free(llp);
    free(mins);
    oracle_tree_free(T);
    return rc;
}

/* ------------------------------------------------------- brute force ---- */

void oracle_stage1_brute(const double *x, const double *y, const double *z, uint64_t n,
                         const oracle_params *p, uint32_t *min_ids)
{
    int64_t W = (int64_t)p->ncols * p->nrows;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t s = 0; s < W; ++s) {
        double b[4];
        window_box(p, (uint32_t)(s % p->ncols), (uint32_t)(s / p->ncols), b);
        uint32_t cnt = 0, best = ORACLE_NONE;
        for (uint64_t i = 0; i < n; ++i) {
            if (x[i] > b[0] && y[i] > b[1] && x[i] < b[2] && y[i] < b[3]) {
                cnt++;
                if (eligible(z[i]))
                    best = better(best, (uint32_t)i, z);
            }
        }
        min_ids[s] = (cnt >= p->min_num_points && best != ORACLE_NONE) ? best : ORACLE_NONE;
    }
}

uint32_t oracle_stage3_brute(const double *x, const double *y, const double *z, uint64_t n,
                             const oracle_params *p, const uint32_t *llp, uint32_t nllp,
                             uint32_t *adds)
{
    if (p->bsize == 0)
        return 0;
    uint32_t na = 0;
    for (uint32_t jj = 0; jj < p->nrowsg; ++jj)
        for (uint32_t ii = 0; ii < p->ncolsg; ++ii) {
            double b[4];
            cell_box(p, ii, jj, b);
            int has = 0;
            for (uint32_t k = 0; k < nllp && !has; ++k) {
                double px = x[llp[k]], py = y[llp[k]];
                has = px > b[0] && py > b[1] && px < b[2] && py < b[3];
            }
            if (has)
                continue;
            uint32_t cnt = 0, best = ORACLE_NONE;
            for (uint64_t i = 0; i < n; ++i)
                if (x[i] > b[0] && y[i] > b[1] && x[i] < b[2] && y[i] < b[3]) {
                    cnt++;
                    if (eligible(z[i]))
                        best = better(best, (uint32_t)i, z);
                }
            if (cnt > 0 && best != ORACLE_NONE)
                adds[na++] = best;
        }
    return na;
}

void oracle_set_threads(int n)
{
    if (n > 0)
        omp_set_num_threads(n);
}

int oracle_get_threads(void) { return omp_get_max_threads(); }

/* readXYZfile's point loop (src/optim3_func.cpp:709-716): three numbers per
 * line read as fscanf("%lf") reads them (strtod), the rest of the line
 * skipped; blank lines skipped.  Returns the points read (at most cap), or
 * -1 for a line with fewer than three numbers.  Single-threaded, like the
 * reference. */
int64_t oracle_parse_xyz(const char *text, uint64_t bytes, double *x, double *y, double *z,
                         uint64_t cap)
{
    const char *p = text, *end = text + bytes;
    uint64_t n = 0;
    while (p < end && n < cap) {
        const char *eol = memchr(p, '\n', (size_t)(end - p));
        const char *e = eol ? eol : end;
        const char *q = p;
        while (q < e && (*q == ' ' || *q == '\t' || *q == '\r' || *q == '\v' || *q == '\f'))
            ++q;
        if (q < e) {
            char buf[256];
            size_t len = (size_t)(e - p) < sizeof buf - 1 ? (size_t)(e - p) : sizeof buf - 1;
            memcpy(buf, p, len);
            buf[len] = 0;
            char *r = buf, *s2;
            double v[3];
            for (int k = 0; k < 3; ++k) {
                v[k] = strtod(r, &s2);
                if (s2 == r)
                    return -1;
                r = s2;
            }
            x[n] = v[0];
            y[n] = v[1];
            z[n] = v[2];
            ++n;
        }
        p = eol ? eol + 1 : end;
    }
    return (int64_t)n;
}

/* ---- quadtree leaf statistics (the reference's owm_stats) ----------------
 * A literal sequential restatement of parallel_qtree
 * (CPP-TBB-CPU/src/optim3_func.cpp:218-253) for both node policies, walked
 * the way owm_stats walks it (CPP-TBB-CPU/include/stats_func.hpp:31-56):
 *   - tree_phase1 (:126-139) creates every node of depths 0..level-1 with its
 *     four children; the depth-`level` nodes collect their points
 *     (insertPointConcurrent, :142-154), then each gets createQuadrants and
 *     its points are inserted into those children in index order (:243-248);
 *   - maxNumber (statsmaxnum, -DMAXNUMBER): insertPointMaxNum (:160-184), a
 *     full leaf (maxNumber points) splits and re-inserts its points when the
 *     next one arrives;
 *   - minRadius (stats): insertPoint (:103-123), a leaf whose radius * 0.5
 *     exceeds minRadius splits before it stores a point.
 * Child centres and radii follow createQuadrants (:59-73).  The reference
 * recurses without bound when more than maxNumber points coincide; here a
 * split below depth max_depth returns -2 instead. */
typedef struct {
    pool P;
    const double *x, *y;
    int policy; /* 0 maxNumber, 1 minRadius */
    uint32_t maxnum;
    float minr;
    int max_depth;
    int err;
} qsbuild;

static void qs_insert(qsbuild *B, int32_t node, int depth, uint32_t p);

static void qs_split(qsbuild *B, int32_t node, int depth)
{
    if (depth + 1 > B->max_depth) {
        B->err = 1;
        return;
    }
    create_quadrants(&B->P, node);
}

static void qs_insert(qsbuild *B, int32_t node, int depth, uint32_t p)
{
    if (B->err)
        return;
    onode *nd = &B->P.nodes[node];
    if (nd->child >= 0) { /* internal: descend */
        qs_insert(B, nd->child + quadrant(nd, B->x[p], B->y[p]), depth + 1, p);
        return;
    }
    if (B->policy == 0) {
        if (nd->npts < B->maxnum) {
            push_pt(nd, p);
            return;
        }
        uint32_t *old = nd->pts;
        uint32_t nold = nd->npts;
        nd->pts = NULL;
        nd->npts = nd->cap = 0;
        qs_split(B, node, depth);
        if (B->err) {
            free(old);
            return;
        }
        for (uint32_t k = 0; k < nold; ++k) {
            const onode *cur = &B->P.nodes[node];
            qs_insert(B, cur->child + quadrant(cur, B->x[old[k]], B->y[old[k]]), depth + 1, old[k]);
        }
        free(old);
        const onode *cur = &B->P.nodes[node];
        qs_insert(B, cur->child + quadrant(cur, B->x[p], B->y[p]), depth + 1, p);
    } else {
        if ((double)nd->r * 0.5 > (double)B->minr) {
            qs_split(B, node, depth);
            if (B->err)
                return;
            const onode *cur = &B->P.nodes[node];
            qs_insert(B, cur->child + quadrant(cur, B->x[p], B->y[p]), depth + 1, p);
        } else {
            push_pt(nd, p);
        }
    }
}

typedef struct {
    uint32_t level, count;
} qsleaf;

static int cmp_qsleaf(const void *a, const void *b)
{
    const qsleaf *u = (const qsleaf *)a, *v = (const qsleaf *)b;
    if (u->level != v->level)
        return u->level < v->level ? -1 : 1;
    return u->count < v->count ? -1 : u->count > v->count;
}

int64_t oracle_qtree_stats(const double *x, const double *y, uint64_t n, double minx, double maxx,
                           double miny, double maxy, int policy, uint32_t max_number,
                           float min_radius, int level, int max_depth, uint32_t *cls_level,
                           uint32_t *cls_count, uint64_t *cls_leaves, uint64_t cap)
{
    if ((policy == 0 && max_number == 0) || level < 0 || level > max_depth)
        return -1;
    qsbuild B;
    memset(&B, 0, sizeof B);
    B.x = x;
    B.y = y;
    B.policy = policy;
    B.maxnum = max_number;
    B.minr = min_radius;
    B.max_depth = max_depth;
    /* getRadius / getCenter: src/optim3_func.cpp:6-26 */
    double rx = (maxx - minx) / 2.0, ry = (maxy - miny) / 2.0;
    float maxr = (float)(rx >= ry ? rx : ry);
    int32_t root = pool_new(&B.P, minx + rx, miny + ry, maxr);
    /* tree_phase1: the depth-`level` nodes, in creation order */
    int64_t nwork = 1;
    for (int l = 0; l < level; ++l)
        nwork *= 4;
    int32_t *work = (int32_t *)malloc(sizeof(int32_t) * (size_t)nwork);
    int32_t nw = 0;
    build_top(&B.P, root, 0, level, work, &nw);
    /* insertPointConcurrent: the points of every depth-`level` node */
    int32_t *slot = (int32_t *)malloc(sizeof(int32_t) * (size_t)B.P.n);
    for (int32_t i = 0; i < B.P.n; ++i)
        slot[i] = -1;
    for (int32_t w = 0; w < nw; ++w)
        slot[work[w]] = w;
    uint32_t *sub = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint64_t *off = (uint64_t *)calloc((size_t)nw + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) {
        int32_t nd = root;
        while (B.P.nodes[nd].child >= 0)
            nd = B.P.nodes[nd].child + quadrant(&B.P.nodes[nd], x[i], y[i]);
        sub[i] = (uint32_t)slot[nd];
        off[sub[i] + 1]++;
    }
    for (int32_t w = 0; w < nw; ++w)
        off[w + 1] += off[w];
    uint32_t *ord = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)nw + 1));
    memcpy(cur, off, sizeof(uint64_t) * ((size_t)nw + 1));
    for (uint64_t i = 0; i < n; ++i)
        ord[cur[sub[i]]++] = (uint32_t)i;
    free(cur);
    free(sub);
    /* second phase: createQuadrants on each, then its points in order */
    for (int32_t w = 0; w < nw && !B.err; ++w) {
        qs_split(&B, work[w], level);
        for (uint64_t k = off[w]; k < off[w + 1] && !B.err; ++k) {
            const onode *t = &B.P.nodes[work[w]];
            qs_insert(&B, t->child + quadrant(t, x[ord[k]], y[ord[k]]), level + 1, ord[k]);
        }
    }
    free(ord);
    free(off);
    free(slot);
    free(work);
    int64_t ret = -2;
    if (!B.err) {
        /* leaves with their depth: iterative walk from the root */
        int32_t *stk = (int32_t *)malloc(sizeof(int32_t) * (size_t)B.P.n);
        uint32_t *dep = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)B.P.n);
        qsleaf *lv = (qsleaf *)malloc(sizeof(qsleaf) * (size_t)B.P.n);
        int64_t sp = 0, nl = 0;
        stk[sp] = root;
        dep[sp++] = 0;
        while (sp) {
            --sp;
            const int32_t nd = stk[sp];
            const uint32_t d = dep[sp];
            const onode *o = &B.P.nodes[nd];
            if (o->child < 0) {
                lv[nl].level = d;
                lv[nl++].count = o->npts;
            } else {
                for (int i = 0; i < 4; ++i) {
                    stk[sp] = o->child + i;
                    dep[sp++] = d + 1;
                }
            }
        }
        qsort(lv, (size_t)nl, sizeof(qsleaf), cmp_qsleaf);
        int64_t nc = 0;
        for (int64_t i = 0; i < nl;) {
            int64_t j = i + 1;
            while (j < nl && lv[j].level == lv[i].level && lv[j].count == lv[i].count)
                ++j;
            if ((uint64_t)nc < cap) {
                cls_level[nc] = lv[i].level;
                cls_count[nc] = lv[i].count;
                cls_leaves[nc] = (uint64_t)(j - i);
            }
            ++nc;
            i = j;
        }
        ret = nc;
        free(stk);
        free(dep);
        free(lv);
    }
    for (int32_t i = 0; i < B.P.n; ++i)
        free(B.P.nodes[i].pts);
    free(B.P.nodes);
    return ret;
}
