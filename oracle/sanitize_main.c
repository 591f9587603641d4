/*
 * Sanitizer driver for the CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Built by `make -C oracle sanitize` with gcc's AddressSanitizer and
 * UndefinedBehaviorSanitizer (every report fatal) and run by
 * tests/test_oracle_sanitize.py.  It drives every oracle entry point over
 * seeded clouds that hit the edge cases the parity tests use (grid-aligned
 * coordinates that fall on window and node edges, z ties, one-point and
 * two-point clouds, all leaf sizes the product accepts, one and several
 * threads) and cross-checks the tree-based stages against brute force, so
 * that a memory or arithmetic fault in the checker surfaces here rather than
 * as a wrong answer in a parity test.  Exit status 0 = clean and consistent.
 */
#include "owm_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t rng_state;

static double urand(void)
{
    /* splitmix64 */
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static int fails;

#define CHECK(cond, ...)                                                                 \
    do {                                                                                 \
        if (!(cond)) {                                                                   \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);                         \
            fprintf(stderr, __VA_ARGS__);                                                \
            fputc('\n', stderr);                                                         \
            ++fails;                                                                     \
        }                                                                                \
    } while (0)

static int cmp_u32(const void *a, const void *b)
{
    const uint32_t u = *(const uint32_t *)a, v = *(const uint32_t *)b;
    return u < v ? -1 : u > v;
}

/* One cloud through every stage: tree vs brute force, the assembled stages vs
 * oracle_run. */
static void one_cloud(const char *name, const double *x, const double *y, const double *z,
                      uint64_t n, uint32_t w, uint32_t b, double ov, uint32_t leaf)
{
    double minx = x[0], maxx = x[0], miny = y[0], maxy = y[0];
    for (uint64_t i = 1; i < n; ++i) {
        minx = x[i] < minx ? x[i] : minx;
        maxx = x[i] > maxx ? x[i] : maxx;
        miny = y[i] < miny ? y[i] : miny;
        maxy = y[i] > maxy ? y[i] : maxy;
    }
    oracle_params p;
    memset(&p, 0, sizeof p);
    p.wsize = w, p.bsize = b, p.overlap = ov, p.leaf_size = leaf, p.level = 5;
    p.minx = minx, p.maxx = maxx, p.miny = miny, p.maxy = maxy;
    if (oracle_derive(&p, n) != 0)
        return; /* rejected parameters: nothing to run */
    const uint32_t nwin = p.ncols * p.nrows;
    const uint32_t ngrid = p.ncolsg * p.nrowsg > 0 ? p.ncolsg * p.nrowsg : 1;
    uint32_t *m_tree = malloc(sizeof(uint32_t) * (nwin + 1));
    uint32_t *m_brute = malloc(sizeof(uint32_t) * (nwin + 1));
    uint32_t *llp = malloc(sizeof(uint32_t) * (nwin + 1));
    uint32_t *a_tree = malloc(sizeof(uint32_t) * ngrid);
    uint32_t *a_brute = malloc(sizeof(uint32_t) * ngrid);
    void *tree = oracle_tree_build(x, y, z, n, &p);
    CHECK(tree != NULL, "%s: tree build", name);
    if (!tree)
        goto out;
    CHECK(oracle_tree_check(tree) == 0, "%s: tree check", name);
    CHECK(oracle_tree_nodes(tree) >= 1, "%s: no nodes", name);
    oracle_stage1(tree, &p, m_tree);
    oracle_stage1_brute(x, y, z, n, &p, m_brute);
    CHECK(memcmp(m_tree, m_brute, sizeof(uint32_t) * nwin) == 0, "%s: stage 1 tree != brute",
          name);
    /* the column split: three column ranges give the whole stage 1 */
    memset(m_brute, 0xAB, sizeof(uint32_t) * nwin);
    oracle_stage1_cols(tree, &p, 0, p.ncols / 3, m_brute);
    oracle_stage1_cols(tree, &p, p.ncols / 3, p.ncols / 2, m_brute);
    oracle_stage1_cols(tree, &p, p.ncols / 2, p.ncols + 5, m_brute);
    CHECK(memcmp(m_tree, m_brute, sizeof(uint32_t) * nwin) == 0, "%s: stage 1 by columns", name);
    const uint32_t nllp = oracle_stage2(&p, m_tree, nwin, llp);
    CHECK(nllp <= nwin, "%s: stage 2 count %u > %u", name, nllp, nwin);
    uint32_t na = 0, nb = 0;
    if (p.bsize > 0) {
        na = oracle_stage3(tree, &p, llp, nllp, a_tree);
        nb = oracle_stage3_brute(x, y, z, n, &p, llp, nllp, a_brute);
        CHECK(na == nb && memcmp(a_tree, a_brute, sizeof(uint32_t) * na) == 0,
              "%s: stage 3 tree (%u) != brute (%u)", name, na, nb);
    }
    {
        /* the assembled stages equal oracle_run's seed list */
        const uint64_t cap = n + 16;
        uint32_t *seeds = malloc(sizeof(uint32_t) * cap);
        uint32_t *want = malloc(sizeof(uint32_t) * (nllp + na + 1));
        uint64_t ns = 0;
        double t4[4];
        oracle_params q = p;
        CHECK(oracle_run(x, y, z, n, &q, seeds, cap, &ns, t4) == 0, "%s: run rejected", name);
        memcpy(want, llp, sizeof(uint32_t) * nllp);
        memcpy(want + nllp, a_tree, sizeof(uint32_t) * na);
        qsort(want, nllp + na, sizeof(uint32_t), cmp_u32);
        uint64_t k = 0;
        for (uint64_t i = 0; i < (uint64_t)nllp + na; ++i)
            if (i == 0 || want[i] != want[i - 1])
                want[k++] = want[i];
        CHECK(ns == k && memcmp(seeds, want, sizeof(uint32_t) * k) == 0,
              "%s: run (%llu seeds) != stages (%llu)", name, (unsigned long long)ns,
              (unsigned long long)k);
        free(seeds);
        free(want);
    }
    oracle_tree_free(tree);
out:
    free(m_tree), free(m_brute), free(llp), free(a_tree), free(a_brute);
}

static void gen(uint64_t n, int kind, double *x, double *y, double *z)
{
    for (uint64_t i = 0; i < n; ++i) {
        double u = urand() * 60.0, v = urand() * 45.0;
        if (kind == 1) { /* 0.5 m grid: points on window and node edges */
            u = (double)(int)(u * 2.0) * 0.5;
            v = (double)(int)(v * 2.0) * 0.5;
        } else if (kind == 2 && i % 3 == 0) { /* a dense cluster */
            u = 20.0 + urand() * 2.0;
            v = 15.0 + urand() * 2.0;
        }
        x[i] = 500000.0 + u;
        y[i] = 4200000.0 + v;
        z[i] = kind == 1 ? (double)(int)(urand() * 4.0) : 300.0 + urand() * 10.0; /* z ties */
    }
}

static void reader(void)
{
    const char *text = "1.5 2.25 3\n\n-4e1 5.000000000000000001 6 7 8\n0.1 0.2 0.3";
    double x[4], y[4], z[4];
    const int64_t k = oracle_parse_xyz(text, strlen(text), x, y, z, 4);
    CHECK(k == 3, "reader: %lld points", (long long)k);
    if (k == 3)
        CHECK(x[0] == 1.5 && y[0] == 2.25 && z[0] == 3.0 && x[1] == -40.0 && z[2] == 0.3,
              "reader: values");
    const char *shortline = "1 2 3\n4 5\n";
    CHECK(oracle_parse_xyz(shortline, strlen(shortline), x, y, z, 4) == -1, "reader: short line");
    CHECK(oracle_parse_xyz(text, strlen(text), x, y, z, 1) >= 0, "reader: capped");
}

static void stats(const double *x, const double *y, uint64_t n)
{
    uint32_t lv[256], cnt[256];
    uint64_t leaves[256];
    double minx = x[0], maxx = x[0], miny = y[0], maxy = y[0];
    for (uint64_t i = 1; i < n; ++i) {
        minx = x[i] < minx ? x[i] : minx;
        maxx = x[i] > maxx ? x[i] : maxx;
        miny = y[i] < miny ? y[i] : miny;
        maxy = y[i] > maxy ? y[i] : maxy;
    }
    for (int policy = 0; policy < 2; ++policy) {
        const int64_t k = oracle_qtree_stats(x, y, n, minx, maxx, miny, maxy, policy, 32, 1.0f, 4,
                                             40, lv, cnt, leaves, 256);
        CHECK(k >= 1, "stats policy %d: %lld", policy, (long long)k);
        uint64_t pts = 0;
        for (int64_t i = 0; i < k && i < 256; ++i)
            pts += (uint64_t)cnt[i] * leaves[i];
        CHECK(k > 256 || pts == n, "stats policy %d: %llu of %llu points in leaves", policy,
              (unsigned long long)pts, (unsigned long long)n);
    }
    CHECK(oracle_qtree_stats(x, y, n, minx, maxx, miny, maxy, 0, 0, 1.0f, 4, 40, lv, cnt, leaves,
                             256) == -1,
          "stats: maxNumber 0 accepted");
}

int main(void)
{
    static const uint32_t leaves[] = {16, 32, 64, 128};
    const uint64_t sizes[] = {1, 2, 37, 900, 6000};
    for (int th = 1; th <= 4; th += 3) {
        oracle_set_threads(th);
        for (int kind = 0; kind < 3; ++kind)
            for (size_t s = 0; s < sizeof sizes / sizeof sizes[0]; ++s) {
                const uint64_t n = sizes[s];
                rng_state = 1000u * (uint64_t)kind + s;
                double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n),
                       *z = malloc(sizeof(double) * n);
                gen(n, kind, x, y, z);
                char name[64];
                for (int l = 0; l < 4; ++l) {
                    snprintf(name, sizeof name, "t%d kind%d n%llu leaf%u", th, kind,
                             (unsigned long long)n, leaves[l]);
                    one_cloud(name, x, y, z, n, 10, 20, 0.8, leaves[l]);
                    one_cloud(name, x, y, z, n, 6, 0, 0.5, leaves[l]);
                }
                if (n > 1)
                    stats(x, y, n);
                free(x), free(y), free(z);
            }
    }
    reader();
    if (fails) {
        fprintf(stderr, "%d check(s) failed\n", fails);
        return 1;
    }
    puts("oracle sanitize: clean");
    return 0;
}
