/*
 * Deterministic synthetic LiDAR tiles for the OWM ground-seed pipeline.
 *
 * Every value of point i depends only on (seed, stream, i), through a
 * counter-based SplitMix64 hash, so generation is embarrassingly parallel and
 * identical on every machine.  A tile is:
 *   - ground: z0 + sum of 4 sinusoids (wavelength log-uniform in [2, 40] m,
 *     random direction and phase, amplitude 0.75 m each, 3 m in total);
 *   - vegetation: a fraction of points lifted by Exp(mean) metres;
 *   - optional Gaussian clusters (x,y ~ centre + sigma * N(0,1)), which skew
 *     the leaf occupancy of a spatial tree;
 *   - optional flat rectangular roofs whose z is rounded to 1 cm, so that many
 *     points share exactly the same z (exact ties).
 * Coordinates carry a UTM-like offset (x0, y0), like real LiDAR clouds.
 *
 * The .xyz writer emits the 5-number header "N minx maxx miny maxy" that the
 * reference reader accepts (CPP-TBB-CPU/src/optim1_qtree_func.cpp readers,
 * application/owm_optim1_qtree.cpp:127-133) followed by "x y z" lines printed
 * with %.17g so that the text round-trips to the same doubles.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

typedef struct {
    uint64_t seed;
    uint64_t stream;
    uint64_t n;
    double x0, y0;          /* tile origin (UTM-like offset) */
    double width, height;   /* tile extent in metres */
    double z0;              /* base elevation */
    double veg_frac;        /* fraction of vegetation points */
    double veg_mean;        /* mean of the exponential vegetation height */
    double cluster_frac;    /* fraction of points in Gaussian clusters */
    uint32_t n_clusters;
    double cluster_sigma;
    double roof_frac;       /* fraction of points on flat roofs */
    uint32_t n_roofs;
    double roof_side_min, roof_side_max;
    double roof_h_min, roof_h_max;
} owm_synth_spec;

static inline uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* k-th random word of item i in (seed, stream, tag). */
static inline uint64_t ctr(uint64_t seed, uint64_t stream, uint64_t tag,
                           uint64_t i, uint32_t k)
{
    uint64_t s = mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ull + tag));
    return mix64(s ^ mix64(i * 8u + k));
}

static inline double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

enum { TAG_POINT = 1, TAG_WAVE = 2, TAG_CLUSTER = 3, TAG_ROOF = 4 };

typedef struct {
    double kx[4], ky[4], ph[4];
} terrain;

static void terrain_init(const owm_synth_spec *s, terrain *t)
{
    for (int w = 0; w < 4; ++w) {
        double lam = 2.0 * pow(20.0, u01(ctr(s->seed, s->stream, TAG_WAVE, w, 0)));
        double th = 2.0 * M_PI * u01(ctr(s->seed, s->stream, TAG_WAVE, w, 1));
        t->kx[w] = 2.0 * M_PI * cos(th) / lam;
        t->ky[w] = 2.0 * M_PI * sin(th) / lam;
        t->ph[w] = 2.0 * M_PI * u01(ctr(s->seed, s->stream, TAG_WAVE, w, 2));
    }
}

/* Ground height at tile-local (lx, ly). */
static inline double ground(const owm_synth_spec *s, const terrain *t, double lx, double ly)
{
    double g = 0.0;
    for (int w = 0; w < 4; ++w)
        g += 0.75 * sin(t->kx[w] * lx + t->ky[w] * ly + t->ph[w]);
    return s->z0 + g;
}

static inline void roof_rect(const owm_synth_spec *s, uint32_t r, double *cx, double *cy,
                             double *hx, double *hy, double *h)
{
    double sx = s->roof_side_min + (s->roof_side_max - s->roof_side_min) * u01(ctr(s->seed, s->stream, TAG_ROOF, r, 0));
    double sy = s->roof_side_min + (s->roof_side_max - s->roof_side_min) * u01(ctr(s->seed, s->stream, TAG_ROOF, r, 1));
    *hx = 0.5 * sx;
    *hy = 0.5 * sy;
    *cx = *hx + (s->width - sx) * u01(ctr(s->seed, s->stream, TAG_ROOF, r, 2));
    *cy = *hy + (s->height - sy) * u01(ctr(s->seed, s->stream, TAG_ROOF, r, 3));
    *h = s->roof_h_min + (s->roof_h_max - s->roof_h_min) * u01(ctr(s->seed, s->stream, TAG_ROOF, r, 4));
}

static void gen_point(const owm_synth_spec *s, const terrain *t, uint64_t i,
                      double *px, double *py, double *pz)
{
    double cat = u01(ctr(s->seed, s->stream, TAG_POINT, i, 0));
    double lx, ly, z;
    int vegetation_allowed = 1;
    if (cat < s->cluster_frac && s->n_clusters > 0) {
        uint32_t c = (uint32_t)(u01(ctr(s->seed, s->stream, TAG_POINT, i, 1)) * s->n_clusters);
        double sg = s->cluster_sigma;
        double m = 3.0 * sg;
        double ccx = m + (s->width - 2.0 * m) * u01(ctr(s->seed, s->stream, TAG_CLUSTER, c, 0));
        double ccy = m + (s->height - 2.0 * m) * u01(ctr(s->seed, s->stream, TAG_CLUSTER, c, 1));
        /* Box-Muller pairs, re-drawn until the sample lands inside the tile. */
        lx = -1.0;
        ly = -1.0;
        for (uint32_t k = 2; k < 64; k += 2) {
            double u1 = u01(ctr(s->seed, s->stream, TAG_POINT, i, k));
            double u2 = u01(ctr(s->seed, s->stream, TAG_POINT, i, k + 1));
            double rad = sqrt(-2.0 * log(1.0 - u1));
            lx = ccx + sg * rad * cos(2.0 * M_PI * u2);
            ly = ccy + sg * rad * sin(2.0 * M_PI * u2);
            if (lx >= 0.0 && lx < s->width && ly >= 0.0 && ly < s->height)
                break;
        }
        if (!(lx >= 0.0 && lx < s->width && ly >= 0.0 && ly < s->height)) {
            lx = ccx;
            ly = ccy;
        }
        z = ground(s, t, lx, ly);
    } else if (cat < s->cluster_frac + s->roof_frac && s->n_roofs > 0) {
        uint32_t r = (uint32_t)(u01(ctr(s->seed, s->stream, TAG_POINT, i, 1)) * s->n_roofs);
        double cx, cy, hx, hy, h;
        roof_rect(s, r, &cx, &cy, &hx, &hy, &h);
        lx = cx - hx + 2.0 * hx * u01(ctr(s->seed, s->stream, TAG_POINT, i, 2));
        ly = cy - hy + 2.0 * hy * u01(ctr(s->seed, s->stream, TAG_POINT, i, 3));
        /* flat roof: one z per roof, quantised to 1 cm */
        z = round((ground(s, t, cx, cy) + h) * 100.0) / 100.0;
        vegetation_allowed = 0;
    } else {
        lx = s->width * u01(ctr(s->seed, s->stream, TAG_POINT, i, 1));
        ly = s->height * u01(ctr(s->seed, s->stream, TAG_POINT, i, 2));
        z = ground(s, t, lx, ly);
    }
    if (vegetation_allowed && u01(ctr(s->seed, s->stream, TAG_POINT, i, 5)) < s->veg_frac)
        z += -s->veg_mean * log(1.0 - u01(ctr(s->seed, s->stream, TAG_POINT, i, 6)));
    *px = s->x0 + lx;
    *py = s->y0 + ly;
    *pz = z;
}

/* Fill x[i], y[i], z[i] for i in [begin, end). */
void owm_synth_generate_range(const owm_synth_spec *s, uint64_t begin, uint64_t end,
                              double *x, double *y, double *z)
{
    terrain t;
    terrain_init(s, &t);
#pragma omp parallel for schedule(static)
    for (int64_t i = (int64_t)begin; i < (int64_t)end; ++i)
        gen_point(s, &t, (uint64_t)i, &x[i - begin], &y[i - begin], &z[i - begin]);
}

void owm_synth_generate(const owm_synth_spec *s, double *x, double *y, double *z)
{
    owm_synth_generate_range(s, 0, s->n, x, y, z);
}

/* Exact fp64 bounding box (NaNs ignored): out = {minx, maxx, miny, maxy}. */
void owm_synth_bbox(const double *x, const double *y, uint64_t n, double *out)
{
    double mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY;
#pragma omp parallel for reduction(min : mnx, mny) reduction(max : mxx, mxy)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        if (x[i] < mnx) mnx = x[i];
        if (x[i] > mxx) mxx = x[i];
        if (y[i] < mny) mny = y[i];
        if (y[i] > mxy) mxy = y[i];
    }
    out[0] = mnx;
    out[1] = mxx;
    out[2] = mny;
    out[3] = mxy;
}

/* Write a headed .xyz file.  Returns 0 on success, -1 on I/O failure. */
int owm_synth_write_xyz(const char *path, const double *x, const double *y, const double *z,
                        uint64_t n, const double *bbox)
{
    FILE *f = fopen(path, "w");
    if (!f)
        return -1;
    fprintf(f, "%llu\n%.17g\n%.17g\n%.17g\n%.17g\n", (unsigned long long)n, bbox[0], bbox[1],
            bbox[2], bbox[3]);
    for (uint64_t i = 0; i < n; ++i)
        fprintf(f, "%.17g %.17g %.17g\n", x[i], y[i], z[i]);
    return fclose(f) == 0 ? 0 : -1;
}

/* Public counter RNG so host code can draw reproducible tile parameters. */
uint64_t owm_synth_ctr(uint64_t seed, uint64_t stream, uint64_t tag, uint64_t i, uint32_t k)
{
    return ctr(seed, stream, tag, i, k);
}

/* Format points as .xyz text lines "x y z\n" with `decimals` fixed decimals
 * (the layout of LiDAR exports); parallel over chunks.  Returns the bytes
 * written, 0 if cap is too small (80 bytes per point always suffice). */
uint64_t owm_synth_format_xyz(const double *x, const double *y, const double *z, uint64_t n,
                              int decimals, char *out, uint64_t cap)
{
    const int nt = omp_get_max_threads();
    const uint64_t per = (n + (uint64_t)nt - 1) / (uint64_t)nt;
    uint64_t *len = (uint64_t *)calloc((size_t)nt + 1, sizeof(uint64_t));
    char **buf = (char **)calloc((size_t)nt, sizeof(char *));
    if (!len || !buf)
        return 0;
#pragma omp parallel num_threads(nt)
    {
        const int t = omp_get_thread_num();
        const uint64_t a = (uint64_t)t * per, b = a + per < n ? a + per : n;
        if (a < b) {
            char *w = (char *)malloc((size_t)((b - a) * 80 + 1));
            uint64_t o = 0;
            if (w)
                for (uint64_t i = a; i < b; ++i)
                    o += (uint64_t)snprintf(w + o, 80, "%.*f %.*f %.*f\n", decimals, x[i], decimals,
                                            y[i], decimals, z[i]);
            buf[t] = w;
            len[t + 1] = w ? o : 0;
        }
    }
    uint64_t total = 0;
    for (int t = 0; t < nt; ++t)
        total += len[t + 1];
    if (total <= cap) {
        uint64_t o = 0;
        for (int t = 0; t < nt; ++t) {
            if (buf[t])
                memcpy(out + o, buf[t], (size_t)len[t + 1]);
            o += len[t + 1];
        }
    } else {
        total = 0;
    }
    for (int t = 0; t < nt; ++t)
        free(buf[t]);
    free(buf);
    free(len);
    return total;
}
