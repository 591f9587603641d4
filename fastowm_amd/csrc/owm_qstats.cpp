// Quadtree leaf statistics (owm_qtree_stats) and the statistics tool's text
// reports (owm_qtree_stats_write): the reference's owm_stats
// (CPP-TBB-CPU/application/owm_stats.cpp:60-136) with the tree resolved on the
// device one depth at a time (qstats.hip).
#include "owm_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <vector>

using namespace owm;

namespace {

uint64_t al256(uint64_t v) { return (v + 255) / 256 * 256; }

double round2d(double z) { return std::round(z * 100.0) / 100.0; }

// Radius of the depth-d nodes: the root's float radius, halved by
// createQuadrants as float(radius * 0.5) (CPP-TBB-CPU/src/optim3_func.cpp:63).
float depth_radius(float r0, int d)
{
    float r = r0;
    for (int k = 0; k < d; ++k)
        r = (float)((double)r * 0.5);
    return r;
}

// Density of a leaf as getDensity computes it: points.size() / (4 * radius *
// radius) with a float radius, so float arithmetic (stats_func.hpp:6-16).
double leaf_density(uint32_t points, float r) { return (double)((float)points / (4 * r * r)); }

}  // namespace

extern "C" {

owm_status owm_qtree_stats(owm_ctx *ctx, const double *x, const double *y, uint64_t n,
                           int device_ptrs, const owm_qtree_params *p, owm_leaf_bin *bins,
                           uint64_t cap, owm_qtree_summary *out)
{
    if (!ctx || !p || !out || (cap && !bins) || (n && (!x || !y)))
        return OWM_E_ARG;
    memset(out, 0, sizeof *out);
    if (p->policy > OWM_QT_MINRADIUS || (p->policy == OWM_QT_MAXNUMBER && p->max_number == 0)) {
        set_err(ctx, "policy %u with max_number %u: maxNumber must be at least 1", p->policy,
                p->max_number);
        return OWM_E_ARG;
    }
    if (p->level > OWM_QT_MAX_LEVEL) {
        set_err(ctx, "level %u above %u", p->level, OWM_QT_MAX_LEVEL);
        return OWM_E_ARG;
    }
    if (p->policy == OWM_QT_MINRADIUS && !(p->min_radius == p->min_radius)) {
        set_err(ctx, "min_radius is not a number");
        return OWM_E_ARG;
    }
    try {
        (void)hipSetDevice(ctx->device);
        hipStream_t s = ctx->stream;
        // open nodes per depth: the forced 4^level, then nodes holding more
        // than maxNumber points (maxNumber) or at least one point (minRadius)
        const uint64_t forced = 1ull << (2 * p->level);
        const uint64_t by_points = p->policy == OWM_QT_MAXNUMBER ? n / ((uint64_t)p->max_number + 1) : n;
        const uint64_t open_cap = std::max(forced, by_points) + 1;
        if (4 * open_cap >= (1ull << 32) || n >= (1ull << 32) - 1) {
            set_err(ctx, "%llu points: more child slots than 32-bit state indexes",
                    (unsigned long long)n);
            return OWM_E_ARG;
        }
        const uint64_t slots = 4 * open_cap;
        const uint64_t ovf_cap = n / QS_HB + 1;
        const uint64_t hist_words = (uint64_t)(QS_MAX_DEPTH + 1) * QS_HB;
        const bool need_bbox = !(p->flags & OWM_F_HAVE_BBOX);
        // layout: [x y] | st | cnt | remap | cen0 | cen1 | nopen | hist | ovf | novf flags | bbox
        const uint64_t w_xy = device_ptrs ? 0 : al256(n * 8);
        const uint64_t w_st = al256(std::max<uint64_t>(n, 1) * 4), w_sl = al256(slots * 4);
        const uint64_t w_cen = al256(open_cap * 16), w_rng = al256(open_cap * 8);
        const uint64_t w_nopen = al256((QS_MAX_DEPTH + 1) * 4);
        const uint64_t w_hist = al256(hist_words * 8), w_ovf = al256(ovf_cap * 8);
        const uint64_t w_bb = need_bbox ? al256(bbox_key_words(1) * 8) + 256 : 0;
        // sort scratch for up to 16 sorted depths (4 passes of one key byte)
        const uint64_t w_sort = al256(sort_scratch_bytes(std::max<uint64_t>(n, 1), 4, 1, 0));
        const uint64_t total = 2 * w_xy + w_st + 2 * w_sl + 2 * w_cen + 2 * w_rng + w_nopen +
                               w_hist + w_ovf + 256 + w_sort + w_bb;
        HIP_TRY(hipStreamSynchronize(s), "sync before the statistics");
        if (ctx->qs_work.bytes < total) {
            if (ctx->budget && ctx->held_bytes() - ctx->qs_work.bytes + total > ctx->budget) {
                set_err(ctx, "quadtree statistics need %llu bytes over the budget",
                        (unsigned long long)total);
                return OWM_E_BUDGET;
            }
            HIP_TRY(ctx->qs_work.reserve(total, ctx->room_for(ctx->qs_work)), "statistics buffer");
        }
        uint8_t *w = (uint8_t *)ctx->qs_work.ptr;
        auto take = [&w](uint64_t b) { uint8_t *r = w; w += b; return r; };
        const double *dx = x, *dy = y;
        if (!device_ptrs) {
            double *ux = (double *)take(w_xy), *uy = (double *)take(w_xy);
            if (n) {
                HIP_TRY(hipMemcpyAsync(ux, x, n * 8, hipMemcpyHostToDevice, s), "points H2D");
                HIP_TRY(hipMemcpyAsync(uy, y, n * 8, hipMemcpyHostToDevice, s), "points H2D");
            }
            dx = ux;
            dy = uy;
        }
        QsDev D{};
        D.x = dx;
        D.y = dy;
        D.n = n;
        D.st = (uint32_t *)take(w_st);
        D.cnt = (uint32_t *)take(w_sl);
        D.remap = (uint32_t *)take(w_sl);
        D.cen[0] = (double2 *)take(w_cen);
        D.cen[1] = (double2 *)take(w_cen);
        D.rng[0] = (uint2 *)take(w_rng);
        D.rng[1] = (uint2 *)take(w_rng);
        uint8_t *zero_from = w;  // nopen, hist, novf and flags start at zero
        D.nopen = (uint32_t *)take(w_nopen);
        D.hist = (unsigned long long *)take(w_hist);
        D.ovf = (uint2 *)take(w_ovf);
        D.novf = (uint32_t *)take(256);
        D.flags = D.novf + 1;
        uint8_t *const zero_end = w;
        uint8_t *const sort_base = take(w_sort);
        SortBufs S = carve_sort(sort_base, std::max<uint64_t>(n, 1), 4);
        D.ovf_cap = (uint32_t)ovf_cap;
        D.open_cap = (uint32_t)open_cap;
        D.policy = p->policy;
        D.max_number = p->max_number;
        D.level = p->level;
        D.min_radius = p->min_radius;

        double box[4];
        if (need_bbox) {
            if (n == 0) {
                set_err(ctx, "no points to take a bbox from");
                return OWM_E_ARG;
            }
            unsigned long long *bk = (unsigned long long *)take(al256(bbox_key_words(1) * 8));
            uint32_t *ptoff = (uint32_t *)take(256);
            uint8_t *h = ctx->pinned(ctx->h_io, ctx->h_io_cap, 256);
            if (!h) {
                set_err(ctx, "pinned host allocation failed");
                return OWM_E_NOMEM;
            }
            const uint32_t po[2] = {0, (uint32_t)n};
            memcpy(h, po, 8);
            HIP_TRY(hipMemcpyAsync(ptoff, h, 8, hipMemcpyHostToDevice, s), "bbox offsets");
            BatchDev B0{};
            B0.ntiles = 1;
            B0.np = (uint32_t)n;
            B0.pt_off = ptoff;
            launch_bbox_init(s, bk, 1);
            launch_bbox(s, B0, dx, dy, bk);
            HIP_TRY(hipMemcpyAsync(h + 64, bk, 32, hipMemcpyDeviceToHost, s), "bbox D2H");
            HIP_TRY(hipStreamSynchronize(s), "bbox");
            const unsigned long long *hk = (const unsigned long long *)(h + 64);
            if (hk[0] == ~0ull || hk[1] == ~0ull) {
                set_err(ctx, "no finite points to take a bbox from");
                return OWM_E_ARG;
            }
            auto dec = [](unsigned long long k) {
                uint64_t b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
                double v;
                memcpy(&v, &b, 8);
                return v;
            };
            box[0] = dec(hk[0]);
            box[1] = dec(hk[2]);
            box[2] = dec(hk[1]);
            box[3] = dec(hk[3]);
        } else {
            memcpy(box, p->bbox, sizeof box);
        }
        // getRadius / getCenter (CPP-TBB-CPU/src/optim3_func.cpp:6-26)
        const double rx = (box[1] - box[0]) / 2.0, ry = (box[3] - box[2]) / 2.0;
        const float r0 = (float)(rx >= ry ? rx : ry);
        const double cx = box[0] + rx, cy = box[2] + ry;
        // depths sorted in the keys: 12, or all 16 when leaves are expected
        // below depth 12 (minRadius: every occupied node splits down to the
        // radius limit, known here; maxNumber: about log4(n / maxNumber) + 1,
        // as for uniform points).  Depths past the sorted ones go point by point.
        int sl = 12;
        if (p->policy == OWM_QT_MINRADIUS) {
            int dleaf = 0;  // occupied nodes split at depths 0 .. dleaf-1
            for (float rr = r0; dleaf < QS_MAX_DEPTH && (double)rr * 0.5 > (double)p->min_radius;
                 rr = (float)((double)rr * 0.5))
                ++dleaf;
            if (dleaf >= QS_MAX_DEPTH && n > 0) {
                // every occupied node splits while radius * 0.5 > minRadius:
                // with minRadius <= 0 (or tiny against the box) that goes on
                // past the deepest level this tool follows
                set_err(ctx, "minRadius %g: occupied nodes of a %g m root would split below "
                             "depth %d",
                        (double)p->min_radius, (double)r0, QS_MAX_DEPTH);
                return OWM_E_ARG;
            }
            sl = dleaf > 12 ? 16 : 12;
        } else if (n > 0) {
            const double est = std::log((double)n / ((double)p->max_number + 1.0)) / std::log(4.0) + 1.0;
            sl = est > 12.0 ? 16 : 12;
        }
        D.sort_levels = sl;
        const int passes = sl / 4;

        int pi = ctx->prof.begin(s, PH_QSTATS);
        launch_fill_u32(s, D.cnt, slots, 0u);
        launch_fill_u32(s, (uint32_t *)zero_from, (uint64_t)(zero_end - zero_from) / 4, 0u);
        launch_qs_root(s, D, cx, cy);
        // path keys, sorted with the point index as value
        QsPath P{};
        P.cx = cx;
        P.cy = cy;
        for (int k = 0; k < QS_KEY_LEVELS; ++k)
            P.off[k] = (double)(depth_radius(r0, k) * 0.5f);
        launch_qs_path(s, dx, dy, n, P, sl, S.k0);
        launch_fill_u32(s, S.hist, (uint64_t)(sort_base + w_sort - (uint8_t *)S.hist) / 4, 0u);
        ctx->prof.end(s, pi, 0, 0, 5);
        // the keys alone: which point sits where is needed only below the key
        // depths, and then the sort is repeated with the index as value (LSD
        // passes are stable, so the key order, hence every node's range, is
        // the same)
        D.key = S.k0;
        D.perm = nullptr;
        if (n) {
            const int which = run_sort(ctx, PH_QSTATS, s, S, (uint32_t)n, nullptr, passes, true, false);
            D.key = which ? S.k1 : S.k0;
        }
        uint32_t *hn = (uint32_t *)ctx->pinned(ctx->h_io, ctx->h_io_cap, 4 * (QS_MAX_DEPTH + 4));
        if (!hn) {
            set_err(ctx, "pinned host allocation failed");
            return OWM_E_NOMEM;
        }
        // depths in groups of up to 6 (the key-range depths and the point
        // depths apart), then one readback of the open-node counts
        int d = 1, last = 0;
        float r = r0;
        while (true) {
            if (d == QS_KEY_LEVELS + 1) {
                // nodes are open at the last key depth: sort again with values
                pi = ctx->prof.begin(s, PH_QSTATS);
                launch_qs_path(s, dx, dy, n, P, sl, S.k0);
                launch_fill_u32(s, S.hist, (uint64_t)(sort_base + w_sort - (uint8_t *)S.hist) / 4, 0u);
                ctx->prof.end(s, pi, 0, 0, 2);
                const int which = run_sort(ctx, PH_QSTATS, s, S, (uint32_t)n, nullptr, passes, false, true);
                D.key = nullptr;  // not read below the key depths
                D.perm = which ? S.v1 : S.v0;
                ctx->prof.add_work(PH_QSTATS, 20 * n, 0);
            }
            int upto = std::min(QS_MAX_DEPTH, d + 5);
            if (d <= sl && upto > sl)
                upto = sl;
            if (d <= QS_KEY_LEVELS && upto > QS_KEY_LEVELS)
                upto = QS_KEY_LEVELS;
            uint32_t launches = 0;
            pi = ctx->prof.begin(s, PH_QSTATS);
            for (; d <= upto; ++d) {
                const float rd = (float)((double)r * 0.5);
                const double off = (double)(r * 0.5f);  // createQuadrants' radius * 0.5f
                if (d <= sl) {
                    launch_qs_split(s, D, d, rd, off);
                    launches += 1;
                } else {
                    if (d == sl + 1) {
                        launch_qs_deep_start(s, D);
                        launches += 2;
                    }
                    launch_qs_level(s, D, d, rd, off);
                    launches += 2;
                }
                r = rd;
            }
            ctx->prof.end(s, pi, 0, 0, launches);
            HIP_TRY(hipMemcpyAsync(hn, D.nopen, 4 * (QS_MAX_DEPTH + 1), hipMemcpyDeviceToHost, s),
                    "open nodes D2H");
            HIP_TRY(hipMemcpyAsync(hn + QS_MAX_DEPTH + 1, D.novf, 8, hipMemcpyDeviceToHost, s),
                    "statistics flags D2H");
            HIP_TRY(hipStreamSynchronize(s), "statistics levels");
            if (hn[upto] == 0) {
                for (last = upto; last > 0 && hn[last - 1] == 0; --last)
                    ;
                break;  // the depth-`last` nodes are all leaves
            }
            if (upto == QS_MAX_DEPTH) {
                if (p->policy == OWM_QT_MAXNUMBER)
                    set_err(ctx, "a node holds more than %u points below depth %d: the "
                                 "reference quadtree never stops splitting there",
                            p->max_number, QS_MAX_DEPTH);
                else
                    set_err(ctx, "minRadius %g: occupied nodes split below depth %d",
                            (double)p->min_radius, QS_MAX_DEPTH);
                return OWM_E_ARG;
            }
        }
        const uint32_t novf = hn[QS_MAX_DEPTH + 1], flags = hn[QS_MAX_DEPTH + 2];
        std::vector<uint32_t> nopen(hn, hn + QS_MAX_DEPTH + 1);
        if (flags) {
            set_err(ctx, "statistics capacity guard tripped (flags %u)", flags);
            return OWM_E_INTERNAL;
        }
        std::vector<unsigned long long> hist((uint64_t)(last + 1) * QS_HB);
        std::vector<uint2> ovf(novf);
        HIP_TRY(hipMemcpyAsync(hist.data(), D.hist, hist.size() * 8, hipMemcpyDeviceToHost, s),
                "histogram D2H");
        if (novf)
            HIP_TRY(hipMemcpyAsync(ovf.data(), D.ovf, (uint64_t)novf * 8, hipMemcpyDeviceToHost, s),
                    "large leaves D2H");
        HIP_TRY(hipStreamSynchronize(s), "statistics readback");
        ctx->prof.collect();

        std::map<std::pair<uint32_t, uint32_t>, uint64_t> big;
        for (const uint2 &e : ovf)
            ++big[{e.x, e.y}];
        std::vector<owm_leaf_bin> cls;
        for (int dd = 0; dd <= last; ++dd) {
            for (uint32_t c = 0; c < QS_HB; ++c)
                if (hist[(uint64_t)dd * QS_HB + c])
                    cls.push_back({(uint32_t)dd, c, hist[(uint64_t)dd * QS_HB + c]});
            for (auto it = big.lower_bound({(uint32_t)dd, 0}); it != big.end() && it->first.first == (uint32_t)dd; ++it)
                cls.push_back({it->first.first, it->first.second, it->second});
        }
        uint64_t leaves = 0, settled = 0, routed = 0;
        for (const auto &c : cls) {
            leaves += c.leaves;
            settled += (uint64_t)c.points * c.leaves;
            if (c.level > (uint32_t)sl)  // routed once per depth below the sorted depths
                routed += (uint64_t)c.points * c.leaves * (c.level - sl);
        }
        if (settled != n) {
            set_err(ctx, "leaves hold %llu points of %llu", (unsigned long long)settled,
                    (unsigned long long)n);
            return OWM_E_INTERNAL;
        }
        // algorithmic bytes besides the sort's own: the path keys (x, y read,
        // key written); per open node of a sorted depth its range and centre
        // read (24) and the keys at its three child boundaries (12), per new
        // open node its range and centre written (24); below the sorted
        // depths, the state fill and per depth every point's state (4 + 4),
        // per routed point its key or coordinates, node and new state (at
        // least 12), per child slot its count (read and cleared) and remap
        // (12), per new open node a centre
        uint64_t alg = 20 * n;
        for (int dd = 1; dd <= last; ++dd) {
            if (dd <= sl)
                alg += 36ull * nopen[dd - 1] + 24ull * nopen[dd];
            else
                alg += 4 * n + 12 * 4ull * nopen[dd - 1] + 16ull * nopen[dd];
        }
        if (last > sl)
            alg += 4 * n + 12 * routed;
        ctx->prof.add_work(PH_QSTATS, alg, n);
        memcpy(out->bbox, box, sizeof box);
        out->max_radius = r0;
        out->depth = (uint32_t)last;
        out->leaves = leaves;
        out->n_bins = cls.size();
        const uint64_t k = std::min<uint64_t>(cap, cls.size());
        if (k)
            memcpy(bins, cls.data(), k * sizeof(owm_leaf_bin));
        return cls.size() > cap ? OWM_E_CAPACITY : OWM_OK;
    } catch (const std::bad_alloc &) {
        set_err(ctx, "host allocation failed");
        return OWM_E_NOMEM;
    } catch (...) {
        set_err(ctx, "internal error");
        return OWM_E_INTERNAL;
    }
}

owm_status owm_qtree_stats_write(const char *path, const owm_qtree_params *p,
                                 const owm_qtree_summary *sm, const owm_leaf_bin *bins,
                                 uint64_t n_bins, uint64_t n_points)
{
    if (!path || !p || !sm || (n_bins && !bins))
        return OWM_E_ARG;
    try {
        const std::string file_name(path);
        // the per-leaf values the tool collects, as histograms (value -> leaves)
        std::map<int, int> by_points, by_level;
        std::map<double, int> by_radius, by_density;
        for (uint64_t k = 0; k < n_bins; ++k) {
            const owm_leaf_bin &c = bins[k];
            const float r = depth_radius(sm->max_radius, (int)c.level);
            by_points[(int)c.points] += (int)c.leaves;
            by_level[(int)c.level] += (int)c.leaves;
            by_radius[(double)r] += (int)c.leaves;
            by_density[leaf_density(c.points, r)] += (int)c.leaves;
        }
        const double width = round2d(sm->bbox[1] - sm->bbox[0]);
        const double height = round2d(sm->bbox[3] - sm->bbox[2]);
        const double density = (double)(uint32_t)n_points / (width * height);
        // save_histogram (stats_func.hpp:120-211)
        {
            std::fstream out(file_name, std::ios::out | std::ios::app);
            if (!out.is_open())
                return OWM_E_ARG;
            size_t nleaf = 0;
            for (auto &it : by_points)
                nleaf += it.second;
            out << "HISTOGRAM: " << file_name << std::endl;
            out << "Number of leaf nodes: " << nleaf << std::endl;
            if (p->policy == OWM_QT_MINRADIUS)
                out << "MinRadius: " << p->min_radius << " m" << std::endl << std::endl;
            else
                out << "MaxNumber: " << (int)p->max_number << " pts" << std::endl << std::endl;
            out << "Observed: " << std::endl;
            uint32_t acc = 0;
            int min_x = 999, max_x = 0;
            for (auto &it : by_points) {
                if (it.first < min_x)
                    min_x = it.first;
                if (max_x < it.first)
                    max_x = it.first;
                acc += it.first * it.second;
            }
            out << "Average number of points/leaf-node: " << (nleaf ? acc / nleaf : 0) << std::endl;
            out << "__min: " << min_x << std::endl;
            out << "__max: " << max_x << std::endl << std::endl;
            out << "Estimation (assuming equal distribution of points): " << std::endl;
            out << "Average density: " << density << std::endl << std::endl;
            out << "x_i n_i x_i*n_i" << std::endl << std::endl;
            for (auto &it : by_points)
                out << it.first << " " << it.second << " " << it.first * it.second << std::endl;
            out << std::endl << std::endl;
        }
        // save_areas (:266-298)
        {
            std::fstream out(file_name, std::ios::out | std::ios::app);
            size_t nleaf = 0;
            for (auto &it : by_radius)
                nleaf += it.second;
            out << "Radius HISTOGRAM:" << std::endl;
            out << "Number of leaf-nodes: " << nleaf << std::endl << std::endl;
            out << "r_i n_i a_i" << std::endl << std::endl;
            for (auto &it : by_radius)
                out << it.first << " " << it.second << " " << 4 * it.first * it.first << std::endl;
            out << std::endl << std::endl;
        }
        // save_density (:300-331)
        {
            std::fstream out(file_name, std::ios::out | std::ios::app);
            size_t nleaf = 0;
            for (auto &it : by_density)
                nleaf += it.second;
            out << "Density HISTOGRAM:" << std::endl;
            out << "Number of leaf-nodes: " << nleaf << std::endl << std::endl;
            out << "d_i n_i" << std::endl << std::endl;
            for (auto &it : by_density)
                out << it.first << " " << it.second << std::endl;
            out << std::endl << std::endl;
        }
        // save_levels (:333-366)
        {
            std::fstream out(file_name, std::ios::out | std::ios::app);
            size_t nleaf = 0;
            for (auto &it : by_level)
                nleaf += it.second;
            out << "Levels HISTOGRAM:" << std::endl;
            out << "Number of leaf-nodes: " << nleaf << std::endl << std::endl;
            out << "level_i n_i" << std::endl << std::endl;
            for (auto &it : by_level)
                out << it.first << " " << it.second << std::endl;
            out << std::endl << std::endl;
            if (!out.good())
                return OWM_E_ARG;
        }
        return OWM_OK;
    } catch (...) {
        return OWM_E_NOMEM;
    }
}

}  // extern "C"
