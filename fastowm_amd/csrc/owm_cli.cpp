// owm: command-line driver of the MI355X OWM pipeline.
//
// Mirrors the reference CPU driver's flags (CPP-TBB-CPU/application/owm_optim3.cpp:14-40):
//   -i <input without .xyz>  -W <Wsize>  -B <Bsize>  -O <Overlap>  -s <maxNumber>
//   -l <loops>  --results <csv>  -n <threads>  -r <minRadius>  -L <level>
// (-n, -r and -L are recorded in the results row as the reference records
// them; the GPU path has no thread count, and -r / -L shape only --stats)
// plus --device <id>, -o <output .xyz of the seeds>, --gold <file> (default
// <input>_salidaGOLD.xyz, as the reference: the results CSV's Correctness) and
// --data-bbox (take the bbox of a header-less file from its points instead of
// refusing it).
//
// --stats maxnum|radius runs the reference's statistics tool instead
// (CPP-TBB-CPU/application/owm_stats.cpp: `statsmaxnum` with -s maxNumber,
// `stats` with -r minRadius, -L level for both): the quadtree's leaf
// histograms are appended to --results (default tbb_histogram.csv).
//
// Input: .xyz text, "x y z" per line.  The bbox comes from the reference's
// 5-number header "N minx maxx miny maxy" (its *H.xyz convention,
// src/optim3_func.cpp:681-700), else from its table of header-less clouds
// (:667-680, owm_xyz_known_cloud), else the file is refused ("No header
// data!") unless --data-bbox.  The file is parsed on the device
// (owm_xyz_parse, its time printed apart) and stays in HBM.  Timing follows
// the reference (owm_optim3.cpp:77-89,145-206): the tree is built and timed
// once (TreeTime), then -l loops of stages 1-3 run on it (OwmTime, averaged).
// -o writes the seeds as its _salida.xyz: LLPs, then stage-3 additions.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "owm.h"

namespace {

bool read_file(const std::string &path, std::vector<char> &buf)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f)
        return false;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    buf.resize(sz > 0 ? (size_t)sz : 0);
    size_t got = sz > 0 ? fread(buf.data(), 1, (size_t)sz, f) : 0;
    fclose(f);
    buf.resize(got);
    return true;
}

void usage()
{
    fprintf(stderr,
            "usage: owm -i <input without .xyz> [-W 10] [-B 20] [-O 0.80] [-s 32] [-l 1]\n"
            "           [-n 1] [-r 0.1] [-L 5] [--results o3_memoization.csv] [-o seeds.xyz]\n"
            "           [--gold <file>] [--data-bbox] [--device 0]\n"
            "       owm --stats maxnum|radius -i <input without .xyz> [-s 32] [-r 0.1] [-L 5]\n"
            "           [--results tbb_histogram.csv] [--device 0]\n");
}

// The statistics tool on a parsed cloud (owm_stats.cpp:60-136).
int run_stats(owm_ctx *ctx, const owm_cloud &cloud, const std::string &mode, uint32_t max_number,
              float min_radius, uint32_t level, const std::string &results)
{
    owm_qtree_params q{};
    q.policy = mode == "radius" ? OWM_QT_MINRADIUS : OWM_QT_MAXNUMBER;
    q.max_number = max_number;
    q.min_radius = min_radius;
    q.level = level;
    if (cloud.flags & OWM_F_HAVE_BBOX) {
        q.flags = OWM_F_HAVE_BBOX;
        memcpy(q.bbox, cloud.bbox, sizeof q.bbox);
    }
    std::vector<owm_leaf_bin> bins(4096);
    owm_qtree_summary sm;
    auto t0 = std::chrono::steady_clock::now();
    owm_status st = owm_qtree_stats(ctx, cloud.x, cloud.y, cloud.n, 1, &q, bins.data(), bins.size(), &sm);
    if (st == OWM_E_CAPACITY) {
        bins.resize(sm.n_bins);
        st = owm_qtree_stats(ctx, cloud.x, cloud.y, cloud.n, 1, &q, bins.data(), bins.size(), &sm);
    }
    auto t1 = std::chrono::steady_clock::now();
    if (st != OWM_OK) {
        fprintf(stderr, "owm_qtree_stats: %s: %s\n", owm_status_str(st), owm_last_error(ctx));
        return 1;
    }
    const double rx = (sm.bbox[1] - sm.bbox[0]) / 2.0, ry = (sm.bbox[3] - sm.bbox[2]) / 2.0;
    printf("QTREE PARAMETERS:\n");
    printf("MaxRadius:  %.3f\n", sm.max_radius);
    printf("Center:     %.2f , %.2f\n", sm.bbox[0] + rx, sm.bbox[2] + ry);
    printf("Radius:     %.2f , %.2f\n", rx, ry);
    if (q.policy == OWM_QT_MINRADIUS)
        printf("Building quadtree with minRadius: %g\n", min_radius);
    else
        printf("Building quadtree with maxNumber: %u\n", max_number);
    printf("Leaf nodes: %llu (depth %u), statistics in %.6f s\n", (unsigned long long)sm.leaves,
           sm.depth, std::chrono::duration<double>(t1 - t0).count());
    st = owm_qtree_stats_write(results.c_str(), &q, &sm, bins.data(), sm.n_bins, cloud.n);
    if (st != OWM_OK) {
        fprintf(stderr, "cannot write %s\n", results.c_str());
        return 1;
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    std::string input = "data/INAER_2011_Alcoy", results, out, gold, stats;
    float min_radius = 0.1f;
    uint32_t level = 5;
    owm_params P{};
    P.wsize = 10;
    P.bsize = 20;
    P.overlap = 0.80;
    P.leaf_size = 32;
    int loops = 1, device = 0, num_threads = 1;
    bool data_bbox = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&]() -> const char * {
            if (i + 1 >= argc) {
                usage();
                exit(2);
            }
            return argv[++i];
        };
        if (a == "-i" || a == "--input") input = val();
        else if (a == "-W" || a == "--Wsize") P.wsize = (uint32_t)strtoul(val(), nullptr, 10);
        else if (a == "-B" || a == "--Bsize") P.bsize = (uint32_t)strtoul(val(), nullptr, 10);
        else if (a == "-O" || a == "--Overlap") P.overlap = strtod(val(), nullptr);
        else if (a == "-s" || a == "--size") P.leaf_size = (uint32_t)strtoul(val(), nullptr, 10);
        else if (a == "-l" || a == "--loop") loops = atoi(val());
        else if (a == "--results") results = val();
        else if (a == "-o") out = val();
        else if (a == "--gold") gold = val();
        else if (a == "--device") device = atoi(val());
        else if (a == "--stats") stats = val();
        else if (a == "-r" || a == "--radius") min_radius = strtof(val(), nullptr);
        else if (a == "-L" || a == "--level") level = (uint32_t)strtoul(val(), nullptr, 10);
        else if (a == "-n" || a == "--num_threads") num_threads = atoi(val());
        else if (a == "--data-bbox") data_bbox = true;
        else if (a == "-h" || a == "--help") {
            usage();
            return 0;
        } else {
            usage();
            return 2;
        }
    }
    if (!stats.empty() && stats != "maxnum" && stats != "radius") {
        usage();
        return 2;
    }
    if (results.empty())
        results = stats.empty() ? "o3_memoization.csv" : "tbb_histogram.csv";
    std::string path = input + ".xyz";
    std::vector<char> text;
    if (!read_file(path, text)) {
        fprintf(stderr, "Unable to open %s\n", path.c_str());
        return 1;
    }
    owm_ctx *ctx = nullptr;
    owm_status st = owm_ctx_create(device, 0, &ctx);
    if (st != OWM_OK) {
        fprintf(stderr, "owm_ctx_create: %s\n", owm_status_str(st));
        return 1;
    }
    // readXYZfile on the device: the cloud stays in HBM for every loop
    owm_cloud cloud;
    auto r0 = std::chrono::steady_clock::now();
    st = owm_xyz_parse(ctx, text.data(), text.size(), &cloud);
    auto r1 = std::chrono::steady_clock::now();
    if (st != OWM_OK) {
        fprintf(stderr, "Unable to read %s: %s\n", path.c_str(), owm_last_error(ctx));
        owm_ctx_destroy(ctx);
        return 1;
    }
    // the bbox: the file's header, else readXYZfile's table of the reference's
    // header-less clouds (which also fixes how many points are read); other
    // header-less files are refused as the reference does, unless --data-bbox
    uint64_t known_n = 0;
    double known_box[4];
    if (cloud.flags & OWM_F_HAVE_BBOX) {
        P.flags |= OWM_F_HAVE_BBOX;
        memcpy(P.bbox, cloud.bbox, sizeof P.bbox);
    } else if (owm_xyz_known_cloud(path.c_str(), &known_n, known_box) == OWM_OK) {
        if (cloud.n < known_n) {
            fprintf(stderr, "Error reading values: %s holds %llu points, %llu expected\n",
                    path.c_str(), (unsigned long long)cloud.n, (unsigned long long)known_n);
            owm_ctx_destroy(ctx);
            return 1;
        }
        cloud.n = known_n;
        P.flags |= OWM_F_HAVE_BBOX;
        memcpy(P.bbox, known_box, sizeof P.bbox);
        cloud.flags |= OWM_F_HAVE_BBOX;
        memcpy(cloud.bbox, known_box, sizeof cloud.bbox);
    } else if (!data_bbox) {
        printf("No header data!\n");
        owm_ctx_destroy(ctx);
        return 1;
    }
    const uint64_t n = cloud.n;
    printf("Input: %s (%llu points) on device %d, read in %.6f s\n", path.c_str(),
           (unsigned long long)n, device, std::chrono::duration<double>(r1 - r0).count());
    if (P.flags & OWM_F_HAVE_BBOX)
        printf("Npoints=%llu; xmin = %.2lf; xmax = %.2lf; ymin = %.2lf; ymax = %.2lf\n",
               (unsigned long long)n, P.bbox[0], P.bbox[1], P.bbox[2], P.bbox[3]);
    if (!stats.empty()) {
        const int rc = run_stats(ctx, cloud, stats, P.leaf_size, min_radius, level, results);
        owm_ctx_destroy(ctx);
        return rc;
    }
    // the tree once (the reference's timed parallel_qtree), then the loops of
    // stages 1-3 on it (its timed OWM loop), as owm_optim3.cpp:77-89,145-206
    auto fail = [&](const char *what, owm_status st) {
        fprintf(stderr, "%s: %s: %s\n", what, owm_status_str(st), owm_last_error(ctx));
        owm_ctx_destroy(ctx);
        return 1;
    };
    owm_tree *tree = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    st = owm_tree_build(ctx, cloud.x, cloud.y, cloud.z, n, 1, P.leaf_size, &tree);
    auto t1 = std::chrono::steady_clock::now();
    if (st != OWM_OK)
        return fail("owm_tree_build", st);
    const double time_tree = std::chrono::duration<double>(t1 - t0).count();
    std::vector<uint32_t> seeds(n + 1);
    uint64_t ns = 0;
    std::vector<double> times;
    for (int r = 0; r < loops; ++r) {
        t0 = std::chrono::steady_clock::now();
        st = owm_tree_run(ctx, tree, &P, seeds.data(), seeds.size(), &ns);
        t1 = std::chrono::steady_clock::now();
        if (st != OWM_OK)
            return fail("owm_tree_run", st);
        const double s = std::chrono::duration<double>(t1 - t0).count();
        times.push_back(s);
        printf("TOTAL time elapsed:     %.6f s\n", s);
        printf("Output ground seed-point cloud with %llu points, %llu fewer points than input cloud\n",
               (unsigned long long)ns, (unsigned long long)(n - ns));
    }
    double avg = 0;
    for (double t : times)
        avg += t;
    avg /= times.size();
    printf("Time elapsed at Quadtree construction:     %.6f s\n\n", time_tree);
    printf("Average OWM: %.6lf\n", avg);
    printf("Total Tree Construction + OWM: %.6lf\n\n", avg + time_tree);
    printf("Mpoints/s (cloud resident in HBM): %.1f\n", n / (avg + time_tree) / 1e6);
    // host copies of the cloud, for the seed file and the GOLD check
    const std::string gold_path = gold.empty() ? input + "_salidaGOLD.xyz" : gold;
    std::vector<char> gold_text;
    const bool have_gold = read_file(gold_path, gold_text);
    std::vector<double> x, y, z;
    if (!out.empty() || have_gold) {
        x.resize(n);
        y.resize(n);
        z.resize(n);
        st = owm_cloud_read(ctx, &cloud, x.data(), y.data(), z.data());
        if (st != OWM_OK)
            return fail("owm_cloud_read", st);
    }
    // the reference's check_results against <input>_salidaGOLD.xyz (or --gold)
    double correctness = -1.0;
    if (have_gold) {
        std::vector<double> gx, gy, gz, sx, sy, sz;
        gold_text.push_back(0);
        for (const char *p = gold_text.data(); *p;) {
            char *q;
            const double a = strtod(p, &q);
            const char *eol = strchr(p, '\n');
            if (q != p) {
                const double b = strtod(q, &q), c = strtod(q, &q);
                gx.push_back(a);
                gy.push_back(b);
                gz.push_back(c);
            }
            p = eol ? eol + 1 : p + strlen(p);
        }
        for (uint64_t i = 0; i < ns; ++i) {
            sx.push_back(x[seeds[i]]);
            sy.push_back(y[seeds[i]]);
            sz.push_back(z[seeds[i]]);
        }
        double bb[4] = {0, 0, 0, 0};
        if (!(P.flags & OWM_F_HAVE_BBOX) && n) {
            bb[0] = bb[1] = x[0];
            bb[2] = bb[3] = y[0];
            for (uint64_t i = 1; i < n; ++i) {
                bb[0] = x[i] < bb[0] ? x[i] : bb[0];
                bb[1] = x[i] > bb[1] ? x[i] : bb[1];
                bb[2] = y[i] < bb[2] ? y[i] : bb[2];
                bb[3] = y[i] > bb[3] ? y[i] : bb[3];
            }
        }
        owm_derived d;
        uint64_t nc = 0;
        if (owm_derive(&P, n, (P.flags & OWM_F_HAVE_BBOX) ? nullptr : bb, &d) == OWM_OK &&
            owm_gold_check(sx.data(), sy.data(), sz.data(), sx.size(), gx.data(), gy.data(),
                           gz.data(), gx.size(), d.displace, &nc, &correctness) == OWM_OK)
            printf("%llu points correct; %.2f%%\n", (unsigned long long)nc, correctness);
        else
            printf("Unable to check results\n");
    }
    if (!results.empty() &&
        owm_save_time(argv[0], results.c_str(), path.c_str(), num_threads, min_radius,
                      (int)P.leaf_size, (int)level, time_tree, avg, correctness) != OWM_OK)
        printf("Unable to create time report file!\n");
    if (!out.empty()) {
        // _salida.xyz as the reference writes it (owm_optim3.cpp:242-262): the
        // LLPs in ascending order, then the stage-3 additions (here in cell
        // order), each "%.2f %.2f %.2f"
        owm_derived d;
        std::vector<uint32_t> mins(1);
        st = owm_stage1(ctx, tree, &P, mins.data(), 0, &d);
        if (st == OWM_E_CAPACITY) {
            mins.resize((size_t)d.ncols * d.nrows);
            st = owm_stage1(ctx, tree, &P, mins.data(), mins.size(), &d);
        }
        if (st != OWM_OK)
            return fail("owm_stage1", st);
        std::vector<uint32_t> llp(mins.size() + 1), adds((size_t)d.ncolsg * d.nrowsg + 1);
        uint64_t n_llp = 0, n_add = 0;
        st = owm_stage2(ctx, tree, &P, mins.data(), mins.size(), llp.data(), llp.size(), &n_llp);
        if (st != OWM_OK)
            return fail("owm_stage2", st);
        st = owm_stage3(ctx, tree, &P, llp.data(), n_llp, adds.data(), adds.size(), &n_add);
        if (st != OWM_OK)
            return fail("owm_stage3", st);
        if (n_llp + n_add != ns)
            fprintf(stderr, "warning: %llu LLPs + %llu additions, %llu seeds\n",
                    (unsigned long long)n_llp, (unsigned long long)n_add, (unsigned long long)ns);
        FILE *of = fopen(out.c_str(), "w");
        if (!of) {
            printf("Unable to create file!\n");
            owm_tree_free(tree);
            owm_ctx_destroy(ctx);
            return 1;
        }
        for (uint64_t i = 0; i < n_llp; ++i)
            fprintf(of, "%.2f %.2f %.2f\n", x[llp[i]], y[llp[i]], z[llp[i]]);
        for (uint64_t i = 0; i < n_add; ++i)
            fprintf(of, "%.2f %.2f %.2f\n", x[adds[i]], y[adds[i]], z[adds[i]]);
        if (fclose(of) != 0) {
            printf("Cannot close the file\n");
            owm_tree_free(tree);
            owm_ctx_destroy(ctx);
            return 1;
        }
    }
    owm_tree_free(tree);
    owm_ctx_destroy(ctx);
    return 0;
}
