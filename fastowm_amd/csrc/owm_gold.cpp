// GOLD comparison of a seed list: the reference's check_results
// (CPP-TBB-CPU/src/optim3_func.cpp:725-794).  A seed is correct when a GOLD
// point lies strictly inside the square of half side Displace/2 (a float, as
// makeBox takes it) around the seed and their z differ by less than 0.01; the
// result is correct seeds / GOLD points x 100.  The reference compares every
// seed with every GOLD point; here the GOLD points are bucketed in a grid of
// cells one half side wide, so a seed looks at the 5x5 cells around it (the
// 3x3 that geometry needs, widened against rounding of the cell index).  A
// seed counts once either way, so the count is the same.  Host code only.
#include "owm_host.h"

#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <unordered_map>
#include <vector>

extern "C" {

owm_status owm_gold_check(const double *sx, const double *sy, const double *sz, uint64_t n_seeds,
                          const double *gx, const double *gy, const double *gz, uint64_t n_gold,
                          double displace, uint64_t *n_correct, double *percent)
{
    if (!n_correct || !percent || (n_seeds && (!sx || !sy || !sz)) ||
        (n_gold && (!gx || !gy || !gz)) || !(displace > 0.0) || !std::isfinite(displace))
        return OWM_E_ARG;
    try {
        // the half side as check_results computes it: float Displace * 0.5,
        // passed to makeBox as a float
        const double r = (double)(float)((double)(float)displace * 0.5);
        const double cell = r > 0.0 ? r : 1.0;
        auto key = [cell](double x, double y) {
            const long long i = (long long)std::floor(x / cell), j = (long long)std::floor(y / cell);
            return ((unsigned long long)i << 32) ^ (unsigned long long)(uint32_t)j;
        };
        std::unordered_map<unsigned long long, std::vector<uint64_t>> grid;
        grid.reserve((size_t)n_gold);
        for (uint64_t k = 0; k < n_gold; ++k)
            if (std::isfinite(gx[k]) && std::isfinite(gy[k]))
                grid[key(gx[k], gy[k])].push_back(k);
        uint64_t count = 0;
        for (uint64_t s = 0; s < n_seeds; ++s) {
            const double x0 = sx[s] - r, x1 = sx[s] + r, y0 = sy[s] - r, y1 = sy[s] + r;
            if (!std::isfinite(x0) || !std::isfinite(y0))
                continue;
            const long long ci = (long long)std::floor(sx[s] / cell),
                            cj = (long long)std::floor(sy[s] / cell);
            bool hit = false;
            for (long long di = -2; di <= 2 && !hit; ++di)
                for (long long dj = -2; dj <= 2 && !hit; ++dj) {
                    const unsigned long long kk =
                        ((unsigned long long)(ci + di) << 32) ^ (unsigned long long)(uint32_t)(cj + dj);
                    auto it = grid.find(kk);
                    if (it == grid.end())
                        continue;
                    for (uint64_t k : it->second)
                        if (gx[k] > x0 && gy[k] > y0 && gx[k] < x1 && gy[k] < y1 &&
                            std::fabs(sz[s] - gz[k]) < 0.01) {
                            hit = true;
                            break;
                        }
                }
            count += hit ? 1 : 0;
        }
        *n_correct = count;
        *percent = n_gold ? (double)count / (double)n_gold * 100.0 : 0.0;
        return OWM_OK;
    } catch (...) {
        return OWM_E_NOMEM;
    }
}

}  // extern "C"

// The driver's results row, save_time (CPP-TBB-CPU/src/optim3_func.cpp:797-828):
// a header line when the file does not exist yet, then one space-separated row.
// The reference streams MinRadius and Correctness in defaultfloat (printf %g
// at the stream's default precision 6) and the two times in fixed (%.6f).
extern "C" owm_status owm_save_time(const char *exec_name, const char *results_path,
                                    const char *input_file, int num_threads, float min_radius,
                                    int max_number, int level, double tree_time,
                                    double owm_time, double correctness)
{
    if (!results_path || !exec_name || !input_file)
        return OWM_E_ARG;
    struct stat sb;
    const bool fresh = stat(results_path, &sb) != 0;
    FILE *f = fopen(results_path, "a");
    if (!f)
        return OWM_E_ARG;
    if (fresh)
        fputs("ExecName InputFile NumThreads MinRadius MaxNumber Level TreeTime OwmTime Correctness\n",
              f);
    fprintf(f, "%s %s %d %g %d %d %.6f %.6f %g\n", exec_name, input_file, num_threads,
            (double)min_radius, max_number, level, tree_time, owm_time, correctness);
    const bool ok = fclose(f) == 0;
    return ok ? OWM_OK : OWM_E_ARG;
}
