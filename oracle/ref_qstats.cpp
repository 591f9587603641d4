// Driver for the reference's own quadtree code -- TEST INFRASTRUCTURE ONLY.
//
// Links CPP-TBB-CPU/src/optim1_qtree_func.cpp (compiled where it lies, by
// `make -C oracle ref`) and builds its minRadius quadtree with insertPointF
// (:250-274) and createQuadrantsF (:215-237), one point at a time in index
// order from a root with the centre and float radius of getRadius/getCenter
// (CPP-TBB-CPU/src/optim3_func.cpp:6-26).  It then walks the leaves as
// owm_stats does (CPP-TBB-CPU/include/stats_func.hpp:31-56) and prints one
// line "depth points leaves" per distinct (depth, points) pair, ascending.
// With level 0 this is the tree that parallel_qtree builds under the
// minRadius policy whenever the root holds a point and radius * 0.5 > minRadius.
//
// Input file: uint64 n, double minx maxx miny maxy, float minRadius, then n
// doubles x and n doubles y.
#include "optim1_qtree_func.h"

#include <cstdio>
#include <map>
#include <utility>

static void walk(Qtree q, unsigned depth, std::map<std::pair<unsigned, size_t>, unsigned long long> &cls)
{
    if (isLeaf(q)) {
        ++cls[{depth, q->points.size()}];
        return;
    }
    for (int i = 0; i < 4; ++i)
        walk(q->quadrants[i], depth + 1, cls);
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: %s points.bin\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f)
        return 2;
    unsigned long long n = 0;
    double box[4];
    float min_radius = 0.f;
    if (fread(&n, 8, 1, f) != 1 || fread(box, 8, 4, f) != 4 || fread(&min_radius, 4, 1, f) != 1)
        return 2;
    std::vector<Lpoint> pts(n);
    std::vector<double> xs(n), ys(n);
    if (n && (fread(xs.data(), 8, n, f) != n || fread(ys.data(), 8, n, f) != n))
        return 2;
    fclose(f);
    for (unsigned long long i = 0; i < n; ++i)
        pts[i] = Lpoint{(unsigned)(i + 1), xs[i], ys[i], 0.0};
    const double rx = (box[1] - box[0]) / 2.0, ry = (box[3] - box[2]) / 2.0;
    const float max_radius = rx >= ry ? rx : ry;
    Vector2D center;
    center.x = box[0] + rx;
    center.y = box[2] + ry;
    Qtree root = new Qtree_t(center, max_radius);
    for (unsigned long long i = 0; i < n; ++i)
        insertPointF(&pts[i], root, min_radius);
    std::map<std::pair<unsigned, size_t>, unsigned long long> cls;
    walk(root, 0, cls);
    for (const auto &c : cls)
        printf("%u %zu %llu\n", c.first.first, c.first.second, c.second);
    return 0;
}
