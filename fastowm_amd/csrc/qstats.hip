// Quadtree leaf statistics: the leaves of the reference's parallel_qtree
// (CPP-TBB-CPU/src/optim3_func.cpp:218-253) as owm_stats walks them
// (CPP-TBB-CPU/application/owm_stats.cpp, CPP-TBB-CPU/include/stats_func.hpp).
//
// The reference tree does not depend on insertion order: a node of depth
// 0..level is always split; below, under the maxNumber policy
// (insertPointMaxNum, :160-184) a node is split exactly when more than
// maxNumber points reach it, and under the minRadius policy (insertPoint,
// :103-123) when at least one point reaches it and radius * 0.5 > minRadius.
// So no pointer tree is built: depth by depth, every point still in a split
// node is routed to its child (quadrantIdx, :91-100, against the child
// centres of createQuadrants, :59-73), the children's points are counted, and
// each child is either split (it gets an open-node index and a centre) or a
// leaf (its (depth, points) pair is counted).  One u32 of state per point;
// the points are first sorted by their path key (their quadrants at depths
// 1..16).  Down to depth 16 a point's quadrant is read from its key, and as a
// node's points are then contiguous, the first point of each child counts the
// child's run by a galloping search and stores it: no atomics, which at the
// top depths would all land on the same few counters.  Below depth 16 the
// quadrant comes from the coordinates against the node's centre and runs of
// equal children within a wave are added atomically, once per run.
#include "owm_kernels.h"

namespace owm {

constexpr int QS_TB = 256;

__global__ __launch_bounds__(QS_TB) void k_qs_path(const double *__restrict__ x,
                                                    const double *__restrict__ y, uint64_t n,
                                                    QsPath P, uint32_t *__restrict__ key)
{
    for (uint64_t i = (uint64_t)blockIdx.x * QS_TB + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * QS_TB) {
        const double px = x[i], py = y[i];
        double cx = P.cx, cy = P.cy;
        uint32_t k = 0;
#pragma unroll
        for (int l = 0; l < QS_KEY_LEVELS; ++l) {
            const bool qx = px >= cx, qy = py >= cy;
            k = (k << 2) | (qx ? 2u : 0u) | (qy ? 1u : 0u);
            cx += qx ? P.off[l] : -P.off[l];
            cy += qy ? P.off[l] : -P.off[l];
        }
        key[i] = k;
    }
}

__global__ void k_qs_root(QsDev D, double cx, double cy)
{
    D.remap[0] = 0;
    D.nopen[0] = 1;
    D.cen[0][0] = make_double2(cx, cy);
}

// Depth d: each point of an open depth-(d-1) node moves to child slot
// 4 * node + quadrant; points whose node became a leaf are marked settled.
__global__ __launch_bounds__(QS_TB) void k_qs_route(QsDev D, int d)
{
    const uint32_t nprev = min(D.nopen[d - 1], D.open_cap);
    if (nprev == 0)
        return;
    const double2 *cen = D.cen[(d - 1) & 1];
    const uint32_t lane = threadIdx.x & 63;
    for (uint64_t base = (uint64_t)blockIdx.x * QS_TB; base < D.n;
         base += (uint64_t)gridDim.x * QS_TB) {
        const uint64_t i = base + threadIdx.x;
        uint32_t child = OWM_NONE;
        if (i < D.n) {
            const uint32_t s = D.st[i];
            if (s != OWM_NONE) {
                const uint32_t node = D.remap[s];
                if (node == OWM_NONE) {
                    D.st[i] = OWM_NONE;
                } else {
                    uint32_t q;
                    if (D.key && d <= QS_KEY_LEVELS) {
                        q = (D.key[i] >> (2 * (QS_KEY_LEVELS - d))) & 3u;
                    } else {
                        const uint64_t pi = D.perm ? D.perm[i] : i;
                        const double2 c = cen[node];
                        q = (D.x[pi] >= c.x ? 2u : 0u) | (D.y[pi] >= c.y ? 1u : 0u);
                    }
                    child = 4u * node + q;
                    D.st[i] = child;
                }
            }
        }
        if (D.key && d <= QS_KEY_LEVELS) {
            // the first point of a child's run (in key order) counts the run
            if (child != OWM_NONE) {
                const int sh = 2 * (QS_KEY_LEVELS - d);
                const uint32_t pre = D.key[i] >> sh;
                if (i == 0 || (D.key[i - 1] >> sh) != pre) {
                    uint64_t lo = i, hi = i + 1, step = 1;  // key[lo] in the run; hi unknown
                    while (hi < D.n && (D.key[hi] >> sh) == pre) {
                        lo = hi;
                        step <<= 1;
                        hi = lo + step;
                    }
                    if (hi > D.n)
                        hi = D.n;
                    while (hi - lo > 1) {  // key[lo] in the run, key[hi] past it (or hi = n)
                        const uint64_t mid = lo + (hi - lo) / 2;
                        if ((D.key[mid] >> sh) == pre)
                            lo = mid;
                        else
                            hi = mid;
                    }
                    D.cnt[child] = (uint32_t)(hi - i);
                }
            }
            continue;
        }
        // a run of lanes with the same child adds its length once
        const uint32_t prev = __shfl_up(child, 1, 64);
        const bool starts = lane == 0 || prev != child;
        const uint64_t bnd = __ballot(starts);
        if (starts && child != OWM_NONE) {
            const uint64_t above = lane == 63 ? 0ull : (bnd & (~0ull << (lane + 1)));
            const uint32_t next = above ? (uint32_t)__builtin_ctzll(above) : 64u;
            atomicAdd(&D.cnt[child], next - lane);
        }
    }
}

// Depth d: every child slot of the open depth-(d-1) nodes is split or a leaf.
__global__ __launch_bounds__(QS_TB) void k_qs_classify(QsDev D, int d, float rd, double off)
{
    __shared__ uint32_t s_hist[QS_HB];
    const uint64_t nslots = 4ull * min(D.nopen[d - 1], D.open_cap);
    if ((uint64_t)blockIdx.x * QS_TB >= nslots)
        return;
    for (uint32_t k = threadIdx.x; k < QS_HB; k += QS_TB)
        s_hist[k] = 0;
    __syncthreads();
    const double2 *pc = D.cen[(d - 1) & 1];
    double2 *nc = D.cen[d & 1];
    const bool forced = d <= (int)D.level;
    const bool radius_split = (double)rd * 0.5 > (double)D.min_radius;
    const uint32_t lane = threadIdx.x & 63;
    for (uint64_t base = (uint64_t)blockIdx.x * QS_TB; base < nslots;
         base += (uint64_t)gridDim.x * QS_TB) {
        const uint64_t j = base + threadIdx.x;
        const bool valid = j < nslots;
        uint32_t c = 0;
        bool open = false;
        if (valid) {
            c = D.cnt[j];
            D.cnt[j] = 0;
            open = forced || (D.policy == 0 ? c > D.max_number : (c >= 1 && radius_split));
        }
        const uint64_t m = __ballot(valid && open);
        uint32_t first = 0;
        if (m) {
            if (lane == 0)
                first = atomicAdd(&D.nopen[d], (uint32_t)__popcll(m));
            first = __shfl(first, 0, 64);
        }
        if (!valid)
            continue;
        if (open) {
            const uint32_t idx = first + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (idx < D.open_cap) {
                D.remap[j] = idx;
                const double2 p = pc[j >> 2];
                const uint32_t q = (uint32_t)(j & 3);
                nc[idx] = make_double2(p.x + ((q & 2) ? off : -off), p.y + ((q & 1) ? off : -off));
            } else {
                D.remap[j] = OWM_NONE;
                atomicOr(D.flags, 2u);
            }
        } else {
            D.remap[j] = OWM_NONE;
            if (c < QS_HB) {
                atomicAdd(&s_hist[c], 1u);
            } else {
                const uint32_t o = atomicAdd(D.novf, 1u);
                if (o < D.ovf_cap)
                    D.ovf[o] = make_uint2((uint32_t)d, c);
                else
                    atomicOr(D.flags, 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < QS_HB; k += QS_TB)
        if (s_hist[k])
            atomicAdd(&D.hist[(uint64_t)d * QS_HB + k], (unsigned long long)s_hist[k]);
}

void launch_qs_root(hipStream_t s, const QsDev &D, double cx, double cy)
{
    hipLaunchKernelGGL(k_qs_root, dim3(1), dim3(1), 0, s, D, cx, cy);
}

void launch_qs_path(hipStream_t s, const double *x, const double *y, uint64_t n, const QsPath &P,
                    uint32_t *key)
{
    if (n == 0)
        return;
    uint64_t b = (n + QS_TB - 1) / QS_TB;
    b = b > 16384 ? 16384 : b;
    hipLaunchKernelGGL(k_qs_path, dim3((unsigned)b), dim3(QS_TB), 0, s, x, y, n, P, key);
}

void launch_qs_level(hipStream_t s, const QsDev &D, int d, float rd, double off)
{
    uint64_t rb = (D.n + QS_TB - 1) / QS_TB;
    rb = rb < 1 ? 1 : (rb > 16384 ? 16384 : rb);
    hipLaunchKernelGGL(k_qs_route, dim3((unsigned)rb), dim3(QS_TB), 0, s, D, d);
    // slots of the open depth-(d-1) nodes: at most 4 * open_cap
    uint64_t cb = (4ull * D.open_cap + QS_TB - 1) / QS_TB;
    cb = cb < 1 ? 1 : (cb > 2048 ? 2048 : cb);
    hipLaunchKernelGGL(k_qs_classify, dim3((unsigned)cb), dim3(QS_TB), 0, s, D, d, rd, off);
}

}  // namespace owm
