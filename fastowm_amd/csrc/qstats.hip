// Quadtree leaf statistics: the leaves of the reference's parallel_qtree
// (CPP-TBB-CPU/src/optim3_func.cpp:218-253) as owm_stats walks them
// (CPP-TBB-CPU/application/owm_stats.cpp, CPP-TBB-CPU/include/stats_func.hpp).
//
// The reference tree does not depend on insertion order: a node of depth
// 0..level is always split; below, under the maxNumber policy
// (insertPointMaxNum, :160-184) a node is split exactly when more than
// maxNumber points reach it, and under the minRadius policy (insertPoint,
// :103-123) when at least one point reaches it and radius * 0.5 > minRadius.
// So no pointer tree is built.  Every point gets a path key, its quadrants
// (quadrantIdx, :91-100, against the child centres of createQuadrants,
// :59-73) at depths 1..16, two bits a depth, and the keys are sorted by the
// quadrants of depths 1..S (S = 12, or 16 when the tree is expected to go
// deeper than 12).  Down to depth S a node's points are then one
// contiguous key range: depth by depth, each split ("open") node's range is
// cut into its four children's by binary searches on the keys, and each
// child is split (it gets an open-node index, a range and a centre) or a leaf
// (its (depth, points) pair is counted).  That work scales with the nodes,
// not the points.  Nodes still open at depth S go on point by point: each
// point in one holds its node and is routed to a child by its key's next
// quadrant (below depth 16 by its coordinates against the node's centre, the
// index of the point at each key position then coming from a second sort with
// values), the children's points are counted with atomics (runs of equal
// children within a wave add once) and classified as above.
#include "owm_kernels.h"

namespace owm {

constexpr int QS_TB = 256;

__global__ __launch_bounds__(QS_TB) void k_qs_path(const double *__restrict__ x,
                                                    const double *__restrict__ y, uint64_t n,
                                                    QsPath P, int sl, uint32_t *__restrict__ key)
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
        // depths 1..sl to bits 0..2sl-1, the deeper ones above
        key[i] = sl >= QS_KEY_LEVELS ? k : (k >> (2 * (QS_KEY_LEVELS - sl))) | (k << (2 * sl));
    }
}

__global__ void k_qs_root(QsDev D, double cx, double cy)
{
    D.nopen[0] = 1;
    D.cen[0][0] = make_double2(cx, cy);
    D.rng[0][0] = make_uint2(0u, (uint32_t)D.n);
}

// Leaves of one block: counts below QS_HB in LDS bins, larger ones listed.
__device__ __forceinline__ void qs_leaf(QsDev &D, uint32_t *s_hist, int d, uint32_t c)
{
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

__device__ __forceinline__ void qs_flush(QsDev &D, uint32_t *s_hist, int d)
{
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < QS_HB; k += QS_TB)
        if (s_hist[k])
            atomicAdd(&D.hist[(uint64_t)d * QS_HB + k], (unsigned long long)s_hist[k]);
}

__device__ __forceinline__ bool qs_split_rule(const QsDev &D, bool forced, bool radius_split,
                                              uint32_t c)
{
    return forced || (D.policy == 0 ? c > D.max_number : (c >= 1 && radius_split));
}

// The quadrant a path key gives at depth d (1..16), sl depths sorted.
__device__ __forceinline__ uint32_t qs_quadrant(uint32_t key, int d, int sl)
{
    return d <= sl ? (key >> (2 * (sl - d))) & 3u : (key >> (2 * sl + 2 * (QS_KEY_LEVELS - d))) & 3u;
}

// Depth d <= sort_levels: one thread per open depth-(d-1) node cuts its key
// range into its children's and classifies them; new open nodes get their
// indices from one atomic per block.
__global__ __launch_bounds__(QS_TB) void k_qs_split(QsDev D, int d, float rd, double off)
{
    __shared__ uint32_t s_hist[QS_HB];
    __shared__ uint32_t s_scan[QS_TB / 64];
    __shared__ uint32_t s_base;
    const uint32_t nprev = min(D.nopen[d - 1], D.open_cap);
    if (blockIdx.x * QS_TB >= nprev)
        return;
    for (uint32_t k = threadIdx.x; k < QS_HB; k += QS_TB)
        s_hist[k] = 0;
    const uint2 *pr = D.rng[(d - 1) & 1];
    const double2 *pc = D.cen[(d - 1) & 1];
    uint2 *nr = D.rng[d & 1];
    double2 *nc = D.cen[d & 1];
    const bool forced = d <= (int)D.level;
    const bool radius_split = (double)rd * 0.5 > (double)D.min_radius;
    const int sh = 2 * (D.sort_levels - d);
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (uint32_t base = blockIdx.x * QS_TB; base < nprev; base += gridDim.x * QS_TB) {
        const uint32_t o = base + threadIdx.x;
        uint32_t b[5] = {0, 0, 0, 0, 0};
        uint32_t nopen = 0;
        bool op[4] = {false, false, false, false};
        if (o < nprev) {
            const uint2 r = pr[o];
            b[0] = r.x;
            b[4] = r.y;
            // keys in [r.x, r.y) share the depth-(d-1) prefix; their depth-d
            // quadrant (key >> sh) & 3 ascends: b[q] = first with quadrant >= q
            for (uint32_t q = 1; q < 4; ++q) {
                uint32_t lo = b[q - 1], hi = r.y;
                while (lo < hi) {
                    const uint32_t mid = lo + (hi - lo) / 2;
                    if (((D.key[mid] >> sh) & 3u) < q)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                b[q] = lo;
            }
            for (int q = 0; q < 4; ++q) {
                op[q] = qs_split_rule(D, forced, radius_split, b[q + 1] - b[q]);
                nopen += op[q] ? 1u : 0u;
            }
        }
        // block-wide exclusive scan of the open children
        uint32_t incl = nopen;
#pragma unroll
        for (int k = 1; k < 64; k <<= 1) {
            const uint32_t v = __shfl_up(incl, k, 64);
            if (lane >= (uint32_t)k)
                incl += v;
        }
        __syncthreads();  // s_scan / s_base of the previous round are consumed
        if (lane == 63)
            s_scan[wv] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int w = 0; w < QS_TB / 64; ++w) {
                const uint32_t v = s_scan[w];
                s_scan[w] = t;
                t += v;
            }
            s_base = t ? atomicAdd(&D.nopen[d], t) : 0u;
        }
        __syncthreads();
        uint32_t idx = s_base + s_scan[wv] + incl - nopen;
        if (o < nprev) {
            const double2 p = pc[o];
            for (int q = 0; q < 4; ++q) {
                if (!op[q]) {
                    qs_leaf(D, s_hist, d, b[q + 1] - b[q]);
                    continue;
                }
                if (idx < D.open_cap) {
                    nr[idx] = make_uint2(b[q], b[q + 1]);
                    nc[idx] = make_double2(p.x + ((q & 2) ? off : -off), p.y + ((q & 1) ? off : -off));
                } else {
                    atomicOr(D.flags, 2u);
                }
                ++idx;
            }
        }
    }
    qs_flush(D, s_hist, d);
}

// Depth sort_levels still has open nodes: every point of one holds its
// index (st was filled with OWM_NONE), and remap is the identity on them.
__global__ __launch_bounds__(QS_TB) void k_qs_deep_start(QsDev D)
{
    const uint32_t nopen = min(D.nopen[D.sort_levels], D.open_cap);
    const uint2 *r = D.rng[D.sort_levels & 1];
    for (uint32_t o = blockIdx.x * QS_TB + threadIdx.x; o < nopen; o += gridDim.x * QS_TB) {
        D.remap[o] = o;
        for (uint32_t i = r[o].x; i < r[o].y; ++i)
            D.st[i] = o;
    }
}

// Depth d > sort_levels: each point of an open depth-(d-1) node moves to
// child slot 4 * node + quadrant; points whose node became a leaf are marked
// settled.  Points are in key order; below depth 16, perm gives each one's
// index.
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
                        q = qs_quadrant(D.key[i], d, D.sort_levels);
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
            open = qs_split_rule(D, forced, radius_split, c);
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
            qs_leaf(D, s_hist, d, c);
        }
    }
    qs_flush(D, s_hist, d);
}

void launch_qs_root(hipStream_t s, const QsDev &D, double cx, double cy)
{
    hipLaunchKernelGGL(k_qs_root, dim3(1), dim3(1), 0, s, D, cx, cy);
}

void launch_qs_path(hipStream_t s, const double *x, const double *y, uint64_t n, const QsPath &P,
                    int sort_levels, uint32_t *key)
{
    if (n == 0)
        return;
    uint64_t b = (n + QS_TB - 1) / QS_TB;
    b = b > 16384 ? 16384 : b;
    hipLaunchKernelGGL(k_qs_path, dim3((unsigned)b), dim3(QS_TB), 0, s, x, y, n, P, sort_levels, key);
}

void launch_qs_split(hipStream_t s, const QsDev &D, int d, float rd, double off)
{
    uint64_t b = ((uint64_t)D.open_cap + QS_TB - 1) / QS_TB;
    b = b < 1 ? 1 : (b > 4096 ? 4096 : b);
    hipLaunchKernelGGL(k_qs_split, dim3((unsigned)b), dim3(QS_TB), 0, s, D, d, rd, off);
}

void launch_qs_deep_start(hipStream_t s, const QsDev &D)
{
    launch_fill_u32(s, D.st, D.n, OWM_NONE);
    uint64_t b = ((uint64_t)D.open_cap + QS_TB - 1) / QS_TB;
    b = b < 1 ? 1 : (b > 4096 ? 4096 : b);
    hipLaunchKernelGGL(k_qs_deep_start, dim3((unsigned)b), dim3(QS_TB), 0, s, D);
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
