// LiDAR .xyz text -> fp64 SoA cloud on the device.
//
// The reference reads its clouds with fscanf("%lf %lf %lf") and skips the
// rest of each line (readXYZfile, CPP-TBB-CPU/src/optim3_func.cpp:665-722).
// Here the text of the data section sits in device memory: one pass counts
// the newlines of every 64 KB chunk, a scan turns the counts into line
// numbers, a second pass writes every newline's position, and a wave takes 64
// lines at a time, copies their text to LDS and parses one line per lane.  A
// number is converted exactly when its decimal mantissa M (up to 19 digits)
// and exponent q satisfy M <= 2^53 and |q| <= 22: M and 10^|q| are then exact
// doubles and one IEEE multiply or divide gives the correctly rounded value
// (Clinger, "How to read floating point numbers accurately", PLDI 1990),
// i.e. the double strtod and fscanf return.  Any other line (more digits,
// large exponents, inf/nan, hex, odd separators) is listed for the host,
// which parses it with strtod; blank lines are listed too and dropped by a
// compaction.
#include "owm_kernels.h"

namespace owm {

constexpr int XY_TB = 256;
constexpr uint32_t XY_PER_THREAD = 64;                    // bytes per thread and round
constexpr uint32_t XY_ROUNDS = 4;                         // rounds per chunk
constexpr uint32_t XY_CHUNK = XY_TB * XY_PER_THREAD * XY_ROUNDS;  // 64 KB

// 10^0 .. 10^22, every one an exact double
__constant__ double XY_P10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                  1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                  1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

// '\n' bytes in a 16-byte word (exact per byte: no carry crosses a byte)
__device__ __forceinline__ uint32_t nl_count16(uint4 w)
{
    uint32_t c = 0;
    const uint32_t v[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t x = v[k] ^ 0x0A0A0A0Au;  // a newline byte becomes zero
        const uint32_t nz = (((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
        c += (uint32_t)__popc(~nz & 0x80808080u);
    }
    return c;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *s_w)
{
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        v += __shfl_xor(v, o, 64);
    if (lane == 0)
        s_w[w] = v;
    __syncthreads();
    uint32_t t = 0;
    for (int k = 0; k < XY_TB / 64; ++k)
        t += s_w[k];
    __syncthreads();
    return t;
}

// Newlines per 64 KB chunk (the buffer is zero-padded to whole chunks).
__global__ __launch_bounds__(XY_TB) void k_xyz_count(const uint4 *__restrict__ text,
                                                     uint64_t nchunks, uint32_t *chunk_cnt)
{
    __shared__ uint32_t s_w[XY_TB / 64];
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        uint32_t n = 0;
#pragma unroll
        for (uint32_t r = 0; r < XY_ROUNDS; ++r) {
            const uint4 *p =
                text + (c * XY_CHUNK + (r * XY_TB + threadIdx.x) * XY_PER_THREAD) / 16;
#pragma unroll
            for (int j = 0; j < (int)(XY_PER_THREAD / 16); ++j)
                n += nl_count16(p[j]);
        }
        n = block_sum(n, s_w);
        if (threadIdx.x == 0)
            chunk_cnt[c] = n;
    }
}

// Exclusive scan of the chunk counts in one workgroup; out[n] = total.
__global__ __launch_bounds__(1024) void k_xyz_scan(const uint32_t *__restrict__ in, uint64_t n,
                                                   uint64_t *__restrict__ out)
{
    __shared__ uint64_t s_part[1024];
    const uint32_t t = threadIdx.x;
    const uint64_t per = (n + 1023) / 1024, a = t * per, b = a + per < n ? a + per : n;
    uint64_t sum = 0;
#pragma unroll 8
    for (uint64_t i = a; i < b; ++i)
        sum += in[i];
    s_part[t] = sum;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const uint64_t v = t >= o ? s_part[t - o] : 0;
        __syncthreads();
        s_part[t] += v;
        __syncthreads();
    }
    uint64_t run = s_part[t] - sum;
    for (uint64_t i = a; i < b; ++i) {
        out[i] = run;
        run += in[i];
    }
    if (t == 1023)
        out[n] = s_part[1023];
}

// Position of every newline, in order: nlpos[k] = offset of the k-th '\n'.
__global__ __launch_bounds__(XY_TB) void k_xyz_newlines(const uint4 *__restrict__ text,
                                                        uint64_t nchunks,
                                                        const uint64_t *__restrict__ chunk_off,
                                                        uint64_t *__restrict__ nlpos)
{
    __shared__ uint32_t s_c[XY_TB];
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        uint64_t run = chunk_off[c];  // newlines before this round
        for (uint32_t r = 0; r < XY_ROUNDS; ++r) {
            const uint64_t base = c * XY_CHUNK + (r * XY_TB + threadIdx.x) * XY_PER_THREAD;
            const uint4 *p = text + base / 16;
            uint4 w[XY_PER_THREAD / 16];
            uint32_t n = 0;
#pragma unroll
            for (int j = 0; j < (int)(XY_PER_THREAD / 16); ++j) {
                w[j] = p[j];
                n += nl_count16(w[j]);
            }
            s_c[threadIdx.x] = n;
            __syncthreads();
            // inclusive prefix over the block's threads (a simple LDS scan)
            for (uint32_t o = 1; o < XY_TB; o <<= 1) {
                const uint32_t v = threadIdx.x >= o ? s_c[threadIdx.x - o] : 0u;
                __syncthreads();
                s_c[threadIdx.x] += v;
                __syncthreads();
            }
            uint64_t k = run + s_c[threadIdx.x] - n;
            run += s_c[XY_TB - 1];
            __syncthreads();
#pragma unroll
            for (int j = 0; j < (int)(XY_PER_THREAD / 16); ++j) {
                const uint32_t v[4] = {w[j].x, w[j].y, w[j].z, w[j].w};
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        if (((v[q] >> (8 * b)) & 255u) == 10u)
                            nlpos[k++] = base + 16 * j + 4 * q + b;
            }
        }
    }
}

// Byte readers over 8-byte words of the text: from global memory (the buffer
// is padded, so the word holding the last byte is always in bounds) or from
// a wave's LDS copy of its lines (words from base_w on).
struct GlobalWords {
    const unsigned long long *w;
    __device__ __forceinline__ unsigned long long load(uint64_t wi) const { return w[wi]; }
};
constexpr uint32_t XY_WBUF = 4096;  // LDS bytes of text per wave
struct LdsWords {  // positions relative to the wave's copy (zero past its buffer)
    const unsigned long long *w;
    __device__ __forceinline__ unsigned long long load(uint32_t wi) const
    {
        return wi < XY_WBUF / 8 ? w[wi] : 0ull;
    }
};

// Pos: the position type (64-bit in the whole text, 32-bit in an LDS copy)
template <typename Src, typename Pos>
struct TextCursor {
    Src src;
    Pos cached = ~(Pos)0;
    unsigned long long word = 0;
    __device__ __forceinline__ uint32_t at(Pos p)
    {
        if ((p >> 3) != cached) {
            cached = p >> 3;
            word = src.load(cached);
        }
        return (uint32_t)(word >> (8 * (uint32_t)(p & 7))) & 255u;
    }
    // bytes p .. p+7, text[p] in the low byte
    __device__ __forceinline__ unsigned long long word8(Pos p)
    {
        const uint32_t sh = 8 * (uint32_t)(p & 7);
        const unsigned long long w0 = src.load(p >> 3);
        if (sh == 0)
            return w0;
        return (w0 >> sh) | (src.load((p >> 3) + 1) << (64 - sh));
    }
};

__constant__ unsigned long long XY_P10U[9] = {1ull,      10ull,      100ull,      1000ull,    10000ull,
                                              100000ull, 1000000ull, 10000000ull, 100000000ull};

// The digit run at p (up to e), eight bytes a step: a byte is a digit when
// its high nibble is 3 and adding 6 keeps it 3 (a carry from a lower byte
// only reaches bytes past the first non-digit); the run's digits, most
// significant in the low byte, are shifted up so that the missing ones read
// as leading zeros and converted with three multiplies (pairs, quads, eight:
// the method of Lemire's and simdjson's eight-digit parsing).  m and sig
// follow the byte loop exactly: digits before the first non-zero one are not
// significant, and more than 19 significant digits return false.
template <typename Cur, typename Pos>
__device__ __forceinline__ bool xyz_digits(Cur &T, Pos &p, Pos e, unsigned long long &m, int &sig,
                                           int &count)
{
    while (p < e) {
        const unsigned long long w = T.word8(p);
        const unsigned long long bad = ((w & 0xF0F0F0F0F0F0F0F0ull) ^ 0x3030303030303030ull) |
                                       (((w + 0x0606060606060606ull) & 0xF0F0F0F0F0F0F0F0ull) ^
                                        0x3030303030303030ull);
        uint32_t n = bad ? (uint32_t)__builtin_ctzll(bad) >> 3 : 8u;
        if ((Pos)n > e - p)
            n = (uint32_t)(e - p);
        if (n == 0)
            break;
        const unsigned long long x = w - 0x3030303030303030ull;  // digit values in bytes < n
        const unsigned long long low = n == 8 ? ~0ull : (1ull << (8 * n)) - 1ull;
        if (m == 0) {
            const unsigned long long nz = (x + 0x7F7F7F7F7F7F7F7Full) & 0x8080808080808080ull & low;
            if (nz)
                sig += (int)n - (int)(__builtin_ctzll(nz) >> 3);
        } else {
            sig += (int)n;
        }
        if (sig > 19)
            return false;
        unsigned long long v = (x & low) << (8 * (8 - n));
        v = (v * 2561ull) >> 8;
        v = ((v & 0x00FF00FF00FF00FFull) * 6553601ull) >> 16;
        v = ((v & 0x0000FFFF0000FFFFull) * 42949672960001ull) >> 32;
        m = m * XY_P10U[n] + v;
        count += (int)n;
        p += n;
        if (n < 8)
            break;
    }
    return true;
}

__device__ __forceinline__ bool xyz_space(uint32_t c)
{
    return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f';
}

// One number at p (leading blanks skipped); false: not one the device
// converts exactly (the host then parses the line).
template <typename Cur, typename Pos>
__device__ __forceinline__ bool xyz_number(Cur &T, Pos &p, Pos e, double &out)
{
    while (p < e && xyz_space(T.at(p)))
        ++p;
    if (p >= e)
        return false;
    uint32_t c = T.at(p);
    const bool neg = c == '-';
    if (c == '-' || c == '+') {
        ++p;
        c = p < e ? T.at(p) : 0u;
    }
    unsigned long long m = 0;
    int sig = 0, frac = 0, digits = 0;
    if (!xyz_digits(T, p, e, m, sig, digits))
        return false;
    if (p < e && T.at(p) == '.') {
        ++p;
        if (!xyz_digits(T, p, e, m, sig, frac))
            return false;
        digits += frac;
    }
    if (digits == 0)
        return false;
    int ex = 0;
    if (p < e && ((c = T.at(p)) == 'e' || c == 'E')) {
        ++p;
        bool eneg = false;
        if (p < e && ((c = T.at(p)) == '-' || c == '+')) {
            eneg = c == '-';
            ++p;
        }
        int ed = 0;
        while (p < e && (c = T.at(p)) >= '0' && c <= '9') {
            if (++ed > 4)
                return false;
            ex = ex * 10 + (int)(c - '0');
            ++p;
        }
        if (ed == 0)
            return false;
        ex = eneg ? -ex : ex;
    }
    // a number ends at a blank or at the end of the line
    if (p < e && !xyz_space(T.at(p)))
        return false;
    double v;
    if (m == 0) {
        v = 0.0;
    } else {
        const int q = ex - frac;
        if (m > (1ull << 53) || q < -22 || q > 22)
            return false;
        v = (double)m;
        v = q < 0 ? __ddiv_rn(v, XY_P10[-q]) : __dmul_rn(v, XY_P10[q]);
    }
    out = neg ? -v : v;
    return true;
}

struct ParseOut {
    double *x, *y, *z;
    uint64_t *blank, *host, *host_start;
    uint32_t cap;
    uint32_t *counts;
};

// Line i = [s, e): a point, a blank line (listed) or a line left to the host
// (listed with its start); lists hold cap entries, counts[0..1] count past it.
template <typename Src, typename Pos>
__device__ __forceinline__ void xyz_line(Src src, uint64_t i, uint64_t s, Pos ps, Pos e,
                                         const ParseOut &o)
{
    TextCursor<Src, Pos> T{src};
    Pos p = ps;
    while (p < e && xyz_space(T.at(p)))
        ++p;
    if (p >= e) {
        const uint32_t slot = atomicAdd(&o.counts[0], 1u);
        if (slot < o.cap)
            o.blank[slot] = i;
        return;
    }
    double v[3];
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        ok = ok && xyz_number(T, p, e, v[k]);
    if (!ok) {
        const uint32_t slot = atomicAdd(&o.counts[1], 1u);
        if (slot < o.cap) {
            o.host[slot] = i;
            o.host_start[slot] = s;
        }
        return;
    }
    o.x[i] = v[0];
    o.y[i] = v[1];
    o.z[i] = v[2];
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src)
{
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, 64);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, 64);
    return ((uint64_t)hi << 32) | lo;
}

// A wave takes 64 consecutive lines: when their text fits XY_WBUF bytes it is
// copied to LDS with 16-byte loads (the whole wave, coalesced) and every lane
// parses its line from there; longer lines are read from global memory.
__global__ __launch_bounds__(XY_TB) void k_xyz_parse(const uint4 *__restrict__ text, uint64_t bytes,
                                                     const uint64_t *__restrict__ nlpos,
                                                     uint64_t nnl, uint64_t nlines, ParseOut o)
{
    __shared__ uint4 s_txt[XY_TB / 64][XY_WBUF / 16];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t stride = (uint64_t)gridDim.x * XY_TB;
    for (uint64_t i0 = ((uint64_t)blockIdx.x * (XY_TB / 64) + w) * 64; i0 < nlines; i0 += stride) {
        const uint64_t i = i0 + lane;
        const bool valid = i < nlines;
        uint64_t s = 0, e = 0;
        if (valid) {
            s = i ? nlpos[i - 1] + 1 : 0;
            e = i < nnl ? nlpos[i] : bytes;
        }
        const uint64_t last = nlines - 1 - i0 < 63 ? nlines - 1 - i0 : 63;
        const uint64_t base = shfl_u64(s, 0) & ~15ull, span = shfl_u64(e, (int)last) - base;
        if (span <= XY_WBUF) {
            for (uint64_t off = lane * 16; off < span; off += 64 * 16)
                s_txt[w][off >> 4] = text[(base + off) >> 4];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            if (valid)
                xyz_line(LdsWords{(const unsigned long long *)s_txt[w]}, i, s,
                         (uint32_t)(s - base), (uint32_t)(e - base), o);
            // the wave's next lines overwrite the buffer
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        } else if (valid) {
            xyz_line(GlobalWords{(const unsigned long long *)text}, i, s, s, e, o);
        }
    }
}

// The host's values for the lines it parsed.
__global__ void k_xyz_scatter(const uint64_t *__restrict__ line, const double *__restrict__ v,
                              uint64_t n, double *x, double *y, double *z)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = line[k];
        x[i] = v[3 * k];
        y[i] = v[3 * k + 1];
        z[i] = v[3 * k + 2];
    }
}

// Drops the blank lines (blank: ascending line numbers): point
// i - #{blank < i} takes line i.
__global__ void k_xyz_compact(const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ z, uint64_t nlines,
                              const uint64_t *__restrict__ blank, uint64_t nb, double *ox,
                              double *oy, double *oz)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nlines;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t lo = 0, hi = nb;  // #{blank < i}
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (blank[mid] < i)
                lo = mid + 1;
            else
                hi = mid;
        }
        if (lo < nb && blank[lo] == i)
            continue;
        ox[i - lo] = x[i];
        oy[i - lo] = y[i];
        oz[i - lo] = z[i];
    }
}

static uint32_t xy_grid(uint64_t work, uint32_t per, uint32_t cap)
{
    const uint64_t g = (work + per - 1) / per;
    return g < 1 ? 1u : (g > cap ? cap : (uint32_t)g);
}

uint64_t xyz_chunk_bytes() { return XY_CHUNK; }

void launch_xyz_count(hipStream_t s, const void *text, uint64_t padded_bytes, uint32_t *chunk_cnt,
                      uint64_t *chunk_off)
{
    const uint64_t nch = padded_bytes / XY_CHUNK;
    hipLaunchKernelGGL(k_xyz_count, dim3(xy_grid(nch, 1, 16384)), dim3(XY_TB), 0, s,
                       (const uint4 *)text, nch, chunk_cnt);
    hipLaunchKernelGGL(k_xyz_scan, dim3(1), dim3(1024), 0, s, (const uint32_t *)chunk_cnt, nch,
                       chunk_off);
}

void launch_xyz_newlines(hipStream_t s, const void *text, uint64_t padded_bytes,
                         const uint64_t *chunk_off, uint64_t *nlpos)
{
    const uint64_t nch = padded_bytes / XY_CHUNK;
    hipLaunchKernelGGL(k_xyz_newlines, dim3(xy_grid(nch, 1, 16384)), dim3(XY_TB), 0, s,
                       (const uint4 *)text, nch, chunk_off, nlpos);
}

void launch_xyz_parse(hipStream_t s, const void *text, uint64_t bytes, const uint64_t *nlpos,
                      uint64_t nnl, uint64_t nlines, double *x, double *y, double *z,
                      uint64_t *blank, uint64_t *host, uint64_t *host_start, uint32_t cap,
                      uint32_t *counts)
{
    if (nlines == 0)
        return;
    const ParseOut o{x, y, z, blank, host, host_start, cap, counts};
    hipLaunchKernelGGL(k_xyz_parse, dim3(xy_grid(nlines, XY_TB, 16384)), dim3(XY_TB), 0, s,
                       (const uint4 *)text, bytes, nlpos, nnl, nlines, o);
}

void launch_xyz_scatter(hipStream_t s, const uint64_t *line, const double *v, uint64_t n, double *x,
                        double *y, double *z)
{
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_xyz_scatter, dim3(xy_grid(n, 256, 4096)), dim3(256), 0, s, line, v, n, x,
                       y, z);
}

void launch_xyz_compact(hipStream_t s, const double *x, const double *y, const double *z,
                        uint64_t nlines, const uint64_t *blank, uint64_t nb, double *ox, double *oy,
                        double *oz)
{
    if (nlines == 0)
        return;
    hipLaunchKernelGGL(k_xyz_compact, dim3(xy_grid(nlines, 256, 16384)), dim3(256), 0, s, x, y, z,
                       nlines, blank, nb, ox, oy, oz);
}

}  // namespace owm
