// .xyz reader: the reference's readXYZfile (CPP-TBB-CPU/src/optim3_func.cpp:665-722)
// with the text parsed on the device (xyz.hip).  The host detects the header,
// uploads the data section, and parses only the lines the device leaves to it
// (numbers outside the exact fast path), with strtod, as fscanf("%lf") would.
#include "owm_host.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace owm;

namespace {

uint64_t al256(uint64_t v) { return (v + 255) / 256 * 256; }

bool blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// The reference's *H.xyz header: a first line holding one number N, then
// minx, maxx, miny, maxy (fscanf "%d\n%lf\n%lf\n%lf\n%lf\n").  Returns the
// offset of the data section (0 without a header).
uint64_t read_header(const char *t, uint64_t bytes, bool &have, uint64_t &n_hdr, double box[4])
{
    have = false;
    const char *end = t + bytes;
    const char *eol = (const char *)memchr(t, '\n', bytes);
    if (!eol)
        return 0;
    std::string first(t, eol);
    char *q = nullptr;
    const unsigned long long n = strtoull(first.c_str(), &q, 10);
    if (q == first.c_str())
        return 0;
    while (*q && (blank(*q)))
        ++q;
    if (*q)
        return 0;  // the first line holds more than one number: no header
    // the four header values, one per line
    std::string rest(eol + 1, std::min<uint64_t>((uint64_t)(end - eol - 1), 4096));
    const char *p = rest.c_str();
    for (int k = 0; k < 4; ++k) {
        char *r;
        box[k] = strtod(p, &r);
        if (r == p)
            return 0;
        p = r;
    }
    have = true;
    n_hdr = n;
    return (uint64_t)(eol + 1 - t) + (uint64_t)(p - rest.c_str());
}

// strtod parse of one line: three numbers, the rest ignored.
bool host_line(const char *s, const char *end, double v[3])
{
    std::string line(s, (const char *)std::find(s, end, '\n'));
    const char *p = line.c_str();
    for (int k = 0; k < 3; ++k) {
        char *r;
        v[k] = strtod(p, &r);
        if (r == p)
            return false;
        p = r;
    }
    return true;
}

}  // namespace

extern "C" {

owm_status owm_xyz_parse(owm_ctx *ctx, const char *text, uint64_t bytes, owm_cloud *out)
{
    if (!ctx || !out || (bytes && !text))
        return OWM_E_ARG;
    try {
        (void)hipSetDevice(ctx->device);
        hipStream_t s = ctx->stream;
        memset(out, 0, sizeof *out);
        bool have = false;
        uint64_t n_hdr = 0;
        double box[4] = {0, 0, 0, 0};
        const uint64_t off = read_header(text, bytes, have, n_hdr, box);
        const char *data = text + off;
        const uint64_t len = bytes - off;
        const uint64_t chunk = xyz_chunk_bytes();
        const uint64_t body = (len + chunk - 1) / chunk * chunk;  // whole chunks
        const uint64_t padded = body + 64;
        const uint64_t nch = body / chunk;
        // text | chunk counts | chunk offsets | counters
        const uint64_t w_text = al256(padded), w_cnt = al256(nch * 4), w_off = al256((nch + 1) * 8);
        HIP_TRY(hipStreamSynchronize(s), "sync before the reader");
        HIP_TRY(ctx->xyz_text.reserve(w_text + w_cnt + w_off + 256, ctx->room_for(ctx->xyz_text)),
                "reader text buffer");
        uint8_t *dt = (uint8_t *)ctx->xyz_text.ptr;
        uint32_t *cnt = (uint32_t *)(dt + w_text);
        uint64_t *coff = (uint64_t *)(dt + w_text + w_cnt);
        uint32_t *counts = (uint32_t *)(dt + w_text + w_cnt + w_off);
        if (len)
            HIP_TRY(hipMemcpyAsync(dt, data, len, hipMemcpyHostToDevice, s), "text H2D");
        HIP_TRY(hipMemsetAsync(dt + len, 0, padded - len, s), "text padding");
        // the device part is timed as one phase: index passes and the parse
        int pi = ctx->prof.begin(s, PH_XYZ);
        launch_xyz_count(s, dt, body, cnt, coff);
        ctx->prof.end(s, pi, len, len, 2);
        uint64_t nnl = 0;
        HIP_TRY(hipMemcpyAsync(&nnl, coff + nch, 8, hipMemcpyDeviceToHost, s), "newline count");
        HIP_TRY(hipStreamSynchronize(s), "newline count");
        const uint64_t nlines = nnl + (len && data[len - 1] != '\n' ? 1 : 0);
        if (nlines >= (1ull << 32)) {
            set_err(ctx, "%llu lines: more than a cloud can index", (unsigned long long)nlines);
            return OWM_E_ARG;
        }
        // newline positions | x y z | lists (blank, host line, host start)
        uint64_t list_max = 1u << 20;
        // OWM_XYZ_LIST_CAP lowers the list capacity (used to exercise the
        // path where the host reads every line)
        if (const char *ov = getenv("OWM_XYZ_LIST_CAP"))
            list_max = std::max<uint64_t>(1, std::min<uint64_t>(list_max, strtoull(ov, nullptr, 10)));
        const uint32_t cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(nlines, 1), list_max);
        const uint64_t w_nl = al256(std::max<uint64_t>(nnl, 1) * 8);
        const uint64_t w_xyz = al256(std::max<uint64_t>(nlines, 1) * 8);
        const uint64_t w_list = al256((uint64_t)cap * 8);
        HIP_TRY(ctx->xyz_work.reserve(w_nl + 3 * w_xyz + 3 * w_list, ctx->room_for(ctx->xyz_work)),
                "reader work buffer");
        uint8_t *dw = (uint8_t *)ctx->xyz_work.ptr;
        uint64_t *nlpos = (uint64_t *)dw;
        double *px = (double *)(dw + w_nl), *py = (double *)(dw + w_nl + w_xyz),
               *pz = (double *)(dw + w_nl + 2 * w_xyz);
        uint64_t *lb = (uint64_t *)(dw + w_nl + 3 * w_xyz);
        uint64_t *lh = (uint64_t *)(dw + w_nl + 3 * w_xyz + w_list);
        uint64_t *ls = (uint64_t *)(dw + w_nl + 3 * w_xyz + 2 * w_list);
        launch_fill_u32(s, counts, 2, 0u);
        pi = ctx->prof.begin(s, PH_XYZ);
        launch_xyz_newlines(s, dt, body, coff, nlpos);
        launch_xyz_parse(s, dt, len, nlpos, nnl, nlines, px, py, pz, lb, lh, ls, cap, counts);
        // algorithmic bytes: the text read once and 24 bytes per point written
        ctx->prof.end(s, pi, 24ull * nlines, nlines, 2);
        uint32_t hc[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(hc, counts, 8, hipMemcpyDeviceToHost, s), "reader counts");
        HIP_TRY(hipStreamSynchronize(s), "reader parse");
        const uint64_t nb = hc[0], nh = hc[1];
        std::vector<double> hv;
        std::vector<uint64_t> hl;
        if (nh > cap || nb > cap) {
            // more odd lines than the lists hold: the host reads every line
            hv.reserve(nlines * 3);
            hl.reserve(nlines);
            const char *p = data, *end = data + len;
            for (uint64_t i = 0; i < nlines; ++i) {
                const char *eol = (const char *)memchr(p, '\n', (size_t)(end - p));
                const char *q = p;
                while (q < (eol ? eol : end) && (blank(*q)))
                    ++q;
                if (q < (eol ? eol : end)) {
                    double v[3];
                    if (!host_line(p, end, v)) {
                        set_err(ctx, "line %llu: expected three numbers",
                                (unsigned long long)(i + 1));
                        return OWM_E_ARG;
                    }
                    hl.push_back(i);
                    hv.insert(hv.end(), v, v + 3);
                }
                p = eol ? eol + 1 : end;
            }
        } else if (nh) {
            std::vector<uint64_t> lines(nh), starts(nh);
            HIP_TRY(hipMemcpyAsync(lines.data(), lh, nh * 8, hipMemcpyDeviceToHost, s), "host lines");
            HIP_TRY(hipMemcpyAsync(starts.data(), ls, nh * 8, hipMemcpyDeviceToHost, s), "host lines");
            HIP_TRY(hipStreamSynchronize(s), "host lines");
            hv.resize(nh * 3);
            for (uint64_t k = 0; k < nh; ++k)
                if (!host_line(data + starts[k], data + len, &hv[3 * k])) {
                    set_err(ctx, "line %llu: expected three numbers",
                            (unsigned long long)(lines[k] + 1));
                    return OWM_E_ARG;
                }
            hl = lines;
        }
        if (!hl.empty()) {
            DevBuf tmp;
            const uint64_t m = hl.size();
            HIP_TRY(tmp.reserve(al256(m * 8) + m * 24, al256(m * 8) + m * 24), "host values");
            uint64_t *dl = (uint64_t *)tmp.ptr;
            double *dv = (double *)((uint8_t *)tmp.ptr + al256(m * 8));
            HIP_TRY(hipMemcpyAsync(dl, hl.data(), m * 8, hipMemcpyHostToDevice, s), "host values");
            HIP_TRY(hipMemcpyAsync(dv, hv.data(), m * 24, hipMemcpyHostToDevice, s), "host values");
            launch_xyz_scatter(s, dl, dv, m, px, py, pz);
            HIP_TRY(hipStreamSynchronize(s), "host values");
        }
        uint64_t n = nlines;
        const double *ox = px, *oy = py, *oz = pz;
        if (nb) {
            // blank lines: every line that is neither a point nor host-parsed
            std::vector<uint64_t> bl;
            if (nb <= cap) {
                bl.resize(nb);
                HIP_TRY(hipMemcpyAsync(bl.data(), lb, nb * 8, hipMemcpyDeviceToHost, s), "blank lines");
                HIP_TRY(hipStreamSynchronize(s), "blank lines");
            } else {
                const char *p = data, *end = data + len;
                for (uint64_t i = 0; i < nlines; ++i) {
                    const char *eol = (const char *)memchr(p, '\n', (size_t)(end - p));
                    const char *q = p, *e = eol ? eol : end;
                    while (q < e && blank(*q))
                        ++q;
                    if (q == e)
                        bl.push_back(i);
                    p = eol ? eol + 1 : end;
                }
            }
            std::sort(bl.begin(), bl.end());
            const uint64_t w_c = al256(std::max<uint64_t>(nlines, 1) * 8);
            HIP_TRY(ctx->xyz_cloud.reserve(3 * w_c + al256(bl.size() * 8),
                                           ctx->room_for(ctx->xyz_cloud)),
                    "reader cloud");
            uint8_t *dc = (uint8_t *)ctx->xyz_cloud.ptr;
            uint64_t *dbl = (uint64_t *)(dc + 3 * w_c);
            HIP_TRY(hipMemcpyAsync(dbl, bl.data(), bl.size() * 8, hipMemcpyHostToDevice, s),
                    "blank lines");
            launch_xyz_compact(s, px, py, pz, nlines, dbl, bl.size(), (double *)dc,
                               (double *)(dc + w_c), (double *)(dc + 2 * w_c));
            HIP_TRY(hipStreamSynchronize(s), "compaction");
            ox = (const double *)dc;
            oy = (const double *)(dc + w_c);
            oz = (const double *)(dc + 2 * w_c);
            n = nlines - bl.size();
        }
        if (have) {
            if (n_hdr > n) {
                set_err(ctx, "the header announces %llu points, the file holds %llu",
                        (unsigned long long)n_hdr, (unsigned long long)n);
                return OWM_E_ARG;
            }
            n = n_hdr;  // the reference reads N points and no more
            out->flags = OWM_F_HAVE_BBOX;
            memcpy(out->bbox, box, sizeof box);
        }
        out->x = ox;
        out->y = oy;
        out->z = oz;
        out->n = n;
        out->host_lines = hl.size();
        return OWM_OK;
    } catch (const std::bad_alloc &) {
        set_err(ctx, "host allocation failed");
        return OWM_E_NOMEM;
    } catch (...) {
        set_err(ctx, "internal error");
        return OWM_E_INTERNAL;
    }
}

owm_status owm_cloud_read(owm_ctx *ctx, const owm_cloud *cloud, double *x, double *y, double *z)
{
    if (!ctx || !cloud || (cloud->n && (!x || !y || !z || !cloud->x || !cloud->y || !cloud->z)))
        return OWM_E_ARG;
    try {
        (void)hipSetDevice(ctx->device);
        hipStream_t s = ctx->stream;
        if (cloud->n) {
            HIP_TRY(hipMemcpyAsync(x, cloud->x, cloud->n * 8, hipMemcpyDeviceToHost, s), "cloud D2H");
            HIP_TRY(hipMemcpyAsync(y, cloud->y, cloud->n * 8, hipMemcpyDeviceToHost, s), "cloud D2H");
            HIP_TRY(hipMemcpyAsync(z, cloud->z, cloud->n * 8, hipMemcpyDeviceToHost, s), "cloud D2H");
        }
        HIP_TRY(hipStreamSynchronize(s), "cloud D2H");
        return OWM_OK;
    } catch (...) {
        set_err(ctx, "internal error");
        return OWM_E_INTERNAL;
    }
}

// The reference reader's table of its own header-less clouds
// (CPP-TBB-CPU/src/optim3_func.cpp:667-680): a file whose name contains one of
// these names holds that many points over that box.  The values are the
// reference's (data about its published clouds), kept as it lists them.
owm_status owm_xyz_known_cloud(const char *path, uint64_t *n_points, double *bbox)
{
    if (!path || !n_points || !bbox)
        return OWM_E_ARG;
    struct Known {
        const char *name;
        uint64_t n;
        double minx, miny, maxx, maxy;
    };
    static const Known table[] = {
        {"INAER_2011_Alcoy.xyz", 2772832, 715244.96, 4286623.63, 716057.75, 4287447.70},
        {"INAER_2011_Alcoy_Core.xyz", 20380212, 714947.98, 4286501.93, 716361.06, 4288406.23},
        {"BABCOCK_2017_Arzua_3B.xyz", 40706503, 568000.00, 4752320.00, 568999.99, 4753319.99},
        {"V21_group1_densified_point_cloud.xyz", 42384876, 526964.093, 4742610.292, 527664.647,
         4743115.738},
        {"V19_group1_densified_point_cloud.xyz", 48024480, 526955.908, 4742586.025, 527686.445,
         4743124.373},
        {"sample24.xyz", 7492, 513748.12, 5403124.76, 513869.97, 5403197.20},
    };
    for (const Known &k : table)
        if (strstr(path, k.name)) {
            *n_points = k.n;
            bbox[0] = k.minx;
            bbox[1] = k.maxx;
            bbox[2] = k.miny;
            bbox[3] = k.maxy;
            return OWM_OK;
        }
    return OWM_E_NOTFOUND;
}

}  // extern "C"
