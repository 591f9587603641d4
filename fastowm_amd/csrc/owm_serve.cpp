// Serving loop (continuous batching) -- entry points of include/owm.h.
#include "owm_host.h"

namespace owm {
struct ServeState {};
}  // namespace owm

owm_ctx::~owm_ctx()
{
    if (h_tiles)
        (void)hipHostFree(h_tiles);
    for (auto &p : h_off)
        if (p)
            (void)hipHostFree(p);
    if (stream)
        (void)hipStreamDestroy(stream);
}


extern "C" {
owm_status owm_serve_start(owm_ctx *ctx, uint32_t, uint32_t)
{
    owm::set_err(ctx, "serving is not available in this build");
    return OWM_E_ARG;
}
owm_status owm_serve_submit(owm_ctx *ctx, const double *, const double *, const double *, uint64_t,
                            const owm_params *, uint64_t, uint64_t *)
{
    owm::set_err(ctx, "serving is not available in this build");
    return OWM_E_ARG;
}
owm_status owm_serve_poll(owm_ctx *, uint64_t, uint32_t *, uint64_t, uint64_t *, double *) { return OWM_E_NOTFOUND; }
owm_status owm_serve_wait(owm_ctx *, uint64_t, uint32_t *, uint64_t, uint64_t *, double *) { return OWM_E_NOTFOUND; }
owm_status owm_serve_cancel(owm_ctx *, uint64_t) { return OWM_E_NOTFOUND; }
owm_status owm_serve_evict(owm_ctx *, uint64_t) { return OWM_E_NOTFOUND; }
owm_status owm_serve_stop(owm_ctx *) { return OWM_OK; }
}
