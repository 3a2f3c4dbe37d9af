// Error plumbing: every entry point returns a status and leaves a message
// here instead of printing and exiting (the reference's checkGPUStatus,
// Utils.cu:700-717).
#include <cstdarg>

#include "common.hpp"

namespace pxl {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void clear_error() { g_err[0] = '\0'; }

int32_t check_launch(const char *what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return PXL_EHIP;
    }
    return PXL_OK;
}

}  // namespace pxl

extern "C" const char *pxl_last_error(void) { return pxl::g_err; }

extern "C" int32_t pxl_version(void) { return (1 << 16) | 1; }

extern "C" int32_t pxl_abi_sizes(int64_t *out, int32_t n) {
    const int64_t sz[6] = {(int64_t)sizeof(pxl_camera), (int64_t)sizeof(pxl_bbox),
                           (int64_t)sizeof(pxl_rays),   (int64_t)sizeof(pxl_grid),
                           (int64_t)sizeof(pxl_optim),  (int64_t)sizeof(pxl_train_config)};
    if (!out) return 6;
    for (int i = 0; i < n && i < 6; i++) out[i] = sz[i];
    return 6;
}
