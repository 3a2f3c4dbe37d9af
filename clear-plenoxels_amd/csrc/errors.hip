// Error plumbing: every entry point returns a status and leaves a message
// here instead of printing and exiting (the reference's checkGPUStatus,
// Utils.cu:700-717).
#include <atomic>
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

// Library options (pxl_set_option): process-wide, read by each launch.
static std::atomic<int> g_marcher{PXL_MARCHER_GROUP};
static std::atomic<int> g_tv_addr64{0};
static std::atomic<int64_t> g_claim_grid{0};

bool option_group_marcher() { return g_marcher.load(std::memory_order_relaxed) == PXL_MARCHER_GROUP; }
bool option_tv_addr64() { return g_tv_addr64.load(std::memory_order_relaxed) != 0; }
int64_t option_claim_grid() { return g_claim_grid.load(std::memory_order_relaxed); }

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

extern "C" int32_t pxl_version(void) { return (1 << 16) | 2; }

extern "C" int32_t pxl_abi_sizes(int64_t *out, int32_t n) {
    const int64_t sz[9] = {(int64_t)sizeof(pxl_camera),       (int64_t)sizeof(pxl_bbox),
                           (int64_t)sizeof(pxl_rays),         (int64_t)sizeof(pxl_grid),
                           (int64_t)sizeof(pxl_optim),        (int64_t)sizeof(pxl_train_config),
                           (int64_t)sizeof(pxl_level_config), (int64_t)sizeof(pxl_level),
                           (int64_t)sizeof(pxl_schedule)};
    if (!out) return 9;
    for (int i = 0; i < n && i < 9; i++) out[i] = sz[i];
    return 9;
}

extern "C" int32_t pxl_set_option(int32_t key, int64_t value) {
    pxl::clear_error();
    switch (key) {
        case PXL_OPT_MARCHER:
            PXL_REQUIRE(value == PXL_MARCHER_GROUP || value == PXL_MARCHER_WAVE,
                        "pxl_set_option: marcher %lld (0 group, 1 wave)", (long long)value);
            pxl::g_marcher.store((int)value);
            return PXL_OK;
        case PXL_OPT_TV_ADDR64:
            PXL_REQUIRE(value == 0 || value == 1, "pxl_set_option: TV address form %lld (0, 1)",
                        (long long)value);
            pxl::g_tv_addr64.store((int)value);
            return PXL_OK;
        case PXL_OPT_CLAIM_GRID:
            PXL_REQUIRE(value >= 0 && value <= (int64_t)1 << 30,
                        "pxl_set_option: claim grid %lld (0: resident, else 1 .. 2^30 blocks)",
                        (long long)value);
            pxl::g_claim_grid.store(value);
            return PXL_OK;
        default:
            pxl::set_error("pxl_set_option: unknown option %d", key);
            return PXL_EINVAL;
    }
}

extern "C" int32_t pxl_get_option(int32_t key, int64_t *value) {
    pxl::clear_error();
    PXL_REQUIRE(value != nullptr, "pxl_get_option: null value");
    switch (key) {
        case PXL_OPT_MARCHER: *value = pxl::g_marcher.load(); return PXL_OK;
        case PXL_OPT_TV_ADDR64: *value = pxl::g_tv_addr64.load(); return PXL_OK;
        case PXL_OPT_CLAIM_GRID: *value = pxl::g_claim_grid.load(); return PXL_OK;
        default:
            pxl::set_error("pxl_get_option: unknown option %d", key);
            return PXL_EINVAL;
    }
}
