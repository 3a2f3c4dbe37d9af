// Training-ray order on the host: ShuffleRank (Utils.cu:488-506).  It draws
// from the C library's rand(), as the reference does, so a caller that seeds
// with srand(seed) and makes the same rand() calls gets the reference's
// order on the same C library.
#include <cstdlib>

#include "common.hpp"

extern "C" int32_t pxl_shuffle_rank(int32_t *dest, int32_t max_index, int32_t regenerate,
                                    int32_t min_index, int32_t shuffle) {
    using namespace pxl;
    clear_error();
    PXL_REQUIRE(dest, "pxl_shuffle_rank: null destination");
    PXL_REQUIRE(min_index >= 0 && max_index >= min_index,
                "pxl_shuffle_rank: bad range [%d, %d)", min_index, max_index);
    // The draw lands in [min_index, min_index + i), which leaves the range
    // for i = max_index - 1 unless min_index <= 1 (the reference only uses 0).
    PXL_REQUIRE(!shuffle || min_index <= 1,
                "pxl_shuffle_rank: min_index %d would draw positions up to %d, past %d",
                min_index, max_index + min_index - 2, max_index - 1);
    if (regenerate)
        for (int32_t i = min_index; i < max_index; i++) dest[i] = i;
    if (!shuffle) return PXL_OK;
    // Swap i with a position drawn from [min_index, min_index + i) for i
    // from the top down (the reference's loop; with min_index 0 this is
    // Sattolo's algorithm, which yields a single cycle).
    for (int32_t i = max_index - 1; i > min_index; i--) {
        const int32_t j = rand() % i + min_index;
        const int32_t t = dest[i];
        dest[i] = dest[j];
        dest[j] = t;
    }
    return PXL_OK;
}
