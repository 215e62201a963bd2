// Size limits of the ndbx pipelines as plain host arithmetic (no HIP types),
// so that they can be checked on a host without a device.
#pragma once

#include <cstdint>

namespace ndbx {

// The MFMA filter keeps its candidate cell pairs as a CSR list with int32
// offsets; a partition of nc cells can have up to nc * nc ordered pairs.
constexpr int64_t kMfmaMaxPairs = INT32_MAX - 64;

// True when `total` candidate cell pairs fit the int32 offsets.
inline bool mfma_pairs_fit(int64_t total) { return total >= 0 && total <= kMfmaMaxPairs; }

// Largest cell count whose every possible pair total fits (nc^2 <= limit);
// below it the per-cell counts need no 64-bit sum.
inline int64_t mfma_safe_cells() {
    int64_t c = 46340;  // floor(sqrt(2^31 - 65))
    while ((c + 1) * (c + 1) <= kMfmaMaxPairs) ++c;
    while (c * c > kMfmaMaxPairs) --c;
    return c;
}

}  // namespace ndbx
