// Host-side checks of the size limits in csrc/ndbx_limits.h (no device).
// Prints one line per check and exits non-zero on the first failure.
#include <cstdio>
#include <cstdlib>

#include "ndbx_limits.h"

#define CHECK(cond)                                                   \
    do {                                                              \
        if (!(cond)) {                                                \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond); \
            return 1;                                                 \
        }                                                             \
        std::printf("ok %s\n", #cond);                                \
    } while (0)

int main() {
    using namespace ndbx;
    const int64_t safe = mfma_safe_cells();
    // every ordered pair total of `safe` cells fits; one more cell may not
    CHECK(safe * safe <= kMfmaMaxPairs);
    CHECK((safe + 1) * (safe + 1) > kMfmaMaxPairs);
    CHECK(mfma_pairs_fit(0));
    CHECK(mfma_pairs_fit(kMfmaMaxPairs));
    CHECK(!mfma_pairs_fit(kMfmaMaxPairs + 1));
    CHECK(!mfma_pairs_fit(-1));
    // the wrapped int32 sum a large run could reach is refused:
    // ~58.6k cells of ~1024 points (60M points), every pair a candidate
    const int64_t nc = 60000000 / 1024 + 1;
    CHECK(nc > safe);
    CHECK(!mfma_pairs_fit(nc * nc));
    return 0;
}
