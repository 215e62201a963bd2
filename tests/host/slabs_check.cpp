// Host-side checks of the slab plan of chunked runs (csrc/ndbx_slabs.h)
// against brute-force restatements of its invariants, on random sorted
// inputs with many equal values.  Exits non-zero on the first failure.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "ndbx_slabs.h"

#define CHECK(cond)                                                   \
    do {                                                              \
        if (!(cond)) {                                                \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond); \
            return 1;                                                 \
        }                                                             \
    } while (0)

static int check_plan(const std::vector<float> &xs, double reach, int64_t budget, bool expect_ok) {
    using namespace ndbx;
    const int64_t n = (int64_t)xs.size();
    std::vector<Slab> sl;
    int64_t bad = -1;
    const bool ok = plan_slabs(xs.data(), n, reach, budget, sl, &bad);
    if (!ok) {
        CHECK(!expect_ok);
        CHECK(bad >= 0 && bad < n);
        // the group of equal x at `bad` with its halo does not fit
        int64_t g0 = bad, g1 = bad;
        while (g1 < n && xs[g1] == xs[bad]) ++g1;
        int64_t cnt = 0;
        for (int64_t i = 0; i < n; ++i)
            cnt += (double)xs[i] >= (double)xs[g0] - reach && (double)xs[i] <= (double)xs[g1 - 1] + reach;
        CHECK(cnt > budget);
        return 0;
    }
    CHECK(expect_ok);
    int64_t next = 0;
    for (const Slab &s : sl) {
        CHECK(s.own_lo == next && s.own_hi > s.own_lo);  // owned ranges partition [0, n) in order
        next = s.own_hi;
        CHECK(s.own_lo == 0 || xs[s.own_lo - 1] != xs[s.own_lo]);  // equal x never split
        CHECK(s.sub_hi - s.sub_lo <= budget);
        // exactly the points within reach of the owned x range
        for (int64_t i = 0; i < n; ++i) {
            const bool in = (double)xs[i] >= (double)xs[s.own_lo] - reach &&
                            (double)xs[i] <= (double)xs[s.own_hi - 1] + reach;
            CHECK(in == (i >= s.sub_lo && i < s.sub_hi));
        }
    }
    CHECK(next == n);
    return 0;
}

int main() {
    std::mt19937_64 rng(7);
    int cases = 0;
    for (int t = 0; t < 300; ++t) {
        const int64_t n = 1 + (int64_t)(rng() % 3000);
        std::vector<float> xs(n);
        const int levels = 1 + (int)(rng() % 400);  // few levels: many equal values
        for (auto &x : xs) x = (float)(rng() % levels) * 0.25f + ((rng() % 5 == 0) ? 0.0f : 0.125f);
        std::sort(xs.begin(), xs.end());
        const double reach = 0.05 + (double)(rng() % 100) * 0.01;
        const int64_t budget = 1 + (int64_t)(rng() % (n + 50));
        // feasible iff every group of equal x with its halo fits
        bool feasible = true;
        for (int64_t a = 0; a < n && feasible;) {
            int64_t g = a;
            while (g < n && xs[g] == xs[a]) ++g;
            int64_t cnt = 0;
            for (int64_t i = 0; i < n; ++i)
                cnt += (double)xs[i] >= (double)xs[a] - reach && (double)xs[i] <= (double)xs[g - 1] + reach;
            feasible = cnt <= budget;
            a = g;
        }
        if (check_plan(xs, reach, budget, feasible)) return 1;
        ++cases;
    }
    // a budget that holds everything gives one slab over all points
    std::vector<float> xs(1000);
    for (int i = 0; i < 1000; ++i) xs[i] = (float)i * 0.01f;
    std::vector<ndbx::Slab> sl;
    CHECK(ndbx::plan_slabs(xs.data(), 1000, 0.5, 1000, sl) && sl.size() == 1 && sl[0].sub_lo == 0 && sl[0].sub_hi == 1000);
    CHECK(ndbx::plan_slabs(xs.data(), 0, 0.5, 10, sl) && sl.empty());
    std::printf("ok %d random plans\n", cases);
    return 0;
}
