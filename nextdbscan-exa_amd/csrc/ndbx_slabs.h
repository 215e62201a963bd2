// Slab plan of a chunked run, as plain host arithmetic (no HIP types), so it
// can be checked on a host without a device.
//
// The points are sorted by their first coordinate x.  Slab k owns the sorted
// positions [own_lo, own_hi) and works on [sub_lo, sub_hi): its owned points
// plus every point whose x is within `reach` of the owned x range.  With
// `reach` above the largest |dx| of any pair the neighbour predicate accepts,
// every neighbour of an owned point is in the slab, so the owned points'
// counts, core-core edges and border attachments are all seen there.  Points
// of equal x are never split between slabs.
#pragma once

#include <cstdint>
#include <vector>

namespace ndbx {

struct Slab {
    int64_t sub_lo, own_lo, own_hi, sub_hi;  // positions in x-sorted order
};

// Greedy plan: each slab owns as many whole groups of equal x as keep its
// working range within `budget` points.  Returns false (and the sorted
// position that cannot be placed in *bad_at) when one group of equal x with
// its halo already exceeds the budget.
inline bool plan_slabs(const float *xs, int64_t n, double reach, int64_t budget, std::vector<Slab> &out,
                       int64_t *bad_at = nullptr) {
    out.clear();
    int64_t lo = 0, hi = 0, a = 0;
    while (a < n) {
        const double xa = (double)xs[a] - reach;
        while (lo < a && (double)xs[lo] < xa) ++lo;  // first position with x >= x[a] - reach
        int64_t e = a, h = hi > a ? hi : a, best_e = -1, best_h = -1;
        while (e < n) {
            int64_t g = e + 1;  // end of the group of equal x starting at e
            while (g < n && xs[g] == xs[e]) ++g;
            const double xb = (double)xs[g - 1] + reach;
            while (h < n && (double)xs[h] <= xb) ++h;  // first position with x > x[g-1] + reach
            if (h - lo > budget) break;
            best_e = g;
            best_h = h;
            e = g;
        }
        if (best_e < 0) {
            if (bad_at) *bad_at = a;
            return false;
        }
        out.push_back(Slab{lo, a, best_e, best_h});
        a = best_e;
        hi = best_h;
    }
    return true;
}

}  // namespace ndbx
