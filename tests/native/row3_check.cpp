// Exhaustive host check of the halo-grid row test used by the neighbour walks
// (row3_in_grid in sph_pairs.hpp): for every row length hx in [2, 512] and every middle
// cell cx in [-2000, hx + 2000] plus the int32 extremes, the test admits exactly the rows
// whose three cells cx - 1 .. cx + 1 map to entries cx .. cx + 2 inside [0, hx).
#include <cstdint>
#include <cstdio>

#include "sph_pairs.hpp"

int main() {
    long checked = 0, bad = 0;
    auto check = [&](int32_t cx, uint32_t hx) {
        const int64_t lo = (int64_t)cx, hi = (int64_t)cx + 2;  // entries of cells cx-1, cx+1
        const bool exact = lo >= 0 && hi < (int64_t)hx;
        const bool got = sph::row3_in_grid(cx, hx);
        ++checked;
        if (got != exact) {
            if (bad < 10) std::printf("mismatch cx=%d hx=%u: %d vs %d\n", cx, hx, got, exact);
            ++bad;
        }
    };
    for (uint32_t hx = 2; hx <= 512; ++hx) {
        for (int32_t cx = -2000; cx <= (int32_t)hx + 2000; ++cx) check(cx, hx);
        check(INT32_MIN, hx);
        check(INT32_MAX, hx);
        check(INT32_MIN + 1, hx);
        check(INT32_MAX - 2, hx);
    }
    std::printf("checked %ld, mismatches %ld\n", checked, bad);
    return bad == 0 ? 0 : 1;
}
