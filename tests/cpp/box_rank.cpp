// Host check of box_rank: for boxes of cells (straddling zero or not, up to
// level 6 and a few at level 8), sorting the box's cells by compact_key must
// put every cell at the position box_rank gives it.
#include <algorithm>
#include <array>
#include <cstdio>
#include <random>
#include <vector>

#include "fsa_device.hpp"

int main() {
    std::mt19937 rng(12345);
    long boxes = 0, cells = 0, bad = 0;
    auto check = [&](const int32_t lo[3], const int32_t hi[3], uint32_t L) {
        std::vector<std::array<int32_t, 3>> all;
        for (int32_t x = lo[0]; x <= hi[0]; x++)
            for (int32_t y = lo[1]; y <= hi[1]; y++)
                for (int32_t z = lo[2]; z <= hi[2]; z++) all.push_back({x, y, z});
        std::sort(all.begin(), all.end(), [&](const std::array<int32_t, 3>& a, const std::array<int32_t, 3>& b) {
            return fsa::compact_key(a[0], a[1], a[2], L) < fsa::compact_key(b[0], b[1], b[2], L);
        });
        for (size_t i = 0; i < all.size(); i++)
            if (fsa::box_rank(all[i][0], all[i][1], all[i][2], lo, hi, L) != (uint32_t)i) bad++;
        cells += (long)all.size();
        boxes++;
    };
    for (uint32_t L = 1; L <= 6; L++) {
        const int32_t lim = 1 << L;
        std::uniform_int_distribution<int32_t> d(-lim, lim - 1);
        for (int t = 0; t < 60; t++) {
            int32_t lo[3], hi[3];
            for (int a = 0; a < 3; a++) {
                int32_t p = d(rng), q = d(rng);
                if (L > 4) q = p + (q & 15);  // keep the box small at the larger levels
                if (q >= lim) q = lim - 1;
                lo[a] = std::min(p, q);
                hi[a] = std::max(p, q);
            }
            check(lo, hi, L);
        }
        const int32_t lo[3] = {-lim, -lim, -lim}, hi[3] = {lim - 1, lim - 1, lim - 1};  // the whole level
        if (L <= 4) check(lo, hi, L);
    }
    // the breaking-dam shape: a box from -2 to 151 x 76 x 76 at level 8, sampled on a few thin slabs
    {
        const int32_t lo[3] = {-2, -2, -2}, hi[3] = {151, 76, 76};
        std::vector<std::array<int32_t, 3>> all;
        for (int32_t x = lo[0]; x <= hi[0]; x++)
            for (int32_t y = lo[1]; y <= hi[1]; y++)
                for (int32_t z = lo[2]; z <= hi[2]; z++) all.push_back({x, y, z});
        std::sort(all.begin(), all.end(), [&](const std::array<int32_t, 3>& a, const std::array<int32_t, 3>& b) {
            return fsa::compact_key(a[0], a[1], a[2], 8) < fsa::compact_key(b[0], b[1], b[2], 8);
        });
        for (size_t i = 0; i < all.size(); i += 7)
            if (fsa::box_rank(all[i][0], all[i][1], all[i][2], lo, hi, 8) != (uint32_t)i) bad++;
        cells += (long)all.size() / 7;
        boxes++;
    }
    printf("boxes %ld cells %ld bad %ld \n", boxes, cells, bad);
    return bad != 0;
}
