// Host check of the cell key order the neighbor lists follow:
//  * compact_key sorts cells as the reference's key does, libmorton's
//    morton3D_64_encode of c + 2^31 (CompressedNeighbors.cpp:43-56: bit b of
//    x, y, z at key bits 3b, 3b + 1, 3b + 2; x keeps 22 bits, y and z 21);
//  * neighbor_order_word: the 27 codes decode to the 27 offsets one to one,
//    the cell's own code to (0, 0, 0), and the neighbors inside [-2^L, 2^L)
//    taken by ascending code have ascending keys.
// Every cell of levels 1-4, a lattice of cells up to level 9.
#include <algorithm>
#include <array>
#include <cstdio>
#include <vector>

#include "fsa_device.hpp"

static uint64_t ref_key(int32_t x, int32_t y, int32_t z) {
    const uint32_t u[3] = {(uint32_t)((int64_t)x + 2147483648LL), (uint32_t)((int64_t)y + 2147483648LL),
                           (uint32_t)((int64_t)z + 2147483648LL)};
    uint64_t k = 0;
    for (int b = 0; b < 22; b++) {
        k |= (uint64_t)((u[0] >> b) & 1u) << (3 * b);
        if (b < 21) {
            k |= (uint64_t)((u[1] >> b) & 1u) << (3 * b + 1);
            k |= (uint64_t)((u[2] >> b) & 1u) << (3 * b + 2);
        }
    }
    return k;
}

int main() {
    long bad_key = 0, bad_order = 0, cells = 0;
    for (uint32_t L = 1; L <= 9; L++) {
        const int lo = -(1 << L), hi = 1 << L;
        std::vector<int> vals;
        if (L <= 4) {
            for (int v = lo; v < hi; v++) vals.push_back(v);
        } else {
            for (int v = lo; v < hi; v += (L > 7 ? 37 : 13)) vals.push_back(v);
            for (int v : {lo + 1, -2, -1, 0, 1, 2, hi - 2, hi - 1, 1 << (L - 1), -(1 << (L - 1)), (1 << (L - 1)) - 1})
                vals.push_back(v);
            std::sort(vals.begin(), vals.end());
            vals.erase(std::unique(vals.begin(), vals.end()), vals.end());
        }
        // compact_key order against the reference key over every cell of the level's sample
        std::vector<std::array<int, 3>> all;
        for (int x : vals)
            for (int y : vals)
                for (int z : vals) all.push_back({x, y, z});
        std::sort(all.begin(), all.end(), [&](const std::array<int, 3>& a, const std::array<int, 3>& b) {
            return fsa::compact_key(a[0], a[1], a[2], L) < fsa::compact_key(b[0], b[1], b[2], L);
        });
        for (size_t i = 1; i < all.size(); i++)
            if (!(ref_key(all[i - 1][0], all[i - 1][1], all[i - 1][2]) < ref_key(all[i][0], all[i][1], all[i][2])))
                bad_key++;
        for (const auto& c : all) {
            const fsa::NeighborOrder o = fsa::unpack_order(fsa::neighbor_order_word(c[0], c[1], c[2], L));
            int seen[27] = {0}, valid = 0;
            uint64_t prev = 0;
            bool first = true;
            for (uint32_t code = 0; code < 64; code++) {
                int32_t d[3];
                if (!(fsa::neighbor_offset(o, code, 0, d[0]) & fsa::neighbor_offset(o, code, 1, d[1]) &
                      fsa::neighbor_offset(o, code, 2, d[2])))
                    continue;
                valid++;
                const int idx = (d[0] + 1) * 9 + (d[1] + 1) * 3 + (d[2] + 1);
                seen[idx]++;
                if (code == o.self && idx != 13) bad_order++;
                const int x = c[0] + d[0], y = c[1] + d[1], z = c[2] + d[2];
                if (x < lo || x >= hi || y < lo || y >= hi || z < lo || z >= hi) continue;
                const uint64_t k = ref_key(x, y, z);
                if (!first && !(prev < k)) bad_order++;
                prev = k;
                first = false;
            }
            if (valid != 27) bad_order++;
            for (int i = 0; i < 27; i++)
                if (seen[i] != 1) bad_order++;
            cells++;
        }
    }
    printf("cells %ld bad %ld (key order %ld, neighbor order %ld)\n", cells, bad_key + bad_order, bad_key, bad_order);
    return bad_key + bad_order != 0;
}
