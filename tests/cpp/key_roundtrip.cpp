// Host check of the compact Morton key and its inverse (cell_of_key), as the
// neighbor-list kernel decodes the cell of a run from its key.
#include "fsa_device.hpp"
#include <cstdio>
int main() {
    long bad = 0, n = 0;
    for (uint32_t L = 1; L <= 9; L++) {
        const int lim = 1 << L;
        for (int x = -lim; x < lim; x += (L > 6 ? 7 : 1))
            for (int y = -lim; y < lim; y += (L > 6 ? 5 : 1))
                for (int z = -lim; z < lim; z += (L > 6 ? 3 : 1)) {
                    int a, b, c;
                    fsa::cell_of_key(fsa::compact_key(x, y, z, L), L, a, b, c);
                    bad += (a != x || b != y || c != z);
                    n++;
                }
    }
    printf("checked %ld bad %ld\n", n, bad);
    return bad != 0;
}
