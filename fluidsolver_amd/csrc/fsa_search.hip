// fsa_search.hip — neighborhood search on gfx950.
//
// Replaces CompressedNeighborhoodSearch::find_neighbors
// (CompressedNeighbors.cpp:168-240):
//   1. k_keys_hist    cell = floor(x / r), compact Morton key      (:174-189, 37-56)
//   2. k_radix_*      stable LSD radix sort of (key, index)        (:191-201)
//   3. k_permute      gather every component into sorted order     (ParticleCollection::swap)
//   4. k_fluid_scan   fluid counts per block to offsets: a fluid particle's
//                     rank is its block's offset plus its in-block rank
//                     (k_permute); k_lists writes fluid rank -> combined index
//   5. (in k_permute) cell -> [start, end) of its run              (:203-225)
//   6. k_lists        27-cell scan in ascending key order, distance
//                     test dot(x_j - x_i) <= r^2, ascending lists  (:536-595)
// All kernels take a scene table and a block map, so one launch serves any
// number of scenes (batching) and a single scene alike.
#include <type_traits>

#include "fsa_device.hpp"
#include "fsa_kernels.hpp"

namespace fsa {

// ---- wave/block helpers ---------------------------------------------------
__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = __lane_id();
    return l == 0 ? 0ull : ((~0ull) >> (64 - l));
}

// exclusive scan of one value per thread over a block of kBlock (= 256) threads
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds_waves, uint32_t& total) {
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x / 64;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) lds_waves[wave] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    const uint32_t nw = blockDim.x / 64;
    for (uint32_t w = 0; w < nw; w++) {
        uint32_t c = lds_waves[w];
        if (w < wave) off += c;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return off + x - v;
}

// ---- 1. keys (in k_keys_hist, with the first radix histogram) ---------------------

// ---- 2a. counting sort over cell ranks ---------------------------------------
// With rank keys every cell of the grid is a bucket, and the collection comes
// in the order of the last search, so a cell's particles mostly arrive as runs
// of consecutive rows.  k_rank_count gives each particle a slot in its bucket
// (one atomic per run of equal ranks within a wave, slots in row order within
// the run), k_bucket_sums / k_bucket_offsets scan the bucket counts, k_place
// writes each particle at its bucket's start plus its slot, and k_fixup puts
// every bucket's rows in ascending order, which makes the result the stable
// sort.  Buckets above 64 rows are not ordered here: the scene's step is
// retried with the radix sort (kStatusSortRetry).
constexpr uint32_t kFixupFast = 12;  // bucket rows ordered in a lane's registers (up to 64 by the wave)

__global__ __launch_bounds__(kBlock) void k_rank_count(const SceneDev* __restrict__ scenes,
                                                       const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t i = br.block * kBlock + threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t n = S.n;
    const bool valid = i < n;
    uint32_t r = 0xffffffffu;
    if (valid) {
        const float4 p = gload_ptr(S.pos[S.state->cur])[i];
        const float radius = S.search_radius;
        const int32_t x = cell_of(p.x, radius) - S.cmin[0], y = cell_of(p.y, radius) - S.cmin[1],
                      z = cell_of(p.z, radius) - S.cmin[2];
        if (x < 0 || y < 0 || z < 0 || x >= (int32_t)S.cdim[0] || y >= (int32_t)S.cdim[1] || z >= (int32_t)S.cdim[2]) {
            atomicOr(&S.state->status, kStatusGrid);  // the scene fails; any bucket keeps the sort a permutation
            r = 0;
        } else {
            r = S.ranks[(uint32_t)x + S.cdim[0] * ((uint32_t)y + S.cdim[1] * (uint32_t)z)];
        }
    }
    // runs of equal ranks in consecutive lanes (the valid lanes are a prefix of the wave)
    const uint32_t rp = (uint32_t)__builtin_amdgcn_update_dpp((int)0xffffffffu, (int)r, 0x138, 0xf, 0xf, false);
    const bool head = valid && (lane == 0 || rp != r);
    const uint64_t heads = __ballot(head);
    const uint64_t nvalid = __ballot(valid);
    uint32_t base = 0;
    if (head) {
        const uint64_t later = heads & ~((2ull << lane) - 1ull);  // heads above this lane (none above 63)
        const uint32_t next = later ? (uint32_t)__ffsll((unsigned long long)later) - 1u
                                    : (uint32_t)__popcll(nvalid);
        base = atomicAdd(&S.bcnt[r], next - lane);
    }
    const uint64_t upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
    const uint32_t h = upto ? 63u - (uint32_t)__clzll(upto) : 0u;
    const uint32_t slot = (uint32_t)__shfl((int)base, (int)h, 64) + (lane - h);
    if (valid) {
        S.keys[0][i] = r;
        S.vals[0][i] = slot;
    }
}

// A scene's rank range [0, cdim[3]) in parts of about kCountPart ranks, at most
// kCountBlocks of them (the blocks of the map above the part count exit)
constexpr uint32_t kCountPart = 1024;
__device__ __forceinline__ uint32_t count_parts(const SceneDev& S) {
    return max(1u, min(kCountBlocks, (S.cdim[3] + kCountPart - 1) / kCountPart));
}
__device__ __forceinline__ void count_part(const SceneDev& S, uint32_t parts, uint32_t b, uint32_t& lo, uint32_t& hi) {
    const uint32_t c = S.cdim[3];
    lo = (uint32_t)((uint64_t)c * b / parts);
    hi = (uint32_t)((uint64_t)c * (b + 1) / parts);
}

__global__ __launch_bounds__(kBlock) void k_bucket_sums(const SceneDev* __restrict__ scenes,
                                                        const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    __shared__ uint32_t lw[kBlock / 64];
    const uint32_t parts = count_parts(S);
    if (br.block >= parts) return;
    uint32_t lo, hi;
    count_part(S, parts, br.block, lo, hi);
    const auto bcnt = gload_ptr(S.bcnt);
    uint32_t sum = 0;
    for (uint32_t c = lo + threadIdx.x; c < hi; c += kBlock) sum += bcnt[c];
    uint32_t tot;
    (void)block_exclusive_scan(sum, lw, tot);
    if (threadIdx.x == 0) S.bsum[br.block] = tot;
}

__global__ __launch_bounds__(kBlock) void k_bucket_offsets(const SceneDev* __restrict__ scenes,
                                                           const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    __shared__ uint32_t lw[kBlock / 64];
    const uint32_t parts = count_parts(S);
    if (br.block >= parts) return;
    uint32_t lo, hi;
    count_part(S, parts, br.block, lo, hi);
    // rows of the parts before this one
    uint32_t before = 0;
    for (uint32_t q = threadIdx.x; q < br.block; q += kBlock) before += S.bsum[q];
    uint32_t carry;
    (void)block_exclusive_scan(before, lw, carry);
    uint32_t* __restrict__ bcnt = S.bcnt;
    uint32_t* __restrict__ boff = S.boff;
    for (uint32_t c0 = lo; c0 < hi; c0 += kBlock) {
        const uint32_t c = c0 + threadIdx.x;
        const uint32_t v = c < hi ? bcnt[c] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, lw, tot);
        if (c < hi) {
            boff[c] = carry + ex;
            bcnt[c] = 0u;  // zero for the next search
        }
        carry += tot;
    }
    if (br.block == parts - 1 && threadIdx.x == 0) boff[S.cdim[3]] = carry;
}

__global__ __launch_bounds__(kBlock) void k_place(const SceneDev* __restrict__ scenes,
                                                  const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t i = br.block * kBlock + threadIdx.x;
    if (i >= S.n) return;
    const uint32_t r = S.keys[0][i], slot = S.vals[0][i];
    const uint32_t p = S.boff[r] + slot;
    S.keys[1][p] = r;
    S.vals[1][p] = i;
}

// The first row of each bucket orders the bucket's rows (source indices)
// ascending: up to kFixupFast rows in its lane's registers (a sorting network
// of compare-exchanges), up to 64 rows by the whole wave (a bitonic sort across
// the lanes, one bucket at a time), more rows not at all (retry).
__global__ __launch_bounds__(kBlock) void k_fixup(const SceneDev* __restrict__ scenes,
                                                  const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t p = br.block * kBlock + threadIdx.x;
    const uint32_t lane = __lane_id();
    uint32_t k = 0;
    if (p < S.n) {
        const uint32_t* __restrict__ keys = S.keys[1];
        const uint32_t r = keys[p];
        if (p == 0 || keys[p - 1] != r) k = S.boff[r + 1] - S.boff[r];
    }
    uint32_t* __restrict__ vals = S.vals[1];
    const bool fast = k >= 2 && k <= kFixupFast;
    uint32_t v[kFixupFast];
    bool sorted = true;
    if (fast) {
#pragma unroll
        for (uint32_t q = 0; q < kFixupFast; q++) v[q] = q < k ? vals[p + q] : 0xffffffffu;
#pragma unroll
        for (uint32_t q = 1; q < kFixupFast; q++) sorted = sorted && v[q - 1] <= v[q];
    }
    // most buckets arrive in order (one run of rows, or runs whose atomics came in
    // row order): the network runs only where some bucket of the wave did not
    if (__any(fast && !sorted) && fast && !sorted) {
        // insertion network; the padding (all ones) stays at the end
#pragma unroll
        for (uint32_t q = 1; q < kFixupFast; q++) {
#pragma unroll
            for (uint32_t t = q; t > 0; t--) {
                const uint32_t a = v[t - 1], b = v[t];
                v[t - 1] = min(a, b);
                v[t] = max(a, b);
            }
        }
#pragma unroll
        for (uint32_t q = 0; q < kFixupFast; q++)
            if (q < k) vals[p + q] = v[q];
    }
    if (k > 64u) atomicOr(&S.state->status, kStatusSortRetry);
    // buckets of kFixupFast + 1 .. 64 rows, one at a time across the wave
    for (uint64_t m = __ballot(k > kFixupFast && k <= 64u); m; m &= m - 1) {
        const uint32_t src = (uint32_t)__ffsll((unsigned long long)m) - 1u;
        const uint32_t q0 = (uint32_t)__builtin_amdgcn_readlane((int)p, (int)src);
        const uint32_t kk = (uint32_t)__builtin_amdgcn_readlane((int)k, (int)src);
        uint32_t x = lane < kk ? vals[q0 + lane] : 0xffffffffu;
#pragma unroll
        for (uint32_t size = 2; size <= 64; size <<= 1) {
#pragma unroll
            for (uint32_t stride = size / 2; stride > 0; stride >>= 1) {
                const uint32_t y = (uint32_t)__shfl_xor((int)x, (int)stride, 64);
                const bool up = (lane & size) == 0;     // this part sorts ascending
                const bool low = (lane & stride) == 0;  // this lane keeps the smaller value when ascending
                x = (up == low) ? min(x, y) : max(x, y);
            }
        }
        if (lane < kk) vals[q0 + lane] = x;
    }
}

// ---- 2. radix sort -----------------------------------------------------------
static_assert(kRadix % kSortThreads == 0, "each sort thread owns a whole number of digits");
// peers of this lane among `valid` lanes holding the same Bits-bit digit
template <int Bits>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) {
    uint64_t peers = valid;
#pragma unroll
    for (int b = 0; b < Bits; b++) {
        uint64_t bal = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    return peers;
}

template <int Bits>
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const SceneDev* __restrict__ scenes,
                                                             const BlockRef* __restrict__ map, uint32_t shift,
                                                             uint32_t src) {
    constexpr uint32_t kDigits = 1u << Bits;
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t tile = br.block;
    __shared__ uint32_t h[kDigits];
    for (uint32_t d = threadIdx.x; d < kDigits; d += kSortThreads) h[d] = 0;
    const uint32_t n = S.n, ntiles = S.ntiles;
    const auto keys = gload_ptr(S.keys[src]);
    FSA_GLOBAL(uint32_t)* __restrict__ hist = (FSA_GLOBAL(uint32_t)*)S.hist;
    const uint32_t base = tile * kSortTile;
    // every item's key is loaded before the first count (past the end: the last key, not counted)
    uint32_t kk[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) kk[r] = keys[min(base + r * kSortThreads + threadIdx.x, n - 1u)];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        const bool valid = i < n;
        const uint32_t d = (kk[r] >> shift) & (kDigits - 1);
        const uint64_t peers = match_digit<Bits>(d, __ballot(valid));
        if (valid && (peers & lanemask_lt()) == 0) atomicAdd(&h[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < kDigits; d += kSortThreads) hist[d * ntiles + tile] = h[d];
}

// wide single-block scans (one block per scene or per (scene, digit))
constexpr uint32_t kScanThreads = 1024;

// k_keys and the first pass's k_radix_hist in one pass over the sort tiles:
// each tile's keys are computed, stored, and counted by their lowest digit.
// A particle's key is its cell's compact key, or, where the scene keeps a rank
// table (rank_keys), the cell's rank in compact-key order among the cells of
// the step's grid: the same order in fewer bits.
template <int Bits>
__global__ __launch_bounds__(kSortThreads) void k_keys_hist(const SceneDev* __restrict__ scenes,
                                                            const BlockRef* __restrict__ map) {
    constexpr uint32_t kDigits = 1u << Bits;
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t tile = br.block;
    __shared__ uint32_t h[kDigits];
    for (uint32_t d = threadIdx.x; d < kDigits; d += kSortThreads) h[d] = 0;
    // the scene's fields are read once (the stores below could otherwise alias
    // the scene table), and every item's position is loaded before the first store
    const uint32_t n = S.n, ntiles = S.ntiles, levels = S.levels;
    const bool rank_keys = S.rank_keys != 0u;
    const auto ranks = gload_ptr(S.ranks);
    const int32_t cmin0 = S.cmin[0], cmin1 = S.cmin[1], cmin2 = S.cmin[2];
    const uint32_t cdim0 = S.cdim[0], cdim1 = S.cdim[1], cdim2 = S.cdim[2];
    SceneState* const state = S.state;
    const float radius = S.search_radius;
    const auto pos = gload_ptr(S.pos[S.state->cur]);
    FSA_GLOBAL(uint32_t)* __restrict__ okeys = (FSA_GLOBAL(uint32_t)*)S.keys[0];
    FSA_GLOBAL(uint32_t)* __restrict__ ovals = (FSA_GLOBAL(uint32_t)*)S.vals[0];
    FSA_GLOBAL(uint32_t)* __restrict__ hist = (FSA_GLOBAL(uint32_t)*)S.hist;
    const uint32_t base = tile * kSortTile;
    // past the end: the last particle's position, neither stored nor counted
    float4 p[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) p[r] = pos[min(base + r * kSortThreads + threadIdx.x, n - 1u)];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        const bool valid = i < n;
        uint32_t key = 0;
        if (valid) {
            const int32_t cx = cell_of(p[r].x, radius), cy = cell_of(p[r].y, radius), cz = cell_of(p[r].z, radius);
            if (rank_keys) {
                const int32_t x = cx - cmin0, y = cy - cmin1, z = cz - cmin2;
                if (x < 0 || y < 0 || z < 0 || x >= (int32_t)cdim0 || y >= (int32_t)cdim1 || z >= (int32_t)cdim2)
                    atomicOr(&state->status, kStatusGrid);
                else
                    key = ranks[(uint32_t)x + cdim0 * ((uint32_t)y + cdim1 * (uint32_t)z)];
            } else {
                key = compact_key(cx, cy, cz, levels);
            }
            okeys[i] = key;
            ovals[i] = i;
        }
        const uint32_t d = key & (kDigits - 1);
        const uint64_t peers = match_digit<Bits>(d, __ballot(valid));
        if (valid && (peers & lanemask_lt()) == 0) atomicAdd(&h[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < kDigits; d += kSortThreads) hist[d * ntiles + tile] = h[d];
}

// The rank table: each cell of the step's grid gets its rank in compact-key
// order among the grid's cells (box_rank).
__global__ __launch_bounds__(kBlock) void k_rank_table(uint32_t* __restrict__ ranks, int32_t cmin0, int32_t cmin1,
                                                       int32_t cmin2, uint32_t cdim0, uint32_t cdim1, uint32_t cdim2,
                                                       uint32_t levels) {
    const uint32_t ncells = cdim0 * cdim1 * cdim2;
    const uint32_t lin = blockIdx.x * kBlock + threadIdx.x;
    if (lin >= ncells) return;
    const uint32_t x = lin % cdim0, y = (lin / cdim0) % cdim1, z = lin / (cdim0 * cdim1);
    const int32_t lo[3] = {cmin0, cmin1, cmin2};
    const int32_t hi[3] = {cmin0 + (int32_t)cdim0 - 1, cmin1 + (int32_t)cdim1 - 1, cmin2 + (int32_t)cdim2 - 1};
    ranks[lin] = box_rank(cmin0 + (int32_t)x, cmin1 + (int32_t)y, cmin2 + (int32_t)z, lo, hi, levels);
}

// one block per (scene, digit d < kRadix): exclusive scan of the tile counts
// of digits d, d + kRadix, ... below 2^Bits, side by side (their loads in flight together)
template <int Bits>
__global__ __launch_bounds__(kBlock) void k_radix_scan(const SceneDev* __restrict__ scenes,
                                                       const BlockRef* __restrict__ map) {
    constexpr uint32_t kDigits = 1u << Bits;
    constexpr uint32_t kRows = kDigits / (uint32_t)kRadix;
    static_assert(kDigits % (uint32_t)kRadix == 0, "whole digit rows per block");
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    __shared__ uint32_t lw[kRows][kBlock / 64];
    const uint32_t ntiles = S.ntiles;
    uint32_t* row[kRows];
    uint32_t carry[kRows];
#pragma unroll
    for (uint32_t q = 0; q < kRows; q++) {
        row[q] = S.hist + (br.block + q * (uint32_t)kRadix) * ntiles;
        carry[q] = 0;
    }
    for (uint32_t base = 0; base < ntiles; base += kBlock) {
        const uint32_t idx = base + threadIdx.x;
        uint32_t v[kRows];
#pragma unroll
        for (uint32_t q = 0; q < kRows; q++) v[q] = idx < ntiles ? row[q][idx] : 0u;
#pragma unroll
        for (uint32_t q = 0; q < kRows; q++) {
            uint32_t tot;
            const uint32_t ex = block_exclusive_scan(v[q], lw[q], tot);
            if (idx < ntiles) row[q][idx] = carry[q] + ex;
            carry[q] += tot;
        }
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (uint32_t q = 0; q < kRows; q++) S.hist[kDigits * ntiles + br.block + q * (uint32_t)kRadix] = carry[q];
    }
}

template <int Bits>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const SceneDev* __restrict__ scenes,
                                                                const BlockRef* __restrict__ map, uint32_t shift,
                                                                uint32_t src) {
    constexpr uint32_t kDigits = 1u << Bits;
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t tile = br.block;
    const uint32_t t = threadIdx.x;
    const uint32_t wave = t / 64;
    __shared__ uint32_t digit_base[kDigits];
    __shared__ uint32_t tile_start[kDigits];
    __shared__ uint32_t wave_cnt[kSortThreads / 64][kDigits];
    __shared__ uint32_t stage_k[kSortTile];
    __shared__ uint32_t stage_v[kSortTile];
    __shared__ uint32_t lw[kSortThreads / 64];

    // the scene's fields once, and the tile's items and this tile's digit counts
    // loaded before anything waits on them
    const uint32_t ntiles = S.ntiles;
    const auto hist = gload_ptr(S.hist);
    const auto keys = gload_ptr(S.keys[src]);
    const auto vals = gload_ptr(S.vals[src]);
    FSA_GLOBAL(uint32_t)* __restrict__ okeys = (FSA_GLOBAL(uint32_t)*)S.keys[src ^ 1u];
    FSA_GLOBAL(uint32_t)* __restrict__ ovals = (FSA_GLOBAL(uint32_t)*)S.vals[src ^ 1u];
    const uint32_t base = tile * kSortTile;
    const uint32_t tile_n = min((uint32_t)kSortTile, S.n - base);
    uint32_t k[kSortItems], v[kSortItems], rank[kSortItems];
    // wave w owns items [w*512, (w+1)*512) of the tile, in rounds of 64 consecutive
    // items (past the tile's end: its last item, which takes no part)
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t li = min(wave * (kSortTile / (kSortThreads / 64)) + r * 64 + __lane_id(), tile_n - 1u);
        k[r] = keys[base + li];
        v[r] = vals[base + li];
    }

    // global base of each digit for this tile; thread t owns digits t * kDigitsPerThread + q
    constexpr uint32_t kDigitsPerThread = kDigits / kSortThreads;
    {
        uint32_t tot_d[kDigitsPerThread], at_d[kDigitsPerThread], local = 0;
#pragma unroll
        for (uint32_t q = 0; q < kDigitsPerThread; q++) {
            tot_d[q] = hist[kDigits * ntiles + t * kDigitsPerThread + q];
            at_d[q] = hist[(t * kDigitsPerThread + q) * ntiles + tile];
        }
#pragma unroll
        for (uint32_t q = 0; q < kDigitsPerThread; q++) local += tot_d[q];
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(local, lw, tot);
#pragma unroll
        for (uint32_t q = 0; q < kDigitsPerThread; q++) {
            const uint32_t d = t * kDigitsPerThread + q;
            digit_base[d] = ex + at_d[q];
            ex += tot_d[q];
        }
    }
#pragma unroll
    for (int w = 0; w < kSortThreads / 64; w++)
#pragma unroll
        for (uint32_t q = 0; q < kDigitsPerThread; q++) wave_cnt[w][t * kDigitsPerThread + q] = 0;
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t li = wave * (kSortTile / (kSortThreads / 64)) + r * 64 + __lane_id();
        const bool valid = li < tile_n;
        const uint32_t d = (k[r] >> shift) & (kDigits - 1);
        const uint64_t peers = match_digit<Bits>(d, __ballot(valid));
        uint32_t before = valid ? wave_cnt[wave][d] : 0u;
        rank[r] = before + (uint32_t)__popcll(peers & lanemask_lt());
        if (valid && (peers & lanemask_lt()) == 0) wave_cnt[wave][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // per digit: exclusive prefix over waves, then exclusive scan over digits
    {
        uint32_t runs[kDigitsPerThread], local = 0;
#pragma unroll
        for (uint32_t q = 0; q < kDigitsPerThread; q++) {
            const uint32_t d = t * kDigitsPerThread + q;
            uint32_t run = 0;
#pragma unroll
            for (int w = 0; w < kSortThreads / 64; w++) {
                const uint32_t c = wave_cnt[w][d];
                wave_cnt[w][d] = run;
                run += c;
            }
            runs[q] = run;
            local += run;
        }
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(local, lw, tot);
#pragma unroll
        for (uint32_t q = 0; q < kDigitsPerThread; q++) {
            tile_start[t * kDigitsPerThread + q] = ex;
            ex += runs[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t li = wave * (kSortTile / (kSortThreads / 64)) + r * 64 + __lane_id();
        if (li < tile_n) {
            const uint32_t d = (k[r] >> shift) & (kDigits - 1);
            const uint32_t pos = tile_start[d] + wave_cnt[wave][d] + rank[r];
            stage_k[pos] = k[r];
            stage_v[pos] = v[r];
        }
    }
    __syncthreads();
    for (uint32_t idx = t; idx < tile_n; idx += kSortThreads) {
        const uint32_t key = stage_k[idx];
        const uint32_t d = (key >> shift) & (kDigits - 1);
        const uint32_t g = digit_base[d] + (idx - tile_start[d]);
        okeys[g] = key;
        ovals[g] = stage_v[idx];
    }
}

// ---- 3. permute ------------------------------------------------------------
// CellTable: the keys are this search's sorted cell keys, and the cell table
// (CompressedNeighbors.cpp:203-225: the run of each cell in sorted order) is
// written from the permuted positions on the way, with each run's grid
// coordinates and neighbor order (neighbor_order_word) for the list kernel.
template <bool CellTable>
__global__ __launch_bounds__(kBlock) void k_permute(const SceneDev* __restrict__ scenes,
                                                    const BlockRef* __restrict__ map, uint32_t sbuf) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (sort_retry(S)) return;
    const uint32_t i = br.block * kBlock + threadIdx.x;
    const uint32_t c = S.state->cur;
    const uint32_t o = c ^ 1u;
    const uint32_t n = S.n;
    const bool ext = (S.flags & kFlagExternal) != 0;
    bool fluid = false;
    if (i < n) {
        // every gather is issued before the first store (the stores could
        // otherwise alias the scene table and the source arrays)
        const uint32_t s = gload_ptr(S.vals[sbuf])[i];
        const auto keys = gload_ptr(S.keys[sbuf]);
        const float4 p = gload_ptr(S.pos[c])[s];
        const float4 vel = gload_ptr(S.vel[c])[s];
        const float4 acc = gload_ptr(S.acc[c])[s];
        const float4 ex = ext ? gload_ptr(S.ext[c])[s] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float pres = gload_ptr(S.pres[c])[s];
        const float err = gload_ptr(S.err[c])[s];
        const uint32_t idt = gload_ptr(S.idt[c])[s];
        uint32_t key = 0, key_prev = 0, key_next = 0;
        if (CellTable) {
            key = keys[i];
            key_prev = keys[i == 0 ? 0u : i - 1u];
            key_next = keys[i + 1 == n ? i : i + 1u];
        }
        const bool first = CellTable && (i == 0 || key_prev != key);
        const bool last = CellTable && (i + 1 == n || key_next != key);
        // the cell table's fields, read before the stores too
        const float radius = S.search_radius;
        const int32_t cmin0 = S.cmin[0], cmin1 = S.cmin[1], cmin2 = S.cmin[2];
        const uint32_t cdim0 = S.cdim[0], cdim1 = S.cdim[1], cdim2 = S.cdim[2];
        const uint32_t stamp = S.stamp, levels = S.levels;
        uint4* __restrict__ cells = S.cells;
        uint32_t* __restrict__ run_grid = S.keys[sbuf ^ 1u];
        uint32_t* __restrict__ run_order = S.vals[sbuf ^ 1u];
        SceneState* const state = S.state;
        FSA_GLOBAL(fsa_f4)* __restrict__ opos = (FSA_GLOBAL(fsa_f4)*)S.pos[o];
        FSA_GLOBAL(fsa_f4)* __restrict__ ovel = (FSA_GLOBAL(fsa_f4)*)S.vel[o];
        FSA_GLOBAL(fsa_f4)* __restrict__ oacc = (FSA_GLOBAL(fsa_f4)*)S.acc[o];
        FSA_GLOBAL(fsa_f4)* __restrict__ oext = (FSA_GLOBAL(fsa_f4)*)S.ext[o];
        FSA_GLOBAL(float)* __restrict__ opres = (FSA_GLOBAL(float)*)S.pres[o];
        FSA_GLOBAL(float)* __restrict__ oerr = (FSA_GLOBAL(float)*)S.err[o];
        FSA_GLOBAL(uint32_t)* __restrict__ oidt = (FSA_GLOBAL(uint32_t)*)S.idt[o];
        opos[i] = fsa_f4{p.x, p.y, p.z, p.w};
        ovel[i] = fsa_f4{vel.x, vel.y, vel.z, vel.w};
        oacc[i] = fsa_f4{acc.x, acc.y, acc.z, acc.w};
        if (ext) oext[i] = fsa_f4{ex.x, ex.y, ex.z, ex.w};
        opres[i] = pres;
        oerr[i] = err;
        oidt[i] = idt;
        fluid = (idt >> kTypeShift) == 0u;
        if (first || last) {
            const int32_t cx = cell_of(p.x, radius), cy = cell_of(p.y, radius), cz = cell_of(p.z, radius);
            // dense index in the step's cell grid (cell_linear)
            const int32_t x = cx - cmin0, y = cy - cmin1, z = cz - cmin2;
            if (x < 0 || y < 0 || z < 0 || x >= (int32_t)cdim0 || y >= (int32_t)cdim1 || z >= (int32_t)cdim2) {
                atomicOr(&state->status, kStatusGrid);
            } else {
                const uint32_t lin = (uint32_t)x + cdim0 * ((uint32_t)y + cdim1 * (uint32_t)z);
                if (first) {
                    cells[lin].x = i;
                    cells[lin].z = stamp;
                    // for the list kernel, at the run's first position of the sort's
                    // idle buffers: the cell's grid coordinates (10 bits each) and the
                    // key order of its neighbors
                    run_grid[i] = (uint32_t)x | (uint32_t)y << 10 | (uint32_t)z << 20;
                    run_order[i] = neighbor_order_word(cx, cy, cz, levels);
                }
                if (last) cells[lin].y = i + 1;
            }
        }
    }
    if (CellTable) {
        // each fluid particle's rank within its block: k_lists adds the block's
        // scanned count for the fluid rank and writes fluid_idx itself
        __shared__ uint32_t lw[kBlock / 64];
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(fluid ? 1u : 0u, lw, tot);
        if (fluid) S.rank_of[i] = ex;
        if (threadIdx.x == 0) S.block_count[br.block] = tot;
    } else {
        const int cnt = __syncthreads_count(fluid);
        if (threadIdx.x == 0) S.block_count[br.block] = (uint32_t)cnt;
    }
}

// one wide block per scene: exclusive scan of the per-block fluid counts, each
// thread over a contiguous run of counts

__global__ __launch_bounds__(kScanThreads) void k_fluid_scan(const SceneDev* __restrict__ scenes,
                                                             const BlockRef* __restrict__ map) {
    const SceneDev& S = scenes[map[blockIdx.x].scene];
    if (sort_retry(S)) return;  // the permute skipped this scene: the buffers stay as they are
    __shared__ uint32_t lw[kScanThreads / 64];
    const uint32_t n = S.nblocks;
    const uint32_t per = (n + kScanThreads - 1) / kScanThreads;
    const uint32_t b0 = min(n, threadIdx.x * per), b1 = min(n, b0 + per);
    uint32_t sum = 0;
    for (uint32_t b = b0; b < b1; b++) sum += S.block_count[b];
    uint32_t tot;
    uint32_t run = block_exclusive_scan(sum, lw, tot);
    for (uint32_t b = b0; b < b1; b++) {
        const uint32_t v = S.block_count[b];
        S.block_count[b] = run;
        run += v;
    }
    // the permuted buffer becomes current (k_permute has finished reading [cur])
    if (threadIdx.x == 0) S.state->cur ^= 1u;
}

__global__ __launch_bounds__(kBlock) void k_compact(const SceneDev* __restrict__ scenes,
                                                    const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t i = br.block * kBlock + threadIdx.x;
    const uint32_t c = S.state->cur;
    const bool fluid = i < S.n && (S.idt[c][i] >> kTypeShift) == 0u;
    __shared__ uint32_t lw[kBlock / 64];
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(fluid ? 1u : 0u, lw, tot);
    if (fluid) {
        const uint32_t k = S.block_count[br.block] + ex;
        S.fluid_idx[k] = i;
        S.rank_of[i] = k;
    }
}

// ---- 6. neighbor lists ---------------------------------------------------------
// One wave per 64 consecutive sorted particles processes every cell whose run
// starts among them.  For a cell, lane k loads the run of the neighbor cell of
// order code k (NeighborOrder), so the runs lie in ascending key order (the
// reference's std::sort of the 27 keys) without a sort, and candidate g of the
// concatenated runs has ascending combined index.  Up to
// kListChunks chunks of 64 candidates are held one per lane; the cell's fluid
// particles, kStageParticles at a time, are taken in turn, each tested against
// the held chunks with one ballot per chunk, whose prefix (on top of the
// particle's running count, a scalar) gives each accepted candidate its slot in
// the particle's row of an LDS stage.  Rows of consecutive cells gather in the
// stage (the fluid ranks of a wave's cells are consecutive) and leave it row by
// row: one store covers 64 / kStageParticles list rows of up to kStageParticles
// consecutive fluid ranks, instead of one scattered store per particle and chunk.
// With at most kListChunks * 64 candidates, each candidate slot finds its run
// through marks in the stage rows from kMapRow on, and the candidates are loaded
// once for all the cell's particles.
// The next cell's table entries and first particle group are loaded while the
// current cell is processed.

constexpr uint32_t kListChunks = 4;
constexpr uint32_t kStageParticles = 16;
// first stage row of a cell's candidate marks (kListChunks * 64 slots from the
// first 16-byte boundary of the row on)
constexpr uint32_t kMapRow = 13;
constexpr uint32_t kMarkWord = (kMapRow * kMaxNeighbors + 3u) & ~3u;
static_assert(kMarkWord + kListChunks * 64 <= kStageParticles * kMaxNeighbors,
              "the candidate marks fit the stage rows from kMapRow on");
static_assert((kStageParticles * kMaxNeighbors) % 4 == 0, "each wave's stage starts on a 16-byte boundary");

// exclusive prefix sum over the 64 lanes from DPP row shifts and row broadcasts
// (no LDS round trip); total = the sum over all lanes, uniform
__device__ __forceinline__ uint32_t exclusive_scan64(uint32_t v, uint32_t& total) {
    uint32_t x = v;
    // inclusive scans within each row of 16 lanes (row_shr:1, 2, 4, 8; lanes
    // shifted in from outside the row read 0)
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);
    // row_bcast:15 adds lane 15's sum to row 1 and lane 47's to row 3; row_bcast:31
    // then adds lane 31's (rows 0-1) to rows 2 and 3
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);
    total = (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
    return x - v;
}

// maximum over lanes 0..lane of values >= 0, from DPP row shifts and row broadcasts
__device__ __forceinline__ uint32_t inclusive_max64(uint32_t v) {
    uint32_t x = v;
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false));
    return x;
}

__device__ __forceinline__ float readlane_f(float v, uint32_t lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)lane));
}
__device__ __forceinline__ uint32_t readlane_u(uint32_t v, uint32_t lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}
// v with lane `lane` (uniform) replaced by the uniform value x (v_writelane_b32,
// the lane select in M0)
extern "C" __device__ int fsa_writelane_i32(int x, int lane, int v) __asm("llvm.amdgcn.writelane.i32");
__device__ __forceinline__ uint32_t writelane_u(uint32_t v, uint32_t x, uint32_t lane) {
    return (uint32_t)fsa_writelane_i32((int)x, (int)lane, (int)v);
}

// slot * stride for operands below 2^24: the full-rate 24-bit multiply (the
// compiler would otherwise fold the byte shift into the multiplier and issue
// the quarter-rate 32-bit multiply)
__device__ __forceinline__ uint32_t mul_u24(uint32_t v, uint32_t s) {
    uint32_t r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "s"(s), "v"(v));
    return r;
}

// Wide: the list array spans 4 GiB or more (element indices still fit 32 bits),
// so slots are addressed in 64 bits; otherwise a slot is one 24-bit multiply and
// a 32-bit byte offset from the particle's list base.
template <bool Wide>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_lists(const SceneDev* __restrict__ scenes,
                                                  const BlockRef* __restrict__ map, uint32_t sbuf) {
    __shared__ __align__(16) uint32_t stage_all[kBlock / 64][kStageParticles * kMaxNeighbors];
    __shared__ uint32_t stage_cnt_all[kBlock / 64][kStageParticles];
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (sort_retry(S)) return;
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const uint32_t lane = __lane_id();
    const uint32_t base = br.block * kBlock + wave * 64;
    if (base >= S.n) return;
    uint32_t* const stage = stage_all[wave];
    uint32_t* const stage_cnt = stage_cnt_all[wave];
    const uint32_t c = S.state->cur;
    const auto pos = gload_ptr(S.pos[c]);
    const auto idt = gload_ptr(S.idt[c]);
    const auto keys = gload_ptr(S.keys[sbuf]);
    const auto rank_of = gload_ptr(S.rank_of);  // in-block fluid rank (k_permute)
    const auto block_count = gload_ptr(S.block_count);  // fluid before the block (k_fluid_scan)
    FSA_GLOBAL(uint32_t)* __restrict__ nbr = (FSA_GLOBAL(uint32_t)*)S.nbr;
    const float r2 = S.r2;
    const uint32_t stride = S.list_stride;
    // run starts among this wave's particles, with their cells' grid coordinates
    // and neighbor order (left by k_permute in the sort's idle buffers)
    const uint32_t mi = base + lane;
    bool start = false, fl = false;
    uint32_t cgrid = 0, corder = 0;
    if (mi < S.n) {
        const uint32_t mykey = keys[mi];
        start = mi == 0 || keys[mi - 1] != mykey;
        fl = (idt[mi] >> kTypeShift) == 0u;
        if (start) {
            cgrid = S.keys[sbuf ^ 1u][mi];
            corder = S.vals[sbuf ^ 1u][mi];
        }
    }
    // run starts and fluid particles among the next 64 particles, where the
    // wave's last run ends (past the collection every run has ended)
    const uint32_t mj = mi + 64u;
    bool start2 = true, fl2 = false;
    if (mj < S.n) {
        start2 = keys[mj - 1u] != keys[mj];
        fl2 = (idt[mj] >> kTypeShift) == 0u;
    }
    const uint64_t fluid = __ballot(fl), starts2 = __ballot(start2), fluid2 = __ballot(fl2);
    // Only the runs holding a fluid particle need lists (boundary particles get
    // none): the others are dropped before anything of theirs is loaded.  A run
    // that does not end within the next 64 particles is kept.
    uint64_t starts = 0;
    for (uint64_t m = __ballot(start); m; m &= m - 1) {
        const uint32_t s = (uint32_t)__ffsll((unsigned long long)m) - 1u;
        const uint64_t from_s = ~0ull << s;
        const uint64_t later = m & (from_s << 1);  // starts above s
        bool any;
        if (later) {
            const uint32_t e = (uint32_t)__ffsll((unsigned long long)later) - 1u;
            any = (fluid & from_s & ((1ull << e) - 1ull)) != 0;
        } else {
            any = (fluid & from_s) != 0 || starts2 == 0 ||
                  (fluid2 & ((1ull << ((uint32_t)__ffsll((unsigned long long)starts2) - 1u)) - 1ull)) != 0;
        }
        if (any) starts |= 1ull << s;
    }
    if (starts == 0) return;
    bool overflow = false;
    bool codec_check = false;  // some list may exceed the NeighborStorage payload
    // readout lanes: particle column lane % kStageParticles, row offset lane / kStageParticles
    const uint32_t col = lane % kStageParticles;
    const uint32_t row_step = 64u / kStageParticles;
    // Stage row r holds the list of fluid rank fbase + r; rows of consecutive
    // cells gather until the next group's ranks would pass the last row, then
    // leave together: each store covers 64 / kStageParticles list rows of up to
    // kStageParticles consecutive ranks.
    uint32_t fbase = 0, fcount = 0, frows = 0;
    auto flush = [&]() {
        if (fcount == 0) return;
        // the stage rows and counts written by other lanes are complete before they are read
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t ccnt = col < fcount ? stage_cnt[col] : 0u;
        const uint32_t crank = fbase + col;
        for (uint32_t e0 = 0; e0 < frows; e0 += row_step) {
            const uint32_t e = e0 + lane / kStageParticles;
            if (e < ccnt) {
                const uint32_t v = stage[col * (uint32_t)kMaxNeighbors + e];
                if (Wide) {
                    nbr[(uint64_t)e * stride + crank] = v;
                } else {
                    nbr[mul_u24(e, stride) + crank] = v;
                }
            }
        }
        // every lane's reads of the stage are done before it is written again
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        fcount = 0;
        frows = 0;
    };
    // The loads a cell starts from: the table entries of its 27 neighbor cells,
    // lane k holding the neighbor of order code k (so the runs lie in ascending
    // key order by lane; `self` is the cell's own lane), and the first particle
    // group's type, position and rank.  They depend on the cell's key only, so
    // the next cell's are issued while the current one is processed.
    auto cell_loads = [&](uint32_t sl, uint4& ent, uint32_t& self, uint32_t& gidt, float4& gpos, uint32_t& grank) {
        const uint32_t g = readlane_u(cgrid, sl);
        const NeighborOrder o = unpack_order(readlane_u(corder, sl));
        self = o.self;
        ent = make_uint4(0u, 0u, 0u, 0u);  // stamp 0 never matches: an absent cell
        int32_t dx, dy, dz;
        const bool has = neighbor_offset(o, lane, 0, dx) & neighbor_offset(o, lane, 1, dy) &
                         neighbor_offset(o, lane, 2, dz);
        // grid coordinates of the neighbor; -1 wraps past the grid's extent
        const uint32_t x = (g & 1023u) + (uint32_t)dx, y = ((g >> 10) & 1023u) + (uint32_t)dy,
                       z = (g >> 20) + (uint32_t)dz;
        if (has && x < S.cdim[0] && y < S.cdim[1] && z < S.cdim[2]) ent = S.cells[x + S.cdim[0] * (y + S.cdim[1] * z)];
        const uint32_t p = base + sl + lane;
        gidt = 0xffffffffu;  // not a fluid particle
        gpos = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        grank = 0;
        if (lane < kStageParticles && p < S.n) {
            gidt = idt[p];
            gpos = pos[p];
            grank = block_count[p / kBlock] + rank_of[p];
        }
    };
    uint32_t sl0 = (uint32_t)__ffsll((unsigned long long)starts) - 1u;
    starts &= starts - 1;
    uint4 ent;
    uint32_t gidt, grank, eself;
    float4 gpos;
    cell_loads(sl0, ent, eself, gidt, gpos, grank);
    while (true) {
        const uint32_t cs = base + sl0;
        const uint4 e = ent;
        const uint32_t idt0 = gidt, rank0 = grank;
        const float4 pos0 = gpos;
        const uint32_t ce = readlane_u(e.y, eself);
        const bool more = starts != 0;
        uint32_t nsl = 0;
        if (more) {
            nsl = (uint32_t)__ffsll((unsigned long long)starts) - 1u;
            starts &= starts - 1;
        }
        // a cell without fluid particles needs no lists
        const bool f0 = lane < kStageParticles && cs + lane < ce && (idt0 >> kTypeShift) == 0u;
        const bool work = __ballot(f0) != 0 || ce - cs > kStageParticles;
        // the runs of the neighbor cells in key order (the reference's std::sort
        // of the 27 keys) by lane; absent cells are empty runs
        uint32_t rs = 0, rl = 0;
        if (e.z == S.stamp) {
            rs = e.x;
            rl = e.y - e.x;
        }
        if (more) cell_loads(nsl, ent, eself, gidt, gpos, grank);
        if (!work) {
            if (!more) break;
            sl0 = nsl;
            continue;
        }
        uint32_t total;
        const uint32_t roff = exclusive_scan64(rl, total);
        // NeighborStorage bound: the cell's non-empty runs (in key order, so in
        // ascending rows) fall into `segs` groups of consecutive rows (run i
        // continues the previous non-empty run exactly when rs - roff is the same
        // for both).  While no group spans more than 259 rows, a delta inside a
        // group is at most 258 (at most 1 payload byte) and only a move to the
        // next group can take 4, so a list of c entries needs at most
        // c - 1 + 3 (segs - 1) bytes.  With at most 259 candidates no group can
        // be longer; otherwise the longest group is measured.
        uint32_t segs = 0;
        bool long_group = false;
        {
            const bool ne = rl != 0;
            const uint32_t jb = ne ? rs - roff + 1u : 0u;
            const uint32_t upto = inclusive_max64(jb);
            const uint32_t before = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)upto, 0x138, 0xf, 0xf, false);
            const bool newseg = ne && jb > before;
            segs = (uint32_t)__popcll(__ballot(newseg));
            if (total > 259u) {
                // each non-empty run's group starts at the last group start at or before it
                const uint32_t gs = inclusive_max64(newseg ? roff + 1u : 0u);
                const uint32_t glen = ne ? roff + rl - (gs - 1u) : 0u;
                long_group = readlane_u(inclusive_max64(glen), 63) > 259u;
            }
        }
        // Up to kListChunks * 64 candidates (the usual case): each non-empty run
        // marks the slot of its first candidate with its lane + 1, so the run of
        // candidate g = k * 64 + lane is the largest mark at or below slot g (a
        // running maximum over the lanes, carried from chunk to chunk), and its
        // element index is rs - roff of that run, plus g.  The candidates are
        // loaded once for every particle group of the cell.  More candidates take
        // the binary search below, chunk by chunk and group by group.
        const bool mapped = total <= kListChunks * 64u;
        float4 pjm[kListChunks];
        uint32_t ejm[kListChunks];
        if (mapped) {
            // the marks take the stage from row kMapRow on, which no gathered row may hold
            if (fcount > kMapRow) flush();
            uint32_t* const marks = stage + kMarkWord;
            reinterpret_cast<uint4*>(marks)[lane] = make_uint4(0u, 0u, 0u, 0u);  // the last cell's marks
            if (rl != 0) marks[roff] = lane + 1u;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const uint32_t jbase = rs - roff;  // modulo 2^32
            uint32_t carry = 0;
#pragma unroll
            for (uint32_t k = 0; k < kListChunks; k++) {
                pjm[k] = make_float4(__builtin_inff(), __builtin_inff(), __builtin_inff(), 0.0f);
                ejm[k] = 0;
                if (k * 64 < total) {
                    const uint32_t g = k * 64 + lane;
                    // slot 0 holds the first non-empty run's mark, so q1 >= 1
                    const uint32_t q1 = max(inclusive_max64(marks[g]), carry);
                    carry = readlane_u(q1, 63);
                    const uint32_t j = (uint32_t)__shfl((int)jbase, (int)(q1 - 1u), 64) + g;
                    if (g < total) {
                        pjm[k] = pos[j];
                        ejm[k] = j | ((idt[j] >> kTypeShift) == 1u ? kBoundaryBit : 0u);
                    }
                }
            }
            // the marks are read before the particle groups write their lists there
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        for (uint32_t pbase = cs; pbase < ce; pbase += kStageParticles) {
            // the group's particles in lanes 0..kStageParticles-1 (the first group's were loaded ahead)
            const uint32_t pi = pbase + lane;
            uint32_t ti = idt0, kl = rank0;
            float4 xl = pos0;
            if (pbase != cs) {
                ti = 0xffffffffu;
                if (lane < kStageParticles && pi < ce) {
                    ti = idt[pi];
                    xl = pos[pi];
                    kl = block_count[pi / kBlock] + rank_of[pi];
                }
            }
            const bool isf = lane < kStageParticles && pi < ce && (ti >> kTypeShift) == 0u;
            const uint64_t fmask = __ballot(isf);
            if (fmask == 0) continue;
            // the group's fluid ranks are consecutive: r0 .. r0 + nfl - 1
            const uint32_t r0 = readlane_u(kl, (uint32_t)__ffsll((unsigned long long)fmask) - 1u);
            const uint32_t nfl = (uint32_t)__popcll(fmask);
            if (fcount != 0 && (r0 != fbase + fcount || r0 + nfl - fbase > kStageParticles)) flush();
            if (fcount == 0) fbase = r0;
            uint32_t cntl = 0;
            for (uint32_t c0 = 0; c0 < total; c0 += kListChunks * 64) {
                // candidates c0 + 64k + lane; a lane past the last candidate sits
                // at infinity and fails every test
                // the mapped candidates were loaded above; otherwise this chunk's
                float4* const pj = pjm;
                uint32_t* const ej = ejm;
#pragma unroll
                for (uint32_t k = 0; k < kListChunks; k++) {
                    if (mapped) break;
                    pj[k] = make_float4(__builtin_inff(), __builtin_inff(), __builtin_inff(), 0.0f);
                    ej[k] = 0;
                    if (c0 + k * 64 < total) {
                        // run q of candidate g: the last lane whose offset is at most g
                        // (an empty run shares its offset with the next run, so the
                        // last such lane holds a non-empty run); every lane takes
                        // part in each shuffle
                        const uint32_t g = c0 + k * 64 + lane;
                        uint32_t q = 0;
#pragma unroll
                        for (uint32_t step = 32; step > 0; step >>= 1) {
                            const uint32_t probe = q + step;
                            const uint32_t off = __shfl(roff, (int)probe, 64);
                            if (off <= g) q = probe;
                        }
                        const uint32_t qs = __shfl(rs, (int)q, 64);
                        const uint32_t qo = __shfl(roff, (int)q, 64);
                        if (g < total) {
                            const uint32_t j = qs + (g - qo);
                            pj[k] = pos[j];
                            ej[k] = j | ((idt[j] >> kTypeShift) == 1u ? kBoundaryBit : 0u);
                        }
                    }
                }
                // the chunks holding candidates, 1..kListChunks (the loop below is
                // compiled for each count, so no chunk test is gated at run time)
                const uint32_t left = total - c0;
                const uint32_t nch = min((left + 63u) / 64u, kListChunks);
                uint64_t fm = fmask;
                // the fluid lanes' ranks are r0, r0 + 1, ... in lane order: one stage row each
                uint32_t* srow = stage + (r0 - fbase) * (uint32_t)kMaxNeighbors;
                auto test_group = [&](auto nchunks) {
                    while (fm) {
                        const uint32_t b = (uint32_t)__ffsll((unsigned long long)fm) - 1u;
                        fm &= fm - 1;
                        const float xs = readlane_f(xl.x, b), ys = readlane_f(xl.y, b), zs = readlane_f(xl.z, b);
                        uint32_t cnt = readlane_u(cntl, b);  // the particle's running count (uniform)
#pragma unroll
                        for (uint32_t k = 0; k < decltype(nchunks)::value; k++) {
                            const float dx = pj[k].x - xs, dy = pj[k].y - ys, dz = pj[k].z - zs;
                            const bool pass = dot3(dx, dy, dz, dx, dy, dz) <= r2;
                            const uint64_t m = __ballot(pass);
                            if (pass) {
                                // past the capacity the last entry is overwritten; the count reports the overflow
                                const uint32_t slot =
                                    min(__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)m, cnt)),
                                        (uint32_t)kMaxNeighbors - 1u);
                                srow[slot] = ej[k];
                            }
                            cnt += (uint32_t)__popcll(m);
                        }
                        cntl = writelane_u(cntl, cnt, b);
                        // longest list of the gathered rows (uniform; counts only grow)
                        frows = max(frows, min(cnt, (uint32_t)kMaxNeighbors));
                        srow += kMaxNeighbors;
                    }
                };
                static_assert(kListChunks == 4, "one specialisation per chunk count");
                if (nch == 4)
                    test_group(std::integral_constant<uint32_t, 4>{});
                else if (nch == 3)
                    test_group(std::integral_constant<uint32_t, 3>{});
                else if (nch == 2)
                    test_group(std::integral_constant<uint32_t, 2>{});
                else
                    test_group(std::integral_constant<uint32_t, 1>{});
            }
            const uint32_t kept = min(cntl, (uint32_t)kMaxNeighbors);
            if (isf) {
                if (cntl > (uint32_t)kMaxNeighbors) overflow = true;
                if (long_group || cntl + 3u * segs > 100u) codec_check = true;
                S.nbr_cnt[kl] = kept;
                S.fluid_idx[kl] = pi;
                stage_cnt[kl - fbase] = kept;
            }
            fcount = r0 + nfl - fbase;
        }
        if (!more) break;
        sl0 = nsl;
    }
    flush();
    if (__any(overflow) && lane == 0) atomicOr(&S.state->status, kStatusCapacity);
    if ((S.flags & kFlagCodec) && __any(codec_check) && lane == 0) atomicOr(&S.state->codec_check, 1u);
}

// ---- 8. hashed / quadratic 3D searches (collection order kept) -------------------
// HashedNeighborhoodSearch3D (HashedNeighborhoodSearch3D.cpp:69-217) and
// QuadraticNeighborhoodSearch3D (QuadraticNeighborhoodSearch3D.cpp:32-69) do
// not sort the collection each step.  Particles are bucketed by a stable sort
// on their dense cell index, so each cell lists its members ascending (the
// hashed grid's std::set); a list walks the 27 cells dx-outer, dz-inner
// (hashed), or is ascending overall (quadratic).  Neighbor test
// dot(x_i - x_j, x_i - x_j) <= r^2.
__global__ __launch_bounds__(kBlock) void k3u_keys(const SceneDev* __restrict__ scenes,
                                                   const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t i = br.block * kBlock + threadIdx.x;
    if (i >= S.n) return;
    const uint32_t c = S.state->cur;
    const float4 p = S.pos[c][i];
    const int32_t cx = cell_of(p.x, S.search_radius), cy = cell_of(p.y, S.search_radius),
                  cz = cell_of(p.z, S.search_radius);
    S.cell_last[i] = make_int4(cx, cy, cz, 0);
    uint32_t lin = S.cdim[3];
    if (!cell_linear(S, cx, cy, cz, lin)) atomicOr(&S.state->status, kStatusGrid);
    S.keys[0][i] = lin;
    S.vals[0][i] = i;
}

__global__ __launch_bounds__(kBlock) void k3u_cell_table(const SceneDev* __restrict__ scenes,
                                                         const BlockRef* __restrict__ map, uint32_t buf) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t k = br.block * kBlock + threadIdx.x;
    if (k >= S.n) return;
    const uint32_t* __restrict__ keys = S.keys[buf];
    const uint32_t key = keys[k];
    if (key >= S.cdim[3]) return;
    const bool first = k == 0 || keys[k - 1] != key;
    const bool last = k + 1 == S.n || keys[k + 1] != key;
    if (first) {
        S.cells[key].x = k;
        S.cells[key].z = S.stamp;
    }
    if (last) S.cells[key].y = k + 1;
}

// One particle's list by a serial walk of the 27 cells (dx-outer, dz-inner),
// kept ascending by insertion for the quadratic order.  The path for cells
// whose quadratic candidate set exceeds the held chunks.
__device__ void k3u_list_serial(const SceneDev& S, uint32_t k, uint32_t i, uint32_t buf) {
    const uint32_t c = S.state->cur;
    const auto pos = gload_ptr(S.pos[c]);
    const auto idt = gload_ptr(S.idt[c]);
    const float4 xi = pos[i];
    const int4 ci = S.cell_last[i];
    const uint32_t* __restrict__ vals = S.vals[buf];
    uint32_t* __restrict__ out = S.nbr + k;
    const uint32_t stride = S.list_stride;
    const bool quad = (S.flags & kFlagQuadratic) != 0;
    const float r2 = S.r2;
    uint32_t cnt = 0;
    for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dz = -1; dz <= 1; dz++) {
                uint32_t lin;
                if (!cell_linear(S, ci.x + dx, ci.y + dy, ci.z + dz, lin)) continue;
                const uint4 e = S.cells[lin];
                if (e.z != S.stamp) continue;
                for (uint32_t m = e.x; m < e.y; m++) {
                    const uint32_t j = vals[m];
                    const float4 xj = pos[j];
                    const float ddx = xi.x - xj.x, ddy = xi.y - xj.y, ddz = xi.z - xj.z;
                    if (!(dot3(ddx, ddy, ddz, ddx, ddy, ddz) <= r2)) continue;
                    const uint32_t ent = j | ((idt[j] >> kTypeShift) == 1u ? kBoundaryBit : 0u);
                    if (quad) {
                        // ascending index: insert before the first larger entry
                        uint32_t q = min(cnt, (uint32_t)kMaxNeighbors);
                        while (q > 0) {
                            const uint32_t prev = out[(q - 1) * stride];
                            if ((prev & kIndexMask) < j) break;
                            if (q < (uint32_t)kMaxNeighbors) out[q * stride] = prev;
                            q--;
                        }
                        if (q < (uint32_t)kMaxNeighbors) out[q * stride] = ent;
                    } else if (cnt < (uint32_t)kMaxNeighbors) {
                        out[cnt * stride] = ent;
                    }
                    cnt++;
                }
            }
    if (cnt > (uint32_t)kMaxNeighbors) atomicOr(&S.state->status, kStatusCapacity);
    S.nbr_cnt[k] = min(cnt, (uint32_t)kMaxNeighbors);
}

// ascending bitonic sort of R x 64 keys, element r * 64 + lane in v[r]
template <uint32_t R>
__device__ __forceinline__ void sort_regs_u32(uint32_t (&v)[R]) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (uint32_t size = 2; size <= 64 * R; size <<= 1) {
#pragma unroll
        for (uint32_t st = size >> 1; st > 0; st >>= 1) {
            uint32_t o[R];
#pragma unroll
            for (uint32_t r = 0; r < R; r++) {
                const uint32_t p = r * 64 + lane;
                const bool asc = (p & size) == 0;
                uint32_t b;
                bool lower;
                if (st >= 64) {
                    b = v[r ^ (st >> 6)];
                    lower = (r & (st >> 6)) == 0;
                } else {
                    b = (uint32_t)__shfl_xor((int)v[r], (int)st, 64);
                    lower = (lane & st) == 0;
                }
                o[r] = (lower == asc) ? min(v[r], b) : max(v[r], b);
            }
#pragma unroll
            for (uint32_t r = 0; r < R; r++) v[r] = o[r];
        }
    }
}

// candidates the quadratic order sorts in registers: two chunk sets
constexpr uint32_t kQuadSorted = 2 * kListChunks * 64;

// Wave-cooperative lists for the hashed and quadratic orders, in the bucket
// array (particles stably sorted by cell): one wave takes the cells whose runs
// start among 64 consecutive bucket positions.  A cell's candidates are the
// runs of its 27 neighbor cells, concatenated dx-outer, dz-inner (the hashed
// order: each run lists its members ascending); for the quadratic order the
// held candidates are sorted by particle index (ascending overall).  The
// cell's fluid particles, kStageParticles at a time, are tested against the
// held chunks as in k_lists, staged in LDS, and written row by row.
__global__ __launch_bounds__(kBlock) void k3u_lists(const SceneDev* __restrict__ scenes,
                                                    const BlockRef* __restrict__ map, uint32_t buf) {
    __shared__ uint32_t stage_all[kBlock / 64][kStageParticles * kMaxNeighbors];
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const uint32_t lane = __lane_id();
    const uint32_t base = br.block * kBlock + wave * 64;
    if (base >= S.n) return;
    uint32_t* const stage = stage_all[wave];
    const uint32_t c = S.state->cur;
    const auto pos = gload_ptr(S.pos[c]);
    const auto idt = gload_ptr(S.idt[c]);
    const auto keys = gload_ptr(S.keys[buf]);
    const auto vals = gload_ptr(S.vals[buf]);
    const auto rank_of = gload_ptr(S.rank_of);
    FSA_GLOBAL(uint32_t)* __restrict__ nbr = (FSA_GLOBAL(uint32_t)*)S.nbr;
    const float r2 = S.r2;
    const uint32_t stride = S.list_stride;
    const bool quad = (S.flags & kFlagQuadratic) != 0;
    const uint32_t mi = base + lane;
    bool start = false;
    uint32_t mykey = 0;
    if (mi < S.n) {
        mykey = keys[mi];
        start = mi == 0 || keys[mi - 1] != mykey;
    }
    uint64_t starts = __ballot(start);
    bool overflow = false;
    const uint32_t col = lane % kStageParticles;
    const uint32_t row_step = 64u / kStageParticles;
    while (starts) {
        const uint32_t sl0 = (uint32_t)__ffsll((unsigned long long)starts) - 1u;
        starts &= starts - 1;
        const uint32_t cs = base + sl0;
        const uint32_t lin0 = readlane_u(mykey, sl0);
        if (lin0 >= S.cdim[3]) continue;  // outside the grid: reported by k3u_keys
        const int32_t cx = (int32_t)(lin0 % S.cdim[0]) + S.cmin[0];
        const int32_t cy = (int32_t)((lin0 / S.cdim[0]) % S.cdim[1]) + S.cmin[1];
        const int32_t cz = (int32_t)(lin0 / (S.cdim[0] * S.cdim[1])) + S.cmin[2];
        uint4 e = make_uint4(0u, 0u, 0u, 0u);
        if (lane < 27) {
            uint32_t lin;
            if (cell_linear(S, cx + (int)(lane / 9) - 1, cy + (int)((lane / 3) % 3) - 1, cz + (int)(lane % 3) - 1,
                            lin))
                e = S.cells[lin];
        }
        const uint32_t ce = readlane_u(e.y, 13);
        // runs in lane order (dx-outer, dz-inner); absent cells are empty runs
        uint32_t rs = 0, rl = 0;
        if (lane < 27 && e.z == S.stamp) {
            rs = e.x;
            rl = e.y - e.x;
        }
        uint32_t total;
        const uint32_t roff = exclusive_scan64(rl, total);
        const bool serial = quad && total > kQuadSorted;
        // quadratic: every candidate index, sorted ascending (two chunk sets at most)
        uint32_t sj[2 * kListChunks];
        if (quad && !serial) {
#pragma unroll
            for (uint32_t k = 0; k < 2 * kListChunks; k++) {
                sj[k] = 0xffffffffu;
                const uint32_t g = k * 64 + lane;
                uint32_t q = 0;
#pragma unroll
                for (uint32_t step = 16; step > 0; step >>= 1) {
                    const uint32_t probe = q + step;
                    const uint32_t off = __shfl(roff, (int)min(probe, 31u), 64);
                    if (probe < 27u && off <= g) q = probe;
                }
                const uint32_t qs = __shfl(rs, (int)q, 64);
                const uint32_t qo = __shfl(roff, (int)q, 64);
                if (g < total) sj[k] = vals[qs + (g - qo)];
            }
            if (total <= kListChunks * 64) {
                uint32_t lo[kListChunks];
#pragma unroll
                for (uint32_t k = 0; k < kListChunks; k++) lo[k] = sj[k];
                sort_regs_u32<kListChunks>(lo);
#pragma unroll
                for (uint32_t k = 0; k < kListChunks; k++) sj[k] = lo[k];
            } else {
                sort_regs_u32<2 * kListChunks>(sj);
            }
        }
        for (uint32_t pbase = cs; pbase < ce; pbase += kStageParticles) {
            const uint32_t pm = pbase + lane;
            uint32_t pi = 0, ti = 0xffffffffu;
            if (lane < kStageParticles && pm < ce) {
                pi = vals[pm];
                ti = idt[pi];
            }
            const bool isf = lane < kStageParticles && pm < ce && (ti >> kTypeShift) == 0u;
            const uint64_t fmask = __ballot(isf);
            if (fmask == 0) continue;
            float4 xl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            uint32_t kl = 0;
            if (isf) {
                xl = pos[pi];
                kl = rank_of[pi];
            }
            if (serial) {
                // more quadratic candidates than the held chunks: each lane walks its particle
                if (isf) k3u_list_serial(S, kl, pi, buf);
                continue;
            }
            uint32_t cntl = 0;
            for (uint32_t c0 = 0; c0 < total; c0 += kListChunks * 64) {
                // candidate c0 + 64k + lane: bucket position by binary search over the run offsets
                uint32_t jj[kListChunks];
#pragma unroll
                for (uint32_t k = 0; k < kListChunks; k++) {
                    jj[k] = 0xffffffffu;
                    if (quad) {
                        jj[k] = c0 == 0 ? sj[k] : sj[kListChunks + k];
                    } else if (c0 + k * 64 < total) {
                        const uint32_t g = c0 + k * 64 + lane;
                        uint32_t q = 0;
#pragma unroll
                        for (uint32_t step = 16; step > 0; step >>= 1) {
                            const uint32_t probe = q + step;
                            const uint32_t off = __shfl(roff, (int)min(probe, 31u), 64);
                            if (probe < 27u && off <= g) q = probe;
                        }
                        const uint32_t qs = __shfl(rs, (int)q, 64);
                        const uint32_t qo = __shfl(roff, (int)q, 64);
                        if (g < total) jj[k] = vals[qs + (g - qo)];
                    }
                }
                float4 pj[kListChunks];
                uint32_t ej[kListChunks];
#pragma unroll
                for (uint32_t k = 0; k < kListChunks; k++) {
                    pj[k] = make_float4(__builtin_inff(), __builtin_inff(), __builtin_inff(), 0.0f);
                    ej[k] = 0;
                    if (jj[k] != 0xffffffffu) {
                        const uint32_t j = jj[k];
                        pj[k] = pos[j];
                        ej[k] = j | ((idt[j] >> kTypeShift) == 1u ? kBoundaryBit : 0u);
                    }
                }
                const uint32_t left = total - c0;
                uint64_t fm = fmask;
                while (fm) {
                    const uint32_t b = (uint32_t)__ffsll((unsigned long long)fm) - 1u;
                    fm &= fm - 1;
                    const float xs = readlane_f(xl.x, b), ys = readlane_f(xl.y, b), zs = readlane_f(xl.z, b);
                    uint32_t cnt = readlane_u(cntl, b);
                    uint32_t* const srow = stage + b * (uint32_t)kMaxNeighbors;
#pragma unroll
                    for (uint32_t k = 0; k < kListChunks; k++) {
                        if (k * 64 < left) {
                            const float dx = xs - pj[k].x, dy = ys - pj[k].y, dz = zs - pj[k].z;
                            const bool pass = dot3(dx, dy, dz, dx, dy, dz) <= r2;
                            const uint64_t m = __ballot(pass);
                            if (pass) {
                                const uint32_t slot =
                                    min(__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)m, cnt)),
                                        (uint32_t)kMaxNeighbors - 1u);
                                srow[slot] = ej[k];
                            }
                            cnt += (uint32_t)__popcll(m);
                        }
                    }
                    if (lane == b) cntl = cnt;
                }
            }
            const uint32_t kept = min(cntl, (uint32_t)kMaxNeighbors);
            if (isf) {
                if (cntl > (uint32_t)kMaxNeighbors) overflow = true;
                S.nbr_cnt[kl] = kept;
            }
            uint32_t rows = 0;
            for (uint64_t fm = fmask; fm; fm &= fm - 1)
                rows = max(rows, readlane_u(kept, (uint32_t)__ffsll((unsigned long long)fm) - 1u));
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const uint32_t ccnt = (uint32_t)__shfl(kept, (int)col, 64);
            const uint32_t crank = (uint32_t)__shfl(kl, (int)col, 64);
            const bool cfluid = (fmask >> col) & 1u;
            for (uint32_t e0 = 0; e0 < rows; e0 += row_step) {
                const uint32_t ee = e0 + lane / kStageParticles;
                if (cfluid && ee < ccnt) nbr[(uint64_t)ee * stride + crank] = stage[col * (uint32_t)kMaxNeighbors + ee];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    if (__any(overflow) && lane == 0) atomicOr(&S.state->status, kStatusCapacity);
}

// ---- 7. NeighborStorage codec (CompressedNeighbors.cpp:260-330) -----------------
// Size of each list in the reference's delta code: the first index, then per
// delta a 2-bit control and 0 (delta 1, 2), 1 (delta <= 258) or 4 payload
// bytes.  A list needs more than 96 payload bytes or more than 96 deltas
// where the reference's fixed NeighborStorage overflows; the scene then
// reports the capacity error.  Block totals of neighbors and encoded bytes go
// to the scene state.
// With Encode, each list is also written delta-coded (kFlagDeltaLists): one
// 16-bit code per entry, the first entry and deltas beyond kDeltaMax in the
// escape rows (layout at kDeltaCodeRows in the device header).
// Without `force`, a scene's lists are sized only when k_lists could not bound
// them within the NeighborStorage payload, or when they are delta-coded.
__device__ __forceinline__ uint32_t scene_fluid_parts(const SceneDev& S) { return (S.nf + kBlock - 1) / kBlock; }
__device__ __forceinline__ bool codec_due(const SceneDev& S, uint32_t force) {
    return !sort_retry(S) && (force || (S.flags & kFlagDeltaLists) || S.state->codec_check);
}

// one kBlock part of a scene's fluid lists: sizes, capacity, delta codes; the part's totals in partial[br.block]
template <bool Encode>
__device__ __forceinline__ void codec_part(const SceneDev& S, const BlockRef br, uint64_t (*wsum)[kBlock / 64]) {
    const uint32_t k = br.block * kBlock + threadIdx.x;
    const bool encode = Encode && (S.flags & kFlagDeltaLists) != 0;
    uint64_t cnt = 0, bytes = 0, dbytes = 0;
    bool over = false, spill = false;
    if (k < S.nf) {
        const uint32_t n = S.nbr_cnt[k];
        cnt = n;
        if (n) {
            const auto nb = gload_ptr(S.nbr + k);
            const uint32_t stride = S.list_stride;
            FSA_GLOBAL(uint32_t)* __restrict__ codes = (FSA_GLOBAL(uint32_t)*)S.dlist + k;
            FSA_GLOBAL(uint32_t)* __restrict__ esc = codes + kDeltaCodeRows * stride;
            const uint32_t first = nb[0];
            uint32_t prev = first & kIndexMask, payload = 0;
            uint32_t nesc = 1;  // the first entry is escape 0
            uint32_t low = 0;   // code of the last even entry, waiting for its pair (entry 0: an escape)
            auto code = [&](uint32_t e, uint32_t ent) {
                const uint32_t j = ent & kIndexMask;
                const uint32_t d = j - prev;
                prev = j;
                payload += d <= 2u ? 0u : (d <= 258u ? 1u : 4u);
                if (encode) {
                    uint32_t c = 0;
                    if (d <= kDeltaMax) {
                        c = d | ((ent >> 16) & 0x8000u);
                    } else {
                        if (nesc < kDeltaEscRows) esc[nesc * stride] = ent;
                        nesc++;
                    }
                    if (e & 1u)
                        codes[(e / 2u) * stride] = low | c << 16;
                    else
                        low = c;
                }
            };
            uint32_t e = 1;
            // eight independent loads in flight, then the deltas in list order
            for (; e + 8 <= n; e += 8) {
                uint32_t j[8];
#pragma unroll
                for (int q = 0; q < 8; q++) j[q] = nb[(e + q) * stride];
#pragma unroll
                for (int q = 0; q < 8; q++) code(e + q, j[q]);
            }
            for (; e < n; e++) code(e, nb[e * stride]);
            if (encode) {
                if (n & 1u) codes[((n - 1u) / 2u) * stride] = low;  // the last entry is even
                esc[0] = first | (nesc > 2u ? kDeltaManyBit : 0u);
                spill = nesc > kDeltaEscRows;
                dbytes = 4u * ((n + 1u) / 2u) + 4u * min(nesc, kDeltaEscRows);
            }
            over = payload > 96u || n - 1u > 96u;
            bytes = 4u + (2u * (n - 1u) + 7u) / 8u + payload;
        }
    }
    if (encode && __any(spill) && __lane_id() == 0) atomicOr(&S.state->delta_spill, 1u);
    if ((S.flags & kFlagCodec) && __any(over) && __lane_id() == 0) atomicOr(&S.state->status, kStatusCapacity);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        cnt += __shfl_xor(cnt, o, 64);
        bytes += __shfl_xor(bytes, o, 64);
        dbytes += __shfl_xor(dbytes, o, 64);
    }
    if (__lane_id() == 0) {
        wsum[0][threadIdx.x / 64] = cnt;
        wsum[1][threadIdx.x / 64] = bytes;
        wsum[2][threadIdx.x / 64] = dbytes;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t c = 0, b = 0, db = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            c += wsum[0][w];
            b += wsum[1][w];
            db += wsum[2][w];
        }
        // block totals as a partial: neighbors in count, encoded bytes in err, delta-coded bytes in aux
        Partial p{};
        p.count = (uint32_t)c;
        p.err = (double)b;
        p.aux = (uint32_t)db;
        S.partial[br.block] = p;
    }
}

// the scene's totals from the part partials, in part order (one block)
__device__ __forceinline__ void codec_totals(const SceneDev& S, uint64_t (*wsum)[kBlock / 64]) {
    const uint32_t nparts = (S.nf + kBlock - 1) / kBlock;
    uint64_t c = 0, b = 0, db = 0;
    // integer sums: 16 partials per thread are loaded ahead of their adds
    constexpr uint32_t kAhead = 16;
    for (uint32_t q0 = threadIdx.x; q0 < nparts; q0 += kAhead * kBlock) {
        uint32_t qc[kAhead], qd[kAhead];
        double qe[kAhead];
#pragma unroll
        for (uint32_t u = 0; u < kAhead; u++) {
            const uint32_t q = q0 + u * kBlock;
            qc[u] = 0;
            qd[u] = 0;
            qe[u] = 0.0;
            if (q < nparts) {
                qc[u] = S.partial[q].count;
                qd[u] = S.partial[q].aux;
                qe[u] = S.partial[q].err;
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kAhead; u++) {
            c += qc[u];
            db += qd[u];
            b += (uint64_t)qe[u];
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        c += __shfl_xor(c, o, 64);
        b += __shfl_xor(b, o, 64);
        db += __shfl_xor(db, o, 64);
    }
    if (__lane_id() == 0) {
        wsum[0][threadIdx.x / 64] = c;
        wsum[1][threadIdx.x / 64] = b;
        wsum[2][threadIdx.x / 64] = db;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t tc = 0, tb = 0, tdb = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            tc += wsum[0][w];
            tb += wsum[1][w];
            tdb += wsum[2][w];
        }
        S.state->total_neighbors = tc;
        S.state->codec_bytes = tb;
        // this search's delta lists are walked unless one of them did not fit
        S.state->delta_bytes = tdb;
        S.state->delta_plain = S.state->delta_spill;
        S.state->delta_spill = 0u;
        S.state->codec_check = 0u;
    }
}


// The lists of the scenes due (codec_due) in parts of kBlock fluid ranks: the
// blocks of the launch take the map's parts in turn, and the block that
// completes a scene's last part (counted in state->done[0] after a device-wide
// fence) sums the scene's part totals.  Scenes not due leave at once.
template <bool Encode>
__global__ __launch_bounds__(kBlock) void k_codec(const SceneDev* __restrict__ scenes,
                                                  const BlockRef* __restrict__ map, uint32_t nmap, uint32_t force) {
    __shared__ uint64_t wsum[3][kBlock / 64];
    __shared__ uint32_t last;
    for (uint32_t m = blockIdx.x; m < nmap; m += gridDim.x) {
        const BlockRef br = map[m];
        const SceneDev& S = scenes[br.scene];
        if (!codec_due(S, force)) continue;
        codec_part<Encode>(S, br, wsum);
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();  // this part's totals are visible before it is counted
            last = atomicAdd(&S.state->done[0], 1u) == scene_fluid_parts(S) - 1u ? 1u : 0u;
        }
        __syncthreads();
        if (last) {
            __threadfence();  // every part's totals are read after the count
            codec_totals(S, wsum);
            if (threadIdx.x == 0) S.state->done[0] = 0u;
        }
        __syncthreads();
    }
}

// ---- launchers -------------------------------------------------------------------
// keys, then the passes of the radix sort; the first pass's histogram comes with the keys
// the compressed search's sort: kSearchSortBits-bit digits
void launch_keys_sort(const LaunchSet& L, uint32_t passes, hipStream_t st) {
    constexpr int B = kSearchSortBits;
    if (L.nb_tiles) k_keys_hist<B><<<L.nb_tiles, kSortThreads, 0, st>>>(L.scenes, L.map_tiles);
    for (uint32_t p = 0; p < passes; p++) {
        const uint32_t src = p & 1u;
        const uint32_t shift = p * (uint32_t)B;
        if (p > 0) k_radix_hist<B><<<L.nb_tiles, kSortThreads, 0, st>>>(L.scenes, L.map_tiles, shift, src);
        k_radix_scan<B><<<L.nb_scan, kBlock, 0, st>>>(L.scenes, L.map_scan);
        k_radix_scatter<B><<<L.nb_tiles, kSortThreads, 0, st>>>(L.scenes, L.map_tiles, shift, src);
    }
}

// the compressed search's counting sort over rank keys (sorted pairs in buffer 1)
void launch_count_sort(const LaunchSet& L, hipStream_t st) {
    if (!L.nb_all) return;
    k_rank_count<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
    k_bucket_sums<<<L.nb_cnt, kBlock, 0, st>>>(L.scenes, L.map_cnt);
    k_bucket_offsets<<<L.nb_cnt, kBlock, 0, st>>>(L.scenes, L.map_cnt);
    k_place<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
    k_fixup<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
}

void launch_rank_table(uint32_t* ranks, const int32_t cmin[3], const uint32_t cdim[3], uint32_t levels,
                       hipStream_t st) {
    const uint64_t cells = (uint64_t)cdim[0] * cdim[1] * cdim[2];
    if (cells)
        k_rank_table<<<(uint32_t)((cells + kBlock - 1) / kBlock), kBlock, 0, st>>>(ranks, cmin[0], cmin[1], cmin[2],
                                                                                  cdim[0], cdim[1], cdim[2], levels);
}

void launch_sort_range(const LaunchSet& L, uint32_t first_src, uint32_t shift0, uint32_t passes, hipStream_t st) {
    for (uint32_t p = 0; p < passes; p++) {
        const uint32_t src = first_src ^ (p & 1u);
        const uint32_t shift = shift0 + p * kRadixBits;
        k_radix_hist<kRadixBits><<<L.nb_tiles, kSortThreads, 0, st>>>(L.scenes, L.map_tiles, shift, src);
        k_radix_scan<kRadixBits><<<L.nb_scan, kBlock, 0, st>>>(L.scenes, L.map_scan);
        k_radix_scatter<kRadixBits><<<L.nb_tiles, kSortThreads, 0, st>>>(L.scenes, L.map_tiles, shift, src);
    }
}

// the fluid ranks come from k_permute's in-block ranks and k_fluid_scan's block
// offsets; k_lists writes fluid_idx
void launch_reorder(const LaunchSet& L, uint32_t sbuf, hipStream_t st) {
    k_permute<true><<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all, sbuf);
    k_fluid_scan<<<L.nb_scene, kScanThreads, 0, st>>>(L.scenes, L.map_scene);
}

void launch_lists(const LaunchSet& L, uint32_t sbuf, hipStream_t st) {
    if (L.nb_all) {
        if (L.wide_lists)
            k_lists<true><<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all, sbuf);
        else
            k_lists<false><<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all, sbuf);
    }
}

void launch3u_search(const LaunchSet& L, uint32_t passes, hipStream_t st) {
    k3u_keys<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
    launch_sort_range(L, 0u, 0u, passes, st);
    const uint32_t buf = passes & 1u;
    k3u_cell_table<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all, buf);
    if (L.nb_fluid) k3u_lists<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all, buf);
}

void launch_permute_only(const LaunchSet& L, uint32_t sbuf, hipStream_t st) {
    k_permute<false><<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all, sbuf);
    k_fluid_scan<<<L.nb_scene, kScanThreads, 0, st>>>(L.scenes, L.map_scene);
    k_compact<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
}

constexpr uint32_t kCodecGrid = 1024;  // blocks of a k_codec launch (each takes parts in turn)
void launch_codec(const LaunchSet& L, hipStream_t st, bool force) {
    const uint32_t f = force ? 1u : 0u;
    if (!L.nb_fluid) return;
    const uint32_t grid = std::min(L.nb_fluid, kCodecGrid);
    if (L.spec & kSpecDelta)
        k_codec<true><<<grid, kBlock, 0, st>>>(L.scenes, L.map_fluid, L.nb_fluid, f);
    else
        k_codec<false><<<grid, kBlock, 0, st>>>(L.scenes, L.map_fluid, L.nb_fluid, f);
}

}  // namespace fsa
