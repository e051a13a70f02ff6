// fsa_search.hip — neighborhood search on gfx950.
//
// Replaces CompressedNeighborhoodSearch::find_neighbors
// (CompressedNeighbors.cpp:168-240):
//   1. k_keys         cell = floor(x / r), compact Morton key      (:174-189, 37-56)
//   2. k_radix_*      stable LSD radix sort of (key, index)        (:191-201)
//   3. k_permute      gather every component into sorted order     (ParticleCollection::swap)
//   4. k_fluid_scan / k_compact   fluid rank -> combined index
//   5. k_cell_table   cell -> [start, end) of its run              (:203-225)
//   6. k_lists        27-cell scan in ascending key order, distance
//                     test dot(x_j - x_i) <= r^2, ascending lists  (:536-595)
// All kernels take a scene table and a block map, so one launch serves any
// number of scenes (batching) and a single scene alike.
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

// ---- 1. keys --------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_keys(const SceneDev* __restrict__ scenes,
                                                 const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t i = br.block * kBlock + threadIdx.x;
    if (i >= S.n) return;
    const uint32_t c = S.state->cur;
    const float4 p = S.pos[c][i];
    const int32_t cx = cell_of(p.x, S.search_radius);
    const int32_t cy = cell_of(p.y, S.search_radius);
    const int32_t cz = cell_of(p.z, S.search_radius);
    S.keys[0][i] = compact_key(cx, cy, cz, S.levels);
    S.vals[0][i] = i;
}

// ---- 2. radix sort -----------------------------------------------------------
// peers of this lane among `valid` lanes holding the same 8-bit digit
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) {
    uint64_t peers = valid;
#pragma unroll
    for (int b = 0; b < kRadixBits; b++) {
        uint64_t bal = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    return peers;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const SceneDev* __restrict__ scenes,
                                                             const BlockRef* __restrict__ map, uint32_t shift,
                                                             uint32_t src) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t tile = br.block;
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t* __restrict__ keys = S.keys[src];
    const uint32_t base = tile * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t i = base + r * kSortThreads + threadIdx.x;
        const bool valid = i < S.n;
        const uint32_t d = valid ? (keys[i] >> shift) & (kRadix - 1) : 0u;
        const uint64_t peers = match_digit(d, __ballot(valid));
        if (valid && (peers & lanemask_lt()) == 0) atomicAdd(&h[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    S.hist[threadIdx.x * S.ntiles + tile] = h[threadIdx.x];
}

// one block per (scene, digit): exclusive scan of the digit's tile counts
__global__ __launch_bounds__(kBlock) void k_radix_scan(const SceneDev* __restrict__ scenes,
                                                       const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t d = br.block;
    uint32_t* row = S.hist + d * S.ntiles;
    __shared__ uint32_t lw[kBlock / 64];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < S.ntiles; base += kBlock) {
        const uint32_t idx = base + threadIdx.x;
        const uint32_t v = idx < S.ntiles ? row[idx] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, lw, tot);
        if (idx < S.ntiles) row[idx] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) S.hist[kRadix * S.ntiles + d] = carry;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const SceneDev* __restrict__ scenes,
                                                                const BlockRef* __restrict__ map, uint32_t shift,
                                                                uint32_t src) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t tile = br.block;
    const uint32_t t = threadIdx.x;
    const uint32_t wave = t / 64;
    __shared__ uint32_t digit_base[kRadix];
    __shared__ uint32_t tile_start[kRadix];
    __shared__ uint32_t wave_cnt[kSortThreads / 64][kRadix];
    __shared__ uint32_t stage_k[kSortTile];
    __shared__ uint32_t stage_v[kSortTile];
    __shared__ uint32_t lw[kSortThreads / 64];

    // global base of each digit for this tile
    {
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(S.hist[kRadix * S.ntiles + t], lw, tot);
        digit_base[t] = ex + S.hist[t * S.ntiles + tile];
    }
#pragma unroll
    for (int w = 0; w < kSortThreads / 64; w++) wave_cnt[w][t] = 0;
    __syncthreads();

    const uint32_t* __restrict__ keys = S.keys[src];
    const uint32_t* __restrict__ vals = S.vals[src];
    const uint32_t base = tile * kSortTile;
    const uint32_t tile_n = min((uint32_t)kSortTile, S.n - base);
    uint32_t k[kSortItems], v[kSortItems], rank[kSortItems];
    // wave w owns items [w*512, (w+1)*512) of the tile, in rounds of 64 consecutive items
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t li = wave * (kSortTile / (kSortThreads / 64)) + r * 64 + __lane_id();
        const bool valid = li < tile_n;
        k[r] = valid ? keys[base + li] : 0u;
        v[r] = valid ? vals[base + li] : 0u;
        const uint32_t d = (k[r] >> shift) & (kRadix - 1);
        const uint64_t peers = match_digit(d, __ballot(valid));
        uint32_t before = valid ? wave_cnt[wave][d] : 0u;
        rank[r] = before + (uint32_t)__popcll(peers & lanemask_lt());
        if (valid && (peers & lanemask_lt()) == 0) wave_cnt[wave][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // per digit: exclusive prefix over waves, then exclusive scan over digits
    {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < kSortThreads / 64; w++) {
            const uint32_t c = wave_cnt[w][t];
            wave_cnt[w][t] = run;
            run += c;
        }
        uint32_t tot;
        tile_start[t] = block_exclusive_scan(run, lw, tot);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const uint32_t li = wave * (kSortTile / (kSortThreads / 64)) + r * 64 + __lane_id();
        if (li < tile_n) {
            const uint32_t d = (k[r] >> shift) & (kRadix - 1);
            const uint32_t pos = tile_start[d] + wave_cnt[wave][d] + rank[r];
            stage_k[pos] = k[r];
            stage_v[pos] = v[r];
        }
    }
    __syncthreads();
    uint32_t* __restrict__ okeys = S.keys[src ^ 1u];
    uint32_t* __restrict__ ovals = S.vals[src ^ 1u];
    for (uint32_t idx = t; idx < tile_n; idx += kSortThreads) {
        const uint32_t key = stage_k[idx];
        const uint32_t d = (key >> shift) & (kRadix - 1);
        const uint32_t g = digit_base[d] + (idx - tile_start[d]);
        okeys[g] = key;
        ovals[g] = stage_v[idx];
    }
}

// ---- 3. permute ------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_permute(const SceneDev* __restrict__ scenes,
                                                    const BlockRef* __restrict__ map, uint32_t sbuf) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t i = br.block * kBlock + threadIdx.x;
    const uint32_t c = S.state->cur;
    const uint32_t o = c ^ 1u;
    bool fluid = false;
    if (i < S.n) {
        const uint32_t s = S.vals[sbuf][i];
        S.pos[o][i] = S.pos[c][s];
        S.vel[o][i] = S.vel[c][s];
        S.acc[o][i] = S.acc[c][s];
        if (S.flags & kFlagExternal) S.ext[o][i] = S.ext[c][s];
        S.pres[o][i] = S.pres[c][s];
        S.err[o][i] = S.err[c][s];
        const uint32_t idt = S.idt[c][s];
        S.idt[o][i] = idt;
        fluid = (idt >> kTypeShift) == 0u;
    }
    const int cnt = __syncthreads_count(fluid);
    if (threadIdx.x == 0) S.block_count[br.block] = (uint32_t)cnt;
}

// one block per scene: exclusive scan of the per-block fluid counts
__global__ __launch_bounds__(kBlock) void k_fluid_scan(const SceneDev* __restrict__ scenes,
                                                       const BlockRef* __restrict__ map) {
    const SceneDev& S = scenes[map[blockIdx.x].scene];
    __shared__ uint32_t lw[kBlock / 64];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < S.nblocks; base += kBlock) {
        const uint32_t idx = base + threadIdx.x;
        const uint32_t v = idx < S.nblocks ? S.block_count[idx] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, lw, tot);
        if (idx < S.nblocks) S.block_count[idx] = carry + ex;
        carry += tot;
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
    if (fluid) S.fluid_idx[S.block_count[br.block] + ex] = i;
}

// ---- 5. cell table -----------------------------------------------------------
__device__ __forceinline__ bool cell_linear(const SceneDev& S, int32_t cx, int32_t cy, int32_t cz, uint32_t& lin) {
    const int32_t x = cx - S.cmin[0], y = cy - S.cmin[1], z = cz - S.cmin[2];
    if (x < 0 || y < 0 || z < 0 || x >= (int32_t)S.cdim[0] || y >= (int32_t)S.cdim[1] || z >= (int32_t)S.cdim[2])
        return false;
    lin = (uint32_t)x + S.cdim[0] * ((uint32_t)y + S.cdim[1] * (uint32_t)z);
    return true;
}

__global__ __launch_bounds__(kBlock) void k_cell_table(const SceneDev* __restrict__ scenes,
                                                       const BlockRef* __restrict__ map, uint32_t sbuf) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t i = br.block * kBlock + threadIdx.x;
    if (i >= S.n) return;
    const uint32_t c = S.state->cur;
    const uint32_t* __restrict__ keys = S.keys[sbuf];
    const uint32_t key = keys[i];
    const bool first = i == 0 || keys[i - 1] != key;
    const bool last = i + 1 == S.n || keys[i + 1] != key;
    if (!first && !last) return;
    const float4 p = S.pos[c][i];
    uint32_t lin;
    if (!cell_linear(S, cell_of(p.x, S.search_radius), cell_of(p.y, S.search_radius),
                     cell_of(p.z, S.search_radius), lin)) {
        atomicOr(&S.state->status, kStatusGrid);
        return;
    }
    if (first) {
        S.cells[lin].x = i;
        S.cells[lin].z = S.stamp;
    }
    if (last) S.cells[lin].y = i + 1;
}

// ---- 6. neighbor lists ---------------------------------------------------------
template <int N>
__device__ __forceinline__ void bitonic_sort(uint32_t (&a)[N]) {
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (int i = 0; i < N; i++) {
                const int l = i ^ j;
                if (l > i) {
                    const uint32_t x = a[i], y = a[l];
                    const bool up = (i & k) == 0;
                    a[i] = up ? min(x, y) : max(x, y);
                    a[l] = up ? max(x, y) : min(x, y);
                }
            }
        }
    }
}

// Distance from x to the face of its own cell in direction d (d = -1, 0, +1).
__device__ __forceinline__ float face_gap(float x, int32_t c, int d, float r) {
    const float lo = (float)c * r;
    return d < 0 ? x - lo : (d > 0 ? (lo + r) - x : 0.0f);
}

__global__ __launch_bounds__(kBlock) void k_lists(const SceneDev* __restrict__ scenes,
                                                  const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    __shared__ uint32_t s_run[27][kBlock];  // sorted runs of this thread, column per thread
    const uint32_t t = threadIdx.x;
    const uint32_t k = br.block * kBlock + t;
    if (k >= S.nf) return;
    const uint32_t c = S.state->cur;
    const uint32_t i = S.fluid_idx[k];
    const float4 p = S.pos[c][i];
    const float r = S.search_radius;
    const int32_t cx = cell_of(p.x, r);
    const int32_t cy = cell_of(p.y, r);
    const int32_t cz = cell_of(p.z, r);
    // cells whose box lies farther than r (with a rounding margin) hold no neighbor
    const float cull2 = S.r2 * 1.0001f;
    float gx[3], gy[3], gz[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const float ax = face_gap(p.x, cx, d - 1, r), ay = face_gap(p.y, cy, d - 1, r), az = face_gap(p.z, cz, d - 1, r);
        gx[d] = ax * ax;
        gy[d] = ay * ay;
        gz[d] = az * az;
    }
    uint32_t run[32];
#pragma unroll
    for (int s = 27; s < 32; s++) run[s] = kEmptyCell;
#pragma unroll
    for (int s = 0; s < 27; s++) {
        const int dx = s / 9, dy = (s / 3) % 3, dz = s % 3;
        run[s] = kEmptyCell;
        uint32_t lin;
        if (gx[dx] + gy[dy] + gz[dz] <= cull2 && cell_linear(S, cx + dx - 1, cy + dy - 1, cz + dz - 1, lin)) {
            const uint4 e = S.cells[lin];
            if (e.z == S.stamp) run[s] = (e.x << 5) | (uint32_t)s;
        }
    }
    bitonic_sort<32>(run);
#pragma unroll
    for (int q = 0; q < 27; q++) s_run[q][t] = run[q];
    const float r2 = S.r2;
    const uint32_t stride = S.list_stride;
    uint32_t cnt = 0;
    bool overflow = false;
    const float4* __restrict__ pos = S.pos[c];
    const uint32_t* __restrict__ idt = S.idt[c];
    uint32_t* __restrict__ out = S.nbr + k;
    for (int q = 0; q < 27; q++) {
        const uint32_t rv = s_run[q][t];
        if (rv == kEmptyCell) break;
        const uint32_t s = rv & 31u;
        uint32_t lin;
        cell_linear(S, cx + (int)(s / 9) - 1, cy + (int)((s / 3) % 3) - 1, cz + (int)(s % 3) - 1, lin);
        const uint32_t end = S.cells[lin].y;
        for (uint32_t j = rv >> 5; j < end; j++) {
            const float4 pj = pos[j];
            const float ddx = pj.x - p.x, ddy = pj.y - p.y, ddz = pj.z - p.z;
            if (dot3(ddx, ddy, ddz, ddx, ddy, ddz) <= r2) {
                if (cnt < (uint32_t)kMaxNeighbors) {
                    const uint32_t bnd = (idt[j] >> kTypeShift) == 1u ? kBoundaryBit : 0u;
                    out[cnt * stride] = j | bnd;
                    cnt++;
                } else {
                    overflow = true;
                }
            }
        }
    }
    S.nbr_cnt[k] = cnt;
    if (overflow) atomicOr(&S.state->status, kStatusCapacity);
}

// ---- launchers -------------------------------------------------------------------
void launch_keys(const LaunchSet& L, hipStream_t st) {
    if (L.nb_all) k_keys<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
}

void launch_sort(const LaunchSet& L, uint32_t passes, hipStream_t st) {
    for (uint32_t p = 0; p < passes; p++) {
        const uint32_t src = p & 1u;
        const uint32_t shift = p * kRadixBits;
        k_radix_hist<<<L.nb_tiles, kSortThreads, 0, st>>>(L.scenes, L.map_tiles, shift, src);
        k_radix_scan<<<L.nb_scan, kBlock, 0, st>>>(L.scenes, L.map_scan);
        k_radix_scatter<<<L.nb_tiles, kSortThreads, 0, st>>>(L.scenes, L.map_tiles, shift, src);
    }
}

void launch_reorder(const LaunchSet& L, uint32_t sbuf, hipStream_t st) {
    k_permute<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all, sbuf);
    k_fluid_scan<<<L.nb_scene, kBlock, 0, st>>>(L.scenes, L.map_scene);
    k_compact<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
    k_cell_table<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all, sbuf);
}

void launch_lists(const LaunchSet& L, hipStream_t st) {
    if (L.nb_fluid) k_lists<<<L.nb_fluid, kBlock, 0, st>>>(L.scenes, L.map_fluid);
}

}  // namespace fsa
