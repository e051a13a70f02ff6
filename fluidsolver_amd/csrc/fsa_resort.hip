// fsa_resort.hip — the hashed-3D search's re-sort on gfx950.
//
// HashedNeighborhoodSearch3D::improve_cache_efficiency (HashedNeighborhoodSearch3D.cpp:
// 219-247) re-sorts the collection, inactive particles included, with
// Sort::quick_sort (ParticleCollectionAlgorithm.hpp:28-127) on the Morton key of
// each particle's cell from the last grid update.  The quick sort is unstable,
// so its permutation, not merely a sorted order, is what the reference's
// collection holds afterwards; this file computes exactly that permutation.
//
// One Hoare partition of [lo, hi] around P = key[lo + (hi - lo) / 2] swaps, for
// t = 0, 1, ..., the t-th position from the left holding a key >= P (A[t]) with
// the t-th position from the right holding a key <= P (B[t]), while A[t] < B[t].
// A scan position never passes a swapped one before the two scans cross, so the
// swaps are exactly the first K rank pairs, K = #{t : A[t] < B[t]} (A rises, B
// falls, so the pairs with A[t] < B[t] come first), and the partition returns
// J = max(B[K], A[K - 1]) (B[K] alone when K = 0, A[K - 1] alone when B has
// only K entries).  Ranks are prefix counts, so a partition of any size runs in
// parallel:
//   k_qs_count   per chunk of a part: the counts of keys >= P and <= P
//   k_qs_offsets per part: the chunks' offsets (A from the left, B from the right)
//   k_qs_ranks   per chunk: A[rank] and B[rank] for every element
//   k_qs_split   per part: K (a binary search, the predicate is monotone) and J
//   k_qs_swap    per chunk of ranks: the K swaps, in place (disjoint pairs)
// Parts split level by level until they hold at most kQsSmall elements.  Each
// level's split writes the next level's part and chunk tables itself (atomic
// appends: the parts are disjoint, so their order in a table does not matter,
// and a part's chunks take one contiguous range), so the host only checks every
// few levels whether any part is left.  The small parts are then sorted by the
// reference's own sequential partition, one part per lane (disjoint parts do not
// interact, so the order in which parts are done does not change the permutation).
#include <algorithm>

#include "fsa_device.hpp"
#include "fsa_kernels.hpp"

namespace fsa {

namespace {

constexpr uint32_t kQsThreads = 256;
constexpr uint32_t kQsPerThread = kQsChunk / kQsThreads;
static_assert(kQsChunk % kQsThreads == 0, "a chunk is a whole number of items per thread");

// exclusive prefix of one value per thread over the block; total = the block's sum
__device__ __forceinline__ uint32_t qs_block_scan(uint32_t v, uint32_t* lds_waves, uint32_t& total) {
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) lds_waves[wave] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; w++) {
        const uint32_t c = lds_waves[w];
        if (w < wave) off += c;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return off + x - v;
}

__global__ __launch_bounds__(kQsThreads) void k_qs_inv_fill(uint32_t* __restrict__ inv, uint32_t rows) {
    const uint64_t r = (uint64_t)blockIdx.x * kQsThreads + threadIdx.x;
    if (r < rows) inv[r] = 0xffffffffu;
}

// device index of each resident row
__global__ __launch_bounds__(kQsThreads) void k_qs_inv(const uint32_t* __restrict__ idt, uint32_t n,
                                                       uint32_t* __restrict__ inv) {
    const uint32_t k = blockIdx.x * kQsThreads + threadIdx.x;
    if (k < n) inv[idt[k] & kIdMask] = k;
}

// key of collection position p: its row's cell from the last grid update, or
// the key of the never-updated cell (INT_MIN, INT_MIN, INT_MIN) for inactive rows
__global__ __launch_bounds__(kQsThreads) void k_qs_keys(const uint32_t* __restrict__ ref, const uint32_t* __restrict__ inv,
                                                        const int4* __restrict__ cell_last, uint32_t rows,
                                                        uint64_t* __restrict__ key, uint32_t* __restrict__ pay) {
    const uint32_t p = blockIdx.x * kQsThreads + threadIdx.x;
    if (p >= rows) return;
    const uint32_t d = inv[ref[p]];
    if (d == 0xffffffffu) {
        key[p] = reference_cell_key(INT32_MIN, INT32_MIN, INT32_MIN);
    } else {
        const int4 c = cell_last[d];
        key[p] = reference_cell_key(c.x, c.y, c.z);
    }
    pay[p] = p;
}

__device__ __forceinline__ uint64_t qs_pivot(const uint64_t* key, const QsSeg& s) {
    return key[s.lo + (s.hi - s.lo) / 2];
}

__global__ __launch_bounds__(kQsThreads) void k_qs_count(QsBuffers B, uint32_t cur) {
    // the next level's table starts empty (the level before used it last)
    if (blockIdx.x == 0 && threadIdx.x == 0) B.ctr[cur ^ 1u] = QsCtr{0u, 0u};
    const uint32_t nchunk = B.ctr[cur].nchunk;
    __shared__ uint32_t lw[kQsThreads / 64];
    for (uint32_t ci = blockIdx.x; ci < nchunk; ci += gridDim.x) {
        QsChunk& c = B.chunk[cur][ci];
        const QsSeg s = B.seg[cur][c.seg];
        const uint64_t P = qs_pivot(B.key, s);
        uint32_t a = 0, b = 0;
        for (uint32_t x = c.lo + threadIdx.x; x <= c.hi; x += kQsThreads) {
            const uint64_t k = B.key[x];
            a += k >= P ? 1u : 0u;
            b += k <= P ? 1u : 0u;
        }
        uint32_t ta, tb;
        (void)qs_block_scan(a, lw, ta);
        (void)qs_block_scan(b, lw, tb);
        if (threadIdx.x == 0) {
            c.cA = ta;
            c.cB = tb;
        }
    }
}

// per part: A offsets of its chunks from the left, B offsets from the right
__global__ __launch_bounds__(kQsThreads) void k_qs_offsets(QsBuffers B, uint32_t cur) {
    const uint32_t nseg = B.ctr[cur].nseg;
    __shared__ uint32_t lw[kQsThreads / 64];
    for (uint32_t si = blockIdx.x; si < nseg; si += gridDim.x) {
        QsSeg& s = B.seg[cur][si];
        QsChunk* const ch = B.chunk[cur];
        uint32_t carry_a = 0;
        for (uint32_t c0 = s.c0; c0 < s.c1; c0 += kQsThreads) {
            const uint32_t c = c0 + threadIdx.x;
            const uint32_t v = c < s.c1 ? ch[c].cA : 0u;
            uint32_t tot;
            const uint32_t ex = qs_block_scan(v, lw, tot);
            if (c < s.c1) ch[c].offA = carry_a + ex;
            carry_a += tot;
        }
        uint32_t carry_b = 0;  // B elements in the chunks right of the current window
        const uint32_t nc = s.c1 - s.c0;
        for (uint32_t r0 = 0; r0 < nc; r0 += kQsThreads) {
            // a window of chunks taken from the right: thread t holds chunk c1 - 1 - (r0 + t)
            const uint32_t r = r0 + threadIdx.x;
            const uint32_t c = s.c1 - 1u - r;
            const uint32_t v = r < nc ? ch[c].cB : 0u;
            uint32_t tot;
            const uint32_t ex = qs_block_scan(v, lw, tot);
            if (r < nc) ch[c].offB = carry_b + ex;
            carry_b += tot;
        }
        if (threadIdx.x == 0) {
            s.nA = carry_a;
            s.nB = carry_b;
        }
    }
}

// A[rank] and B[rank] of every element of each chunk
__global__ __launch_bounds__(kQsThreads) void k_qs_ranks(QsBuffers B, uint32_t cur) {
    const uint32_t nchunk = B.ctr[cur].nchunk;
    __shared__ uint32_t lw[kQsThreads / 64];
    for (uint32_t ci = blockIdx.x; ci < nchunk; ci += gridDim.x) {
        const QsChunk c = B.chunk[cur][ci];
        const QsSeg s = B.seg[cur][c.seg];
        const uint64_t P = qs_pivot(B.key, s);
        const uint32_t x0 = c.lo + threadIdx.x * kQsPerThread;
        uint32_t fa = 0, fb = 0;  // flags of the thread's items
        uint32_t na = 0, nb = 0;
#pragma unroll
        for (uint32_t q = 0; q < kQsPerThread; q++) {
            const uint32_t x = x0 + q;
            if (x <= c.hi) {
                const uint64_t k = B.key[x];
                if (k >= P) {
                    fa |= 1u << q;
                    na++;
                }
                if (k <= P) {
                    fb |= 1u << q;
                    nb++;
                }
            }
        }
        uint32_t ta, tb;
        uint32_t ra = c.offA + qs_block_scan(na, lw, ta);
        // B ranks count from the right: the B items after this thread's in the
        // chunk, plus the chunk's offset
        const uint32_t before_b = qs_block_scan(nb, lw, tb);
        uint32_t rb = c.offB + (tb - before_b - nb);
#pragma unroll
        for (uint32_t q = 0; q < kQsPerThread; q++)
            if (fa & (1u << q)) B.apos[s.lo + ra++] = x0 + q;
        // this thread's B items from its last to its first
#pragma unroll
        for (int q = (int)kQsPerThread - 1; q >= 0; q--)
            if (fb & (1u << q)) B.bpos[s.lo + rb++] = x0 + (uint32_t)q;
    }
}

// a child part of the split: into the next level's tables, or the small list
__device__ __forceinline__ void qs_child(const QsBuffers& B, uint32_t nxt, uint32_t lo, uint32_t hi) {
    if (lo >= hi) return;
    if (hi - lo + 1u <= kQsSmall) {
        const uint32_t q = atomicAdd(&B.ctr[2].nseg, 1u);
        if (q < B.small_cap)
            B.small[q] = make_uint2(lo, hi);
        else
            atomicOr(&B.ctr[2].nchunk, 1u);
        return;
    }
    const uint32_t nch = (hi - lo) / kQsChunk + 1u;
    const uint32_t si = atomicAdd(&B.ctr[nxt].nseg, 1u);
    const uint32_t c0 = atomicAdd(&B.ctr[nxt].nchunk, nch);
    if (si >= B.seg_cap || c0 + nch > B.chunk_cap) {
        atomicOr(&B.ctr[2].nchunk, 1u);
        return;
    }
    B.seg[nxt][si] = QsSeg{lo, hi, c0, c0 + nch, 0u, 0u, 0u, 0u};
    for (uint32_t q = 0; q < nch; q++) {
        const uint32_t clo = lo + q * kQsChunk;
        B.chunk[nxt][c0 + q] = QsChunk{si, clo, min(hi, clo + kQsChunk - 1u), 0u, 0u, 0u, 0u, 0u};
    }
}

// per part (one thread each): the number of swaps K and the returned index J;
// the two halves become parts of the next level (or small parts)
__global__ __launch_bounds__(64) void k_qs_split(QsBuffers B, uint32_t cur) {
    const uint32_t nseg = B.ctr[cur].nseg;
    for (uint32_t i = blockIdx.x * 64 + threadIdx.x; i < nseg; i += gridDim.x * 64) {
        QsSeg& s = B.seg[cur][i];
        const uint32_t* A = B.apos + s.lo;
        const uint32_t* Bp = B.bpos + s.lo;
        // first t in [0, m) with A[t] >= B[t], or m
        uint32_t lo = 0, hi = min(s.nA, s.nB);
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (A[mid] < Bp[mid])
                lo = mid + 1;
            else
                hi = mid;
        }
        const uint32_t K = lo;
        uint32_t J;
        if (K == 0)
            J = Bp[0];  // the pivot's position is in B
        else if (K < s.nB)
            J = max(Bp[K], A[K - 1]);
        else
            J = A[K - 1];
        s.K = K;
        s.J = J;
        if (J < s.lo || J >= s.hi) {
            // a Hoare partition of [lo, hi] returns J in [lo, hi - 1] (internal error)
            atomicOr(&B.ctr[2].nchunk, 1u);
            s.K = 0;
            continue;
        }
        qs_child(B, cur ^ 1u, s.lo, J);
        qs_child(B, cur ^ 1u, J + 1u, s.hi);
    }
}

// the first K rank pairs of each part swap their elements
__global__ __launch_bounds__(kQsThreads) void k_qs_swap(QsBuffers B, uint32_t cur) {
    const uint32_t nchunk = B.ctr[cur].nchunk;
    for (uint32_t ci = blockIdx.x; ci < nchunk; ci += gridDim.x) {
        const QsChunk c = B.chunk[cur][ci];
        const QsSeg s = B.seg[cur][c.seg];
        for (uint32_t x = c.lo + threadIdx.x; x <= c.hi; x += kQsThreads) {
            const uint32_t t = x - s.lo;
            if (t >= s.K) break;
            const uint32_t pa = B.apos[s.lo + t], pb = B.bpos[s.lo + t];
            const uint64_t ka = B.key[pa], kb = B.key[pb];
            const uint32_t va = B.pay[pa], vb = B.pay[pb];
            B.key[pa] = kb;
            B.key[pb] = ka;
            B.pay[pa] = vb;
            B.pay[pb] = va;
        }
    }
}

// parts of at most kQsSmall elements, one per lane: the reference's sequential
// quick sort (partition around the middle key, the smaller half first, the
// larger one waiting on the lane's stack) in LDS, element i of lane l at i * 64 + l
__global__ __launch_bounds__(64) void k_qs_small(QsBuffers B, uint32_t nsmall) {
    __shared__ uint64_t key[kQsSmall * 64];
    __shared__ uint32_t pay[kQsSmall * 64];
    __shared__ uint32_t stack[8 * 64];  // lo | hi << 16; at most log2(kQsSmall) + 1 parts wait
    static_assert(kQsSmall <= 128, "a small part's stack fits 8 entries");
    const uint32_t l = threadIdx.x;
    const uint32_t p = blockIdx.x * 64 + l;
    if (p >= nsmall) return;
    const uint2 part = B.small[p];
    const uint32_t base = part.x;
    const int32_t n = (int32_t)(part.y - part.x + 1u);
    for (int32_t q = 0; q < n; q++) {
        key[q * 64 + l] = B.key[base + q];
        pay[q * 64 + l] = B.pay[base + q];
    }
    uint32_t sp = 0;
    int32_t lo = 0, hi = n - 1;
    for (;;) {
        if (lo >= hi) {
            if (sp == 0) break;
            sp--;
            const uint32_t e = stack[sp * 64 + l];
            lo = (int32_t)(e & 0xffffu);
            hi = (int32_t)(e >> 16);
            continue;
        }
        const uint64_t P = key[(lo + (hi - lo) / 2) * 64 + l];
        int32_t i = lo - 1, j = hi + 1;
        for (;;) {
            do {
                i++;
            } while (key[i * 64 + l] < P);
            do {
                j--;
            } while (key[j * 64 + l] > P);
            if (i >= j) break;
            const uint64_t tk = key[i * 64 + l];
            key[i * 64 + l] = key[j * 64 + l];
            key[j * 64 + l] = tk;
            const uint32_t tv = pay[i * 64 + l];
            pay[i * 64 + l] = pay[j * 64 + l];
            pay[j * 64 + l] = tv;
        }
        if (j - lo < hi - j) {
            stack[sp * 64 + l] = (uint32_t)(j + 1) | (uint32_t)hi << 16;
            hi = j;
        } else {
            stack[sp * 64 + l] = (uint32_t)lo | (uint32_t)j << 16;
            lo = j + 1;
        }
        sp++;
    }
    for (int32_t q = 0; q < n; q++) {
        B.key[base + q] = key[q * 64 + l];
        B.pay[base + q] = pay[q * 64 + l];
    }
}

// the collection order after the sort: position p takes the row at the old
// position pay[p]; rows still resident are counted per block for the compaction
__global__ __launch_bounds__(kQsThreads) void k_qs_apply(const uint32_t* __restrict__ ref, const uint32_t* __restrict__ pay,
                                                         const uint32_t* __restrict__ inv, uint32_t rows,
                                                         uint32_t* __restrict__ ref_out, uint32_t* __restrict__ block_cnt) {
    const uint32_t p = blockIdx.x * kQsThreads + threadIdx.x;
    bool act = false;
    if (p < rows) {
        const uint32_t r = ref[pay[p]];
        ref_out[p] = r;
        act = inv[r] != 0xffffffffu;
    }
    const int c = __syncthreads_count(act);
    if (threadIdx.x == 0) block_cnt[blockIdx.x] = (uint32_t)c;
}

// exclusive scan of the block counts, one block
__global__ __launch_bounds__(1024) void k_qs_scan_counts(uint32_t* __restrict__ block_cnt, uint32_t nblocks) {
    __shared__ uint32_t lw[1024 / 64];
    const uint32_t per = (nblocks + 1023) / 1024;
    const uint32_t b0 = min(nblocks, threadIdx.x * per), b1 = min(nblocks, b0 + per);
    uint32_t sum = 0;
    for (uint32_t b = b0; b < b1; b++) sum += block_cnt[b];
    uint32_t tot;
    uint32_t run = qs_block_scan(sum, lw, tot);
    for (uint32_t b = b0; b < b1; b++) {
        const uint32_t v = block_cnt[b];
        block_cnt[b] = run;
        run += v;
    }
}

// the device index of every resident row in the new order: the source list of the permute
__global__ __launch_bounds__(kQsThreads) void k_qs_sources(const uint32_t* __restrict__ ref_out,
                                                           const uint32_t* __restrict__ inv, uint32_t rows,
                                                           const uint32_t* __restrict__ block_off, uint32_t* __restrict__ src) {
    const uint32_t p = blockIdx.x * kQsThreads + threadIdx.x;
    uint32_t d = 0xffffffffu;
    if (p < rows) d = inv[ref_out[p]];
    const bool act = d != 0xffffffffu;
    __shared__ uint32_t lw[kQsThreads / 64];
    uint32_t tot;
    const uint32_t ex = qs_block_scan(act ? 1u : 0u, lw, tot);
    if (act) src[block_off[blockIdx.x] + ex] = d;
}

}  // namespace

void launch_qs_keys(const uint32_t* idt, uint32_t n, const int4* cell_last, const uint32_t* ref, uint32_t rows,
                    uint32_t* inv, uint64_t* key, uint32_t* pay, hipStream_t st) {
    const uint32_t gr = (rows + kQsThreads - 1) / kQsThreads, gn = (n + kQsThreads - 1) / kQsThreads;
    if (gr) k_qs_inv_fill<<<gr, kQsThreads, 0, st>>>(inv, rows);
    if (gn) k_qs_inv<<<gn, kQsThreads, 0, st>>>(idt, n, inv);
    if (gr) k_qs_keys<<<gr, kQsThreads, 0, st>>>(ref, inv, cell_last, rows, key, pay);
}

// Grids of a level are fixed (the counts live on the device); blocks loop over
// the level's chunks and parts.
void launch_qs_level(const QsBuffers& B, uint32_t cur, hipStream_t st) {
    const uint32_t gc = std::max(1u, std::min(B.chunk_cap, 2048u));
    const uint32_t gs = std::max(1u, std::min(B.seg_cap, 1024u));
    k_qs_count<<<gc, kQsThreads, 0, st>>>(B, cur);
    k_qs_offsets<<<gs, kQsThreads, 0, st>>>(B, cur);
    k_qs_ranks<<<gc, kQsThreads, 0, st>>>(B, cur);
    k_qs_split<<<std::max(1u, std::min((B.seg_cap + 63) / 64, 512u)), 64, 0, st>>>(B, cur);
    k_qs_swap<<<gc, kQsThreads, 0, st>>>(B, cur);
}

void launch_qs_small(const QsBuffers& B, uint32_t nsmall, hipStream_t st) {
    if (nsmall) k_qs_small<<<(nsmall + 63) / 64, 64, 0, st>>>(B, nsmall);
}

void launch_qs_apply(const QsBuffers& B, const uint32_t* ref, uint32_t rows, const uint32_t* inv, uint32_t* ref_out,
                     uint32_t* block_cnt, uint32_t* src, hipStream_t st) {
    const uint32_t gr = (rows + kQsThreads - 1) / kQsThreads;
    if (!gr) return;
    k_qs_apply<<<gr, kQsThreads, 0, st>>>(ref, B.pay, inv, rows, ref_out, block_cnt);
    k_qs_scan_counts<<<1, 1024, 0, st>>>(block_cnt, gr);
    k_qs_sources<<<gr, kQsThreads, 0, st>>>(ref_out, inv, rows, block_cnt, src);
}

}  // namespace fsa
