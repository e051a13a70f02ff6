// fsa_kernels.hpp — host-side launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fsa_device.hpp"

namespace fsa {

// neighbor-walk variants the autotuner times per kernel family: the unroll in
// the low nibble, an occupancy hint (at least that many waves per SIMD) in the high one
constexpr uint8_t kVariantU2 = 0x02, kVariantU4 = 0x04, kVariantU8 = 0x08;
constexpr uint8_t kVariantU4W8 = 0x84, kVariantU8W6 = 0x68;
constexpr int kTuneVariants = 5;
constexpr uint8_t kTuneVariant[kTuneVariants] = {kVariantU2, kVariantU4, kVariantU8, kVariantU4W8, kVariantU8W6};

// Device-resident launch description of a set of scenes: the scene table and
// one block map per launch domain.
struct LaunchSet {
    SceneDev* scenes = nullptr;   // device table
    BlockRef* map_all = nullptr;  // kBlock-blocks over each scene's combined particles
    BlockRef* map_fluid = nullptr;   // kBlock-blocks over each scene's fluid particles
    BlockRef* map_tiles = nullptr;   // radix tiles
    BlockRef* map_scan = nullptr;    // (scene, digit)
    BlockRef* map_scene = nullptr;   // one block per scene
    BlockRef* map_cnt = nullptr;     // kCountBlocks blocks per scene over its rank range
    uint32_t nb_all = 0, nb_fluid = 0, nb_tiles = 0, nb_scan = 0, nb_scene = 0, nb_cnt = 0;
    uint32_t spec = 0;  // kSpec* bits shared by every scene of the set
    // neighbor-walk variant (kVariant*) per IISPH kernel family (kTune*), chosen by the autotuner
    uint8_t unroll[4] = {kVariantU4, kVariantU4, kVariantU4, kVariantU4};
    bool codec = false;  // some scene of the set checks the NeighborStorage codec each search
    bool wide_lists = false;  // some scene's neighbor-list array spans 4 GiB or more
    bool entities_before = false;  // some scene has a boundary preprocessor or a velocity alteration
    bool remover = false;          // some scene has a particle remover
    bool counting = false;         // compressed search: the counting sort over rank keys (every scene has them)
    bool skip_first_check = false; // 3D IISPH: every scene iterates at least twice (no check after the first iteration)
    bool fluidless_scene = false;  // some scene has no fluid particle (no phase-1 block sets its loop state)
};

// kernel families the autotuner times
constexpr int kTunePrepare = 0;
constexpr int kTuneDiagonal = 1;
constexpr int kTunePassA = 2;
constexpr int kTunePassB = 3;
constexpr int kTuneFamilies = 4;

// simulation entities and sensors
void launch_entities_before(const LaunchSet& L, hipStream_t st);
void launch_remover_check(const LaunchSet& L, hipStream_t st);
void launch_query_count(const SceneDev* d_scene, const float4* q, uint32_t nq, uint32_t* counts, hipStream_t st);
void launch_query_fill(const SceneDev* d_scene, const float4* q, uint32_t nq, const uint64_t* offsets, uint64_t* out,
                       hipStream_t st);
// SensorPlane sample geometry: ax = normalize(span_x) * x_step, ay likewise
struct PlaneParams {
    float origin[3], ax[3], ay[3];
    uint32_t nx, ny, sub, type;
    float sub_den;  // sub_sample_grid_size + 1.0f
    float div;      // Math::pow2(sub_sample_grid_size + 1)
};
void launch_sensor_plane(const SceneDev* d_scene, const PlaneParams& P, float* out, hipStream_t st);
void launch_volume(const SceneDev* d_scene, const float4* q, uint32_t nq, float* out, hipStream_t st);
constexpr uint64_t kStatAccBytes = 80;  // one block's partial of the global sensors
uint64_t statistics_scratch_bytes(uint32_t nblocks);
void launch_statistics(const SceneDev* d_scene, uint32_t nblocks, float zero_height, void* scratch, hipStream_t st);
void read_statistics(const void* host_copy, uint64_t counts[3], float mx[3], float mn[3], double sums[5]);
uint64_t statistics_result_bytes();  // one accumulator; the result sits at the end of the scratch

// neighborhood search
void launch_keys_sort(const LaunchSet& L, uint32_t passes, hipStream_t st);
// the counting sort over rank keys: sorted (rank, row) pairs in keys[1] / vals[1]
void launch_count_sort(const LaunchSet& L, hipStream_t st);
// the rank table of a grid of cdim cells from cmin (box_rank of every cell)
void launch_rank_table(uint32_t* ranks, const int32_t cmin[3], const uint32_t cdim[3], uint32_t levels,
                       hipStream_t st);
void launch_reorder(const LaunchSet& L, uint32_t sbuf, hipStream_t st);
void launch_lists(const LaunchSet& L, uint32_t sbuf, hipStream_t st);
// hashed / quadratic 3D searches over the kept collection order
void launch3u_search(const LaunchSet& L, uint32_t passes, hipStream_t st);
// gather every particle array into the order vals[sbuf] gives, flip the buffers,
// rebuild the fluid rank maps
void launch_permute_only(const LaunchSet& L, uint32_t sbuf, hipStream_t st);
// NeighborStorage sizes and capacity of the lists just built: every scene's with
// `force`, otherwise those whose lists k_lists could not bound within the payload
// (and delta-coded ones, which are encoded here)
void launch_codec(const LaunchSet& L, hipStream_t st, bool force);
// stable LSD radix passes over (keys, vals)[first_src], digits from bit shift0 upward
void launch_sort_range(const LaunchSet& L, uint32_t first_src, uint32_t shift0, uint32_t passes, hipStream_t st);

// the hashed-3D search's re-sort (Sort::quick_sort) on the device: parts of more
// than kQsSmall elements are partitioned level by level in chunks of kQsChunk
// (each level's split writing the next level's tables), smaller parts sorted one
// per lane at the end
constexpr uint32_t kQsChunk = 2048;
constexpr uint32_t kQsSmall = 64;
struct QsSeg {
    uint32_t lo, hi;  // the part [lo, hi]
    uint32_t c0, c1;  // its chunks in the chunk table
    uint32_t nA, nB;  // keys >= pivot, keys <= pivot
    uint32_t K, J;    // swaps, returned index
};
struct QsChunk {
    uint32_t seg, lo, hi;  // part and [lo, hi] of the chunk
    uint32_t cA, cB;       // counts of keys >= pivot and <= pivot
    uint32_t offA, offB;   // A items in the part's chunks before it, B items in those after it
    uint32_t pad;
};
// counters: [0], [1] the parts and chunks of the level tables 0 and 1; [2] the
// small parts and an error flag (a table would overflow: internal error)
struct QsCtr {
    uint32_t nseg, nchunk;
};
struct QsBuffers {
    uint64_t* key;
    uint32_t* pay;   // old collection position of each element
    uint32_t* apos;  // [lo + t]: position of the part's t-th key >= pivot from the left
    uint32_t* bpos;  // [lo + t]: position of the part's t-th key <= pivot from the right
    QsSeg* seg[2];   // level tables, alternating
    QsChunk* chunk[2];
    uint2* small;    // parts of at most kQsSmall elements: {lo, hi}
    QsCtr* ctr;      // 3 entries
    uint32_t seg_cap, chunk_cap, small_cap;
};
// keys of every collection position (rows in collection order ref[0..rows), the
// resident rows' cells from the last grid update), old positions as payload
void launch_qs_keys(const uint32_t* idt, uint32_t n, const int4* cell_last, const uint32_t* ref, uint32_t rows,
                    uint32_t* inv, uint64_t* key, uint32_t* pay, hipStream_t st);
// one level: the parts of table cur are partitioned, their children go to table cur ^ 1 or the small list
void launch_qs_level(const QsBuffers& B, uint32_t cur, hipStream_t st);
void launch_qs_small(const QsBuffers& B, uint32_t nsmall, hipStream_t st);
// the new collection order (ref_out) and the permute's source list of the resident rows
void launch_qs_apply(const QsBuffers& B, const uint32_t* ref, uint32_t rows, const uint32_t* inv, uint32_t* ref_out,
                     uint32_t* block_cnt, uint32_t* src, hipStream_t st);

// IISPH / SESPH
void launch_iisph_prepare(const LaunchSet& L, bool boundary_pass, hipStream_t st);
void launch_iisph_diagonal(const LaunchSet& L, hipStream_t st);
// the Jacobi loop state of scenes without fluid (phase 1 sets the others')
void launch_iisph_begin_loop(const LaunchSet& L, hipStream_t st);
void launch_iisph_pass_a(const LaunchSet& L, hipStream_t st);
void launch_iisph_first_converge(const LaunchSet& L, hipStream_t st);
void launch_iisph_pass_b(const LaunchSet& L, hipStream_t st);
void launch_sesph_density(const LaunchSet& L, hipStream_t st);
void launch_sesph_accel(const LaunchSet& L, hipStream_t st);
void launch_integrate(const LaunchSet& L, hipStream_t st);
// one kernel of a tuned family with the set's current unroll choice
void launch_tuned_kernel(const LaunchSet& L, int family, hipStream_t st);

// 2D (hashed / quadratic-dynamic-allocated searches, IISPH and SESPH)
void launch2_search(const LaunchSet& L, uint32_t entry_passes, uint32_t cell_passes, hipStream_t st);
void launch2_iisph_prepare(const LaunchSet& L, hipStream_t st);
void launch2_iisph_diagonal(const LaunchSet& L, hipStream_t st);
void launch2_iisph_begin_loop(const LaunchSet& L, hipStream_t st);
void launch2_iisph_pass_a(const LaunchSet& L, hipStream_t st);
void launch2_iisph_pass_b(const LaunchSet& L, hipStream_t st);
void launch2_sesph(const LaunchSet& L, hipStream_t st);
void launch2_integrate(const LaunchSet& L, hipStream_t st);

}  // namespace fsa
