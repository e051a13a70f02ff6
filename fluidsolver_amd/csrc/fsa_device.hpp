// fsa_device.hpp — device-side data layout and SPH math of the gfx950 hot path.
//
// Every scene lives in HBM as one combined particle set (fluid + boundary,
// inactive particles excluded) kept in the order libFluid's compressed search
// gives the collection: a stable sort by Morton cell key every step
// (CompressedNeighbors.cpp:191-201).  Because the order, the neighbor order
// (ascending combined index) and the FP32 evaluation order below all follow
// the reference, device results are bit-identical to the reference-order CPU
// restatement.  The translation unit must be compiled with -ffp-contract=off
// (no fused multiply-add) and IEEE-correct division and square root.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsa {

constexpr int kBlock = 256;            // threads per block of the particle kernels
constexpr int kMaxNeighbors = 97;      // NeighborStorage capacity (CompressedNeighbors.hpp:29-34)
constexpr uint32_t kBoundaryBit = 0x80000000u;  // neighbor entry flag: neighbor is a boundary particle
constexpr uint32_t kInactiveBit = 0x40000000u;  // neighbor entry flag (2D searches): neighbor is inactive
constexpr uint32_t kIndexMask = 0x3fffffffu;    // neighbor entry: particle index
constexpr uint32_t kTypeShift = 30;    // idt word: caller index (30 bits) | type << 30
constexpr uint32_t kIdMask = (1u << kTypeShift) - 1u;

// radix sort geometry
constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;  // 2048 keys per tile
constexpr int kRadixBits = 9;  // 512 digits: the 2D and hashed-3D searches' sorts
constexpr int kRadix = 1 << kRadixBits;
// The compressed search sorts 10-bit digits: the cell rank keys of grids up to
// 2^20 cells in 2 passes, compact keys up to 30 bits in 3.
constexpr int kSearchSortBits = 10;
constexpr int kRadixMaxDigits = 1 << kSearchSortBits;  // tile counts the histogram buffer holds per tile
constexpr uint32_t kCountBlocks = 1024;  // blocks per scene over its rank range (counting sort scans)

// SceneDev.flags
constexpr uint32_t kFlagHOne = 1u;          // kernel h == 1: q = |x|, pre = x / |x| exactly
constexpr uint32_t kFlagRho0SqOne = 2u;     // rest_density^2 == 1: p / rho0^2 == p exactly
constexpr uint32_t kFlagExternal = 4u;      // some particle carries a non-zero external acceleration
constexpr uint32_t kFlagDynamicBoundary = 8u;  // boundary particles move (non-zero v, a or ext)
constexpr uint32_t kFlagSingleLayer = 16u;  // IISPH/SESPH single-layer boundary scaling
constexpr uint32_t kFlagSesph = 32u;        // SESPH solver (else IISPH)
constexpr uint32_t kFlagPlainLists = 64u;   // uncompressed 32-bit neighbor lists
constexpr uint32_t kFlagIntAbs = 128u;      // 2D IISPH: integer abs() in the loop condition
constexpr uint32_t kFlagQuadratic = 256u;   // 2D: QuadraticNeighborhoodSearchDynamicAllocated order
constexpr uint32_t kFlagCodec = 512u;       // enforce the NeighborStorage codec capacity each search
constexpr uint32_t kFlagUnsorted = 1024u;   // 3D hashed / quadratic searches: the collection keeps its order
constexpr uint32_t kFlagTinyCoords = 2048u; // some coordinate lies in (0, 2^-36): pair math takes the IEEE path
constexpr uint32_t kFlagDeltaLists = 4096u; // 3D compressed search: the walks read the lists as 16-bit delta codes

// Delta-coded neighbor lists (3D compressed search, kFlagDeltaLists).  Each
// list is re-encoded after the search from its 32-bit entries:
//   code rows: entry e of fluid rank k has a 16-bit code in dword row e / 2
//     (low half for even e) at row * list_stride + k, so a wave reads the
//     codes of two entries of 64 consecutive ranks in one 256-byte load;
//     code = delta to the previous entry's index (1..32767) | boundary << 15,
//     or 0: the entry is the next one of the escape rows (entry 0 always is);
//   escape rows: the x-th escaped entry of rank k (its full 32-bit entry) at
//     (kDeltaCodeRows + x) * list_stride + k.  Entry 0 carries kDeltaManyBit
//     when the list has more than two escapes (the walks keep two in
//     registers and load the others as they are reached).
// A search whose lists need more than kDeltaEscRows escapes somewhere leaves
// SceneState.delta_plain set, and that step's walks read the 32-bit lists.
constexpr uint32_t kDeltaCodeRows = (kMaxNeighbors + 1) / 2;  // 49 rows: entries 0..96
constexpr uint32_t kDeltaEscRows = 15;
constexpr uint32_t kDeltaRows = kDeltaCodeRows + kDeltaEscRows;  // 64 dwords per fluid rank
constexpr uint32_t kDeltaMax = 0x7fffu;
constexpr uint32_t kDeltaManyBit = 0x40000000u;

// 2D scenes (IISPHFluidSolver / SESPHFluidSolver with HashedNeighborhoodSearch or
// QuadraticNeighborhoodSearchDynamicAllocated) keep every particle, inactive
// ones included, in the caller's order (the 2D searches never sort the
// collection); positions carry z = 0, so the 3-component arithmetic below
// gives the 2D values: glm's vec2 dot (x*x' + y*y') plus an exact + 0.

// SceneState.status bits
constexpr uint32_t kStatusCapacity = 1u;    // a particle had more than 97 neighbors
constexpr uint32_t kStatusGrid = 2u;        // a particle left the cell grid of this step
constexpr uint32_t kStatusNonFinite = 4u;
constexpr uint32_t kStatusTinyCoord = 8u;   // a coordinate lies in (0, 2^-36) in magnitude (not an error)
constexpr uint32_t kStatusRemoval = 16u;    // a particle remover entity found members among the active particles
constexpr uint32_t kStatusSortRetry = 32u;  // the counting sort met a cell it does not order: redo the step with the radix sort

// Per-scene state written by the device and read back once per step.
struct alignas(16) SceneState {
    double avg_err;        // average predicted density error of the last iteration (FP64)
    uint32_t cur;          // which particle buffer holds the current (sorted) state
    uint32_t active;       // Jacobi loop still iterating
    uint32_t iterations;   // Jacobi iterations of this step (stat_last_iteration_count)
    uint32_t status;       // kStatus* bits
    uint32_t done[4];      // last-block counters: 0 permute, 1 pass B, 2 integrate, 3 cfl
    float max_a2, max_v2;  // CFL maxima of the last iteration / SESPH step
    float dt_actual, max_a, max_v;
    uint32_t err_count;
    int32_t bbox_min[4], bbox_max[4];  // cell bounding box after the last integration
    float next_v2, next_a2;  // after integration: max |v|^2, |a|^2 over the particles that move
    uint64_t total_neighbors;
    uint64_t codec_bytes;  // NeighborStorage encoded bytes of all lists of the last search
    uint32_t pending;      // 3D: this step's integration has not run yet
    uint32_t delta_spill;  // delta lists: some list of the search in progress exceeds the escape rows
    uint32_t delta_plain;  // delta lists: the last search's walks read the 32-bit lists
    uint64_t delta_bytes;  // delta lists: code and escape bytes of all lists of the last search
    // compressed search: some list of the search in progress may need more than the
    // NeighborStorage payload (k_lists' bound), so k_codec sizes every list exactly
    uint32_t codec_check;
    uint32_t pad_codec[3];
};

// max of (v2, a2) over a kBlock-thread block, NaN ignored; every thread gets the result
__device__ __forceinline__ void block_max2(float& v2, float& a2, float* lds /* 2 * kBlock / 64 */) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v2 = fmaxf(v2, __shfl_xor(v2, o, 64));
        a2 = fmaxf(a2, __shfl_xor(a2, o, 64));
    }
    __syncthreads();
    if (__lane_id() == 0) {
        lds[threadIdx.x / 64] = v2;
        lds[kBlock / 64 + threadIdx.x / 64] = a2;
    }
    __syncthreads();
    v2 = a2 = 0.0f;
    for (int w = 0; w < kBlock / 64; w++) {
        v2 = fmaxf(v2, lds[w]);
        a2 = fmaxf(a2, lds[kBlock / 64 + w]);
    }
}

struct alignas(16) Partial {
    double err;
    uint32_t count;
    float max_a2;
    float max_v2;
    uint32_t aux;  // k_codec: the block's delta-coded list bytes
};

struct alignas(16) BoxPartial {
    int32_t lo[4];
    int32_t hi[4];
};

// Per-scene descriptor, one per scene in a launch table; blocks find their
// scene through a block map {scene, block within scene}.
// a simulation entity of the scene (the C ABI's fsa_entity, same layout)
struct EntityDev {
    uint32_t type, tag, remove_if_outside, reserved;
    float velocity[3];
    float center[3];
    float distance_from_center[3];
};
constexpr uint32_t kEntityBoundaryPreprocessor = 1u;
constexpr uint32_t kEntityVelocityAlteration = 2u;
constexpr uint32_t kEntityParticleRemover = 3u;
// SceneDev.entity_bits: which kinds the scene's list holds
constexpr uint32_t kEntityBitBoundary = 1u;
constexpr uint32_t kEntityBitVelocity = 2u;
constexpr uint32_t kEntityBitRemover = 4u;

struct alignas(16) SceneDev {
    uint32_t n;            // combined active particles
    uint32_t nf;           // fluid particles
    uint32_t list_stride;  // neighbor-list stride (nf rounded up to 64)
    uint32_t flags;

    float rest_density, gravity, particle_size, search_radius;
    float r2;              // Math::pow2(search_radius)
    float h, alpha;        // CubicSplineKernel3D::initialize
    float rho0sq;          // Math::pow2(rest_density)
    float visc_c01;        // 0.01f * particle_size * particle_size
    float visc2;           // 2 * viscosity of the active solver
    float omega, gamma, gamma1, gamma2, max_err;
    uint32_t min_it, max_it;
    float stiffness;
    float dt;              // desired time step of the next step
    float gen_dt, lambda_a, lambda_v;
    float inactive_a2, inactive_v2;  // SESPH CFL contribution of inactive particles
    uint32_t dim;          // 2 or 3
    // 2D hashed grid (HashedNeighborhoodSearch.cpp:117-151): members are the
    // first grid_count particles; a particle that changes cell is appended to
    // its new cell, so a cell lists its members by (entry, index)
    uint32_t search_count; // searches done before this one
    uint32_t grid_count;   // particles the grid update visits (uint16 count)
    uint32_t pad1;
    int2* cell_prev;       // 2D: cell of the last grid update
    uint32_t* entry;       // 2D: search count at which the particle entered its cell
    int4* cell_last;       // 3D hashed / quadratic: cell of the last grid update (GridCellState.current)

    // cell grid of this step (dense, linear x-fastest over the cell bounding box)
    int32_t cmin[4];
    uint32_t cdim[4];      // x, y, z, cell count
    uint32_t levels;       // Morton levels of the compact key
    uint32_t rank_keys;    // compressed search: sort by the rank table instead of the compact key
    const uint32_t* ranks; // rank of each grid cell in compact-key order among the grid's cells
    // counting sort over the ranks: particles per rank (zero between searches),
    // first sorted position of each rank (cdim[3] + 1 entries), sums of the
    // kCountBlocks parts of the rank range
    uint32_t* bcnt;
    uint32_t* boff;
    uint32_t* bsum;
    uint32_t stamp;        // cell-table generation
    uint32_t ntiles;       // radix tiles
    uint32_t nblocks;      // kBlock-blocks over n

    // combined particle arrays, two buffers each ([cur] is current)
    float4* pos[2];        // x, y, z, mass
    float4* vel[2];        // vx, vy, vz, density
    float4* acc[2];        // acceleration (IISPH: pressure acceleration)
    float4* ext[2];        // external (non-pressure) acceleration
    float* pres[2];
    float* err[2];         // IISPH predicted density error (kept across steps)
    uint32_t* idt[2];      // caller index | type << 30
    // per-step scratch (combined index)
    float* rho_new;
    float4* vstar;         // predicted velocity
    float4* anp;           // original non-pressure acceleration
    float* src;            // source term
    float* diag;           // diagonal element a_ii
    float4* dii;           // IISPH d_ii of fluid rank k, summed in phase 1 for phase 2
    // search
    uint32_t* keys[2];
    uint32_t* vals[2];
    uint32_t* hist;        // kRadix * ntiles tile offsets + kRadix digit totals
    uint4* cells;          // {start, end, stamp, -}
    uint32_t* fluid_idx;   // combined index of fluid rank k
    uint32_t* rank_of;     // fluid rank of combined index i (fluid particles only)
    uint32_t* block_count; // fluid count per kBlock-block, scanned in place
    uint32_t* nbr;         // neighbor entries, entry e of fluid rank k at e * list_stride + k
    uint32_t* dlist;       // delta-coded lists (kFlagDeltaLists): kDeltaRows rows of list_stride dwords
    uint32_t* nbr_cnt;
    Partial* partial;      // per fluid block
    BoxPartial* box;       // per combined block
    SceneState* state;
    // simulation entities around the solver
    const EntityDev* entities;
    const uint32_t* tag_of_row;  // ParticleInfo.tag per caller row (velocity alteration)
    uint32_t n_entities;
    uint32_t entity_bits;        // kEntityBit*
    uint32_t vals_buf;           // hashed / quadratic: the bucket buffer of the last search
    float bp_gamma1, bp_max_mass;  // BoundaryPreprocessor::initialize (BoundaryPreprocessor.cpp:50-62)
};

// dense index of cell (cx, cy, cz) in the step's cell grid; false outside it
__device__ __forceinline__ bool cell_linear(const SceneDev& S, int32_t cx, int32_t cy, int32_t cz, uint32_t& lin) {
    const int32_t x = cx - S.cmin[0], y = cy - S.cmin[1], z = cz - S.cmin[2];
    if (x < 0 || y < 0 || z < 0 || x >= (int32_t)S.cdim[0] || y >= (int32_t)S.cdim[1] || z >= (int32_t)S.cdim[2])
        return false;
    lin = (uint32_t)x + S.cdim[0] * ((uint32_t)y + S.cdim[1] * (uint32_t)z);
    return true;
}

struct BlockRef {
    uint32_t scene;
    uint32_t block;
};

// A scene whose search reported an error (neighbor capacity, cell grid) is
// left untouched by the solver kernels of that step: its particle state stays
// as the search left it, and the host reports the error.  A scene whose
// counting sort asked for a retry is left as it was before the search (the
// search kernels after the sort skip it too) and steps again with the radix sort.
__device__ __forceinline__ bool scene_failed(const SceneDev& S) {
    return (S.state->status & (kStatusCapacity | kStatusGrid | kStatusSortRetry)) != 0u;
}
__device__ __forceinline__ bool sort_retry(const SceneDev& S) { return (S.state->status & kStatusSortRetry) != 0u; }

// ---- reference-order arithmetic ------------------------------------------
// glm::dot for vec3: ((x*x') + (y*y')) + (z*z')
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return (ax * bx + ay * by) + az * bz;
}

// ---- correctly rounded sqrt and division on the restricted ranges of the pair math ----
// sqrt_rn_pair(x), x >= 0: the hardware square root (within 1 ulp) corrected by
// comparing x with the products of its neighbors (s - 1 ulp, s + 1 ulp): the
// residual x - t*s of a candidate t is exact in an FMA and its sign tells on
// which side of the rounding midpoint sqrt(x) lies.  For x >= 2^-96 this is
// the IEEE result bit for bit (checked exhaustively on the device); below, it
// returns some value in [0, 2^-40], which gives the same kernel value and a
// zero gradient as the exact root, since both then see q < 2^-25 and a length
// under FLT_EPSILON.
__device__ __forceinline__ float sqrt_rn_pair(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float lo = __int_as_float(__float_as_int(s) - 1);
    const float hi = __int_as_float(__float_as_int(s) + 1);
    float r = s;
    if (__builtin_fmaf(-lo, s, x) <= 0.0f) r = lo;
    if (__builtin_fmaf(-hi, s, x) > 0.0f) r = hi;
    return r;
}

// Reciprocal of b refined by one Newton step, for quotients sharing b.  For b
// in [2^-23, 2^2) this is the correctly rounded 1 / b (checked on the device for
// every mantissa at every exponent of that range).
__device__ __forceinline__ float recip_pair(float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y0, 1.0f);
    return __builtin_fmaf(e, y0, y0);
}

// a / b from y = RN(1 / b) with one FMA correction.  With q0 = RN(a * y) within
// one ulp of a / b, the residual a - b * q0 is exact in an FMA and RN(q0 + r * y)
// is the correctly rounded quotient (Markstein's theorem; Muller et al.,
// Handbook of Floating-Point Arithmetic, on division with a correctly rounded
// reciprocal).  It needs no underflow in the residual: a = 0 or
// 2^-59 <= |a| <= b, with 2^-23 < b <= 2.  The sign of a zero quotient is
// restored from a.
__device__ __forceinline__ float div_pair(float a, float b, float y) {
    const float q0 = a * y;
    const float r0 = __builtin_fmaf(-b, q0, a);
    return __builtin_copysignf(__builtin_fmaf(r0, y, q0), a);
}

// div_pair without restoring the sign of a zero quotient: the same value, but
// a zero quotient may come out as +0 where IEEE gives -0 (a = -0).  Used for
// the pair gradient, whose every consumer adds it (or products with it) into
// a sum started at +0; such a sum is never -0, so the sign of a zero term
// cannot reach a result.
__device__ __forceinline__ float div_pair_zs(float a, float b, float y) {
    const float q0 = a * y;
    const float r0 = __builtin_fmaf(-b, q0, a);
    return __builtin_fmaf(r0, y, q0);
}

// std::abs((int)e) as the reference's x86-64 build evaluates it (IISPHFluidSolver.hpp:400-402):
// the float-to-int conversion (cvttss2si) truncates in range and gives INT_MIN for NaN
// or |e| beyond the int range, and abs(INT_MIN) stays INT_MIN (the negation wraps).
// The C conversion is undefined there and the device's saturates, so both cases
// are spelled out.
__device__ __forceinline__ float x86_int_abs(float e) {
    if (!(e >= -2147483648.0f && e < 2147483648.0f)) return -2147483648.0f;
    const int v = (int)e;
    return v == INT_MIN ? -2147483648.0f : (float)(v < 0 ? -v : v);
}

// Compile-time specialisation of the two parameters that make operations
// exact no-ops: Spec bit 0 = every scene of the launch has h == 1, bit 1 =
// every scene has rest_density^2 == 1.  Without the bit the scene flag decides.
constexpr uint32_t kSpecHOne = 1u;
constexpr uint32_t kSpecRho0SqOne = 2u;
// Bit 2 (with bit 0 only): every scene of the launch has h == 1 and no coordinate
// in (0, 2^-36).  Then every pair difference is 0 or at least 2^-59 in magnitude
// and lengths with a nonzero gradient lie in (2^-23, 2], the ranges on which
// sqrt_rn_pair and div_pair equal the IEEE operations.
constexpr uint32_t kSpecFastMath = 4u;
// Bit 3: every scene of the launch keeps delta-coded neighbor lists (kFlagDeltaLists).
constexpr uint32_t kSpecDelta = 8u;
// smallest magnitude a nonzero coordinate may have for the fast pair math
constexpr float kTinyCoord = 1.4551915228366852e-11f;  // 2^-36

template <uint32_t Spec>
__device__ __forceinline__ bool spec_h_one(const SceneDev& S) {
    return (Spec & kSpecHOne) ? true : (S.flags & kFlagHOne) != 0;
}
template <uint32_t Spec>
__device__ __forceinline__ bool spec_rho_one(const SceneDev& S) {
    return (Spec & kSpecRho0SqOne) ? true : (S.flags & kFlagRho0SqOne) != 0;
}

// CubicSplineKernel3D::GetKernelValue (CubicSplineKernel3D.cpp:7-22) from |x|.
// Both polynomial branches are formed and selected; the selected value is the
// one the reference's if/else computes.
template <uint32_t Spec>
__device__ __forceinline__ float kernel_w(const SceneDev& S, float len) {
    const float q = spec_h_one<Spec>(S) ? len : len / S.h;
    const float a = 2.0f - q, b = 1.0f - q;
    const float a3 = a * a * a;
    const float inner = a3 - 4.0f * (b * b * b);
    // listed pairs have q <= 2 when h == 1 (|d|^2 <= 4), and at q == 2 the
    // cubic branch is (2 - 2)^3 = +0, the reference's zero
    const float ret = (Spec & kSpecFastMath) ? (q < 1.0f ? inner : a3) : (q < 1.0f ? inner : (q < 2.0f ? a3 : 0.0f));
    return S.alpha * ret;
}

// GetKernelDerivativeReversedValue(x_j, x_i) for d = x_i - x_j
// (CubicSplineKernel3D.cpp:24-47, 62-65): alpha * (pre * f'(q)) with
// pre = h * d / |d|, zero when |d| <= FLT_EPSILON.  The reference negates
// twice (derivative and reversal); negation is exact, so both are dropped.
template <uint32_t Spec>
__device__ __forceinline__ void kernel_grad(const SceneDev& S, float dx, float dy, float dz, float len,
                                            float& gx, float& gy, float& gz) {
    const bool hone = spec_h_one<Spec>(S);
    const float q = hone ? len : len / S.h;
    const bool nz = len > 1.1920928955078125e-07f;
    float px, py, pz;
    if (Spec & kSpecFastMath) {
        // h == 1: d / |d| with one shared reciprocal (exact on this range)
        const float y = recip_pair(len);
        px = div_pair_zs(dx, len, y);
        py = div_pair_zs(dy, len, y);
        pz = div_pair_zs(dz, len, y);
    } else if (hone) {
        px = dx / len;
        py = dy / len;
        pz = dz / len;
    } else {
        px = (S.h * dx) / len;
        py = (S.h * dy) / len;
        pz = (S.h * dz) / len;
    }
    px = nz ? px : 0.0f;
    py = nz ? py : 0.0f;
    pz = nz ? pz : 0.0f;
    const float a = 2.0f - q, b = 1.0f - q;
    const float m3a2 = -3.0f * (a * a);
    const float s = q < 1.0f ? m3a2 + 12.0f * (b * b) : m3a2;
    // With h == 1 a listed pair has q <= 2; at q == 2 the product below is a
    // signed zero where the reference has +0.  Every consumer adds the gradient
    // into a sum started at +0 (acc = acc + t or acc - t, denormals kept), which
    // is never -0, so the sign of a zero term cannot reach a result.
    const bool inside = (Spec & kSpecFastMath) ? true : q < 2.0f;
    gx = inside ? S.alpha * (px * s) : 0.0f;
    gy = inside ? S.alpha * (py * s) : 0.0f;
    gz = inside ? S.alpha * (pz * s) : 0.0f;
}

// length(x_i - x_j) from the squared length, correctly rounded where it matters
template <uint32_t Spec>
__device__ __forceinline__ float pair_length(float d2) {
    return (Spec & kSpecFastMath) ? sqrt_rn_pair(d2) : __builtin_sqrtf(d2);
}

template <uint32_t Spec>
__device__ __forceinline__ float div_rho0sq(const SceneDev& S, float v) {
    return spec_rho_one<Spec>(S) ? v : v / S.rho0sq;
}

// (int)std::floor(x / search_radius) (CompressedNeighbors.cpp:37-41)
__device__ __forceinline__ int32_t cell_of(float x, float r) {
    return (int32_t)floorf(x / r);
}

// Compact Morton key.  libFluid keys a cell by morton3D_64(c + 2^31) keeping
// 22/21/21 bits (CompressedNeighbors.cpp:43-56).  When every cell coordinate
// lies in [-2^L, 2^L), the bits above level L of each coordinate equal its
// sign bit, and the 64-bit key order equals the order of
//   sx << (3L+2) | sz << (3L+1) | sy << 3L | interleave_L(x, y, z)
// (the top key bits are x@63, z@62, y@61, x@60, ...).  3L + 3 <= 32 bits.
__device__ __host__ __forceinline__ uint32_t compact_key(int32_t cx, int32_t cy, int32_t cz, uint32_t L) {
    uint32_t mask = (1u << L) - 1u;
    uint32_t x = (uint32_t)cx & mask, y = (uint32_t)cy & mask, z = (uint32_t)cz & mask;
    uint32_t key = 0;
    for (uint32_t b = 0; b < L; b++) {
        key |= ((x >> b) & 1u) << (3 * b);
        key |= ((y >> b) & 1u) << (3 * b + 1);
        key |= ((z >> b) & 1u) << (3 * b + 2);
    }
    uint32_t sx = cx < 0, sy = cy < 0, sz = cz < 0;
    key |= sy << (3 * L);
    key |= sz << (3 * L + 1);
    key |= sx << (3 * L + 2);
    return key;
}


// Rank of cell c in compact_key order among the cells of the box [lo, hi]
// (inclusive, every coordinate in [-2^L, 2^L)): the number of box cells with
// a smaller key.  The key orders first by the sign bits (sx, sz, sy), then by
// the interleaved low L bits (z, y, x within each level, highest level first),
// and the low bits of either sign half are monotone in the coordinate, so the
// count is the box cells of the earlier sign octants plus, inside c's octant,
// those of the earlier children at every level of the octree walk down to c.
__device__ __host__ __forceinline__ uint32_t box_rank(int32_t cx, int32_t cy, int32_t cz, const int32_t lo[3],
                                                      const int32_t hi[3], uint32_t L) {
    const int32_t c[3] = {cx, cy, cz};
    const int32_t half = 1 << L;
    int32_t alo[3], ahi[3];  // the box inside c's sign half, in low-bit coordinates
    uint32_t n0[3], n1[3];   // box cells per axis with sign 0 and with sign 1
    uint32_t s[3];
    for (int a = 0; a < 3; a++) {
        const int32_t p0 = lo[a] > 0 ? lo[a] : 0, q1 = hi[a] < -1 ? hi[a] : -1;
        n0[a] = hi[a] >= p0 ? (uint32_t)(hi[a] - p0 + 1) : 0u;
        n1[a] = q1 >= lo[a] ? (uint32_t)(q1 - lo[a] + 1) : 0u;
        s[a] = c[a] < 0 ? 1u : 0u;
        alo[a] = s[a] ? lo[a] + half : p0;
        ahi[a] = s[a] ? q1 + half : hi[a];
    }
    // earlier octants in (sx, sz, sy) order
    const uint32_t tz = n0[2] + n1[2], ty = n0[1] + n1[1];
    const uint32_t nx = s[0] ? n1[0] : n0[0], nz = s[2] ? n1[2] : n0[2];
    uint32_t rank = (s[0] ? n0[0] * tz * ty : 0u) + nx * ((s[2] ? n0[2] * ty : 0u) + nz * (s[1] ? n0[1] : 0u));
    // inside the octant: Morton rank of the low bits in the box [alo, ahi]
    int32_t base[3] = {0, 0, 0};
    const int32_t u[3] = {c[0] + (int32_t)s[0] * half, c[1] + (int32_t)s[1] * half, c[2] + (int32_t)s[2] * half};
    for (int32_t b = (int32_t)L - 1; b >= 0; b--) {
        const int32_t w = 1 << b;
        uint32_t ov[3][2], bit[3];
        for (int a = 0; a < 3; a++) {
            for (int h = 0; h < 2; h++) {
                const int32_t l0 = base[a] + h * w, l1 = l0 + w - 1;
                const int32_t x0 = alo[a] > l0 ? alo[a] : l0, x1 = ahi[a] < l1 ? ahi[a] : l1;
                ov[a][h] = x1 >= x0 ? (uint32_t)(x1 - x0 + 1) : 0u;
            }
            bit[a] = (uint32_t)(u[a] >> b) & 1u;
        }
        const uint32_t vx = ov[0][0] + ov[0][1], vy = ov[1][0] + ov[1][1];
        rank += (bit[2] ? ov[2][0] * vy * vx : 0u) +
                ov[2][bit[2]] * ((bit[1] ? ov[1][0] * vx : 0u) + ov[1][bit[1]] * (bit[0] ? ov[0][0] : 0u));
        for (int a = 0; a < 3; a++) base[a] += (int32_t)bit[a] * w;
    }
    return rank;
}

// The reference's 64-bit cell key (CompressedNeighbors.cpp:43-56, also the
// hashed search's re-sort key): libmorton's morton3D_64 of c + 2^31, bit b of
// x, y, z at key bits 3b, 3b + 1, 3b + 2 (x keeps 22 bits, y and z 21).
__device__ __host__ __forceinline__ uint64_t reference_cell_key(int32_t cx, int32_t cy, int32_t cz) {
    const uint32_t u[3] = {(uint32_t)((int64_t)cx - INT32_MIN), (uint32_t)((int64_t)cy - INT32_MIN),
                           (uint32_t)((int64_t)cz - INT32_MIN)};
    uint64_t key = 0;
    for (int b = 0; b < 22; b++)
        for (int a = 0; a < 3; a++) {
            const int bit = 3 * b + a;
            if (bit < 64) key |= (uint64_t)((u[a] >> b) & 1u) << bit;
        }
    return key;
}

// Key order of a cell's 27 neighbor cells without sorting.  Along one axis the
// compact keys of c-1, c, c+1 differ in two steps: one changes only coordinate
// bit 0 (key bit 0, 1 or 2 for x, y, z), the other changes the coordinate bits
// up to p (trailing zeros of an even c, trailing ones of an odd c), or, between
// -1 and 0, the sign bit at the top of the key (x, z, y from the top).  So every
// neighbor has, per axis, a side bit B (which side of that axis's big step) and
// a bit S (which value within its side), and the key order of the 27 cells is
// the order of the 6-bit code: the three B bits, the axis whose big step changes
// the highest key bit first, above the S bits of z, y, x.  (Negative
// coordinates carry the sign bit, so in key order 0 < 1 < ... < -2 < -1: for
// c = 0 the big step lies between c and c - 1, for c = -1 between c + 1 and c.)
//
// The order packs into one word: bits 2a..2a+1 the type of axis a (0 even,
// 1 odd, 2 zero, 3 minus one), bits 6+3a..8+3a the code bit of its B bit
// (3, 4 or 5), bits 15..20 the code of the cell itself.
constexpr uint32_t kOrderEven = 0u, kOrderOdd = 1u, kOrderZero = 2u, kOrderMinus1 = 3u;
// per type, 2 bits per (B, S) at bit 2 * (2B + S): the offset + 1, or 3 where no
// neighbor has that pair.  even: (0,0) -1, (1,0) 0, (1,1) +1; odd: (0,0) -1,
// (0,1) 0, (1,0) +1; zero: (0,0) 0, (0,1) +1, (1,0) -1; minus one: (0,0) +1,
// (1,0) -1, (1,1) 0
constexpr uint32_t kOrderLuts = 0x9Cu | 0xE4u << 8 | 0xC9u << 16 | 0x4Eu << 24;

__device__ __host__ __forceinline__ uint32_t neighbor_order_word(int32_t cx, int32_t cy, int32_t cz, uint32_t L) {
    const int32_t c[3] = {cx, cy, cz};
    const uint32_t sign_pos[3] = {3 * L + 2, 3 * L, 3 * L + 1};  // key bit of the sign of x, y, z
    uint32_t big[3], type[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int32_t v = c[a];
        const bool zero = v == 0, minus1 = v == -1, odd = (v & 1) != 0;
        // trailing zeros (even) or ones (odd), at most L - 1: a step past the key
        // range has no neighbor in the grid, and the clamp keeps the positions distinct
        const uint32_t w = (odd ? ~(uint32_t)v : (uint32_t)v) | (1u << (L - 1u));
        big[a] = (zero || minus1) ? sign_pos[a] : 3u * (uint32_t)__builtin_ctz(w) + (uint32_t)a;
        type[a] = zero ? kOrderZero : minus1 ? kOrderMinus1 : odd ? kOrderOdd : kOrderEven;
    }
    uint32_t word = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const uint32_t bpos = 3u + (big[a] > big[(a + 1) % 3] ? 1u : 0u) + (big[a] > big[(a + 2) % 3] ? 1u : 0u);
        // the cell's own (B, S): even (1,0), odd (0,1), zero (0,0), minus one (1,1)
        const uint32_t selfB = (type[a] == kOrderEven || type[a] == kOrderMinus1) ? 1u : 0u;
        const uint32_t selfS = (type[a] == kOrderOdd || type[a] == kOrderMinus1) ? 1u : 0u;
        word |= type[a] << (2 * a) | bpos << (6 + 3 * a) | (selfB << bpos | selfS << a) << 15;
    }
    return word;
}

struct NeighborOrder {
    uint32_t lut[3];   // per axis: 2-bit offset + 1 (3: none) at bit 2 * (2B + S)
    uint32_t bpos[3];  // code bit of the axis's B bit (3, 4 or 5)
    uint32_t self;     // code of the cell itself
};

__device__ __host__ __forceinline__ NeighborOrder unpack_order(uint32_t word) {
    NeighborOrder o;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        o.lut[a] = (kOrderLuts >> (8u * ((word >> (2 * a)) & 3u))) & 0xffu;
        o.bpos[a] = (word >> (6 + 3 * a)) & 7u;
    }
    o.self = (word >> 15) & 63u;
    return o;
}

// offset (-1, 0, +1) of lane `code`'s neighbor along axis a; false where no neighbor has the code
__device__ __host__ __forceinline__ bool neighbor_offset(const NeighborOrder& o, uint32_t code, int a, int32_t& d) {
    const uint32_t idx = ((code >> o.bpos[a]) & 1u) << 1 | ((code >> a) & 1u);
    const uint32_t v = (o.lut[a] >> (2u * idx)) & 3u;
    d = (int32_t)v - 1;
    return v != 3u;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Global-address-space views of device pointers: loads through them are
// global_load (not flat_load, which also waits on the LDS counter).
#define FSA_GLOBAL(T) __attribute__((address_space(1))) T
typedef float fsa_f4 __attribute__((ext_vector_type(4)));

struct GView4 {
    const FSA_GLOBAL(fsa_f4) * p;
    __device__ __forceinline__ float4 operator[](uint32_t i) const {
        const fsa_f4 v = p[i];
        return make_float4(v.x, v.y, v.z, v.w);
    }
    // element i of a particle array (i < 2^27) at a 32-bit byte offset from the
    // array's scalar base: the load takes the base from SGPRs and one VGPR offset
    __device__ __forceinline__ float4 at(uint32_t i) const {
        const fsa_f4 v = *(const FSA_GLOBAL(fsa_f4)*)((const FSA_GLOBAL(char)*)p + (i << 4));
        return make_float4(v.x, v.y, v.z, v.w);
    }
};
struct GViewF {
    const FSA_GLOBAL(float) * p;
    __device__ __forceinline__ float operator[](uint32_t i) const { return p[i]; }
    __device__ __forceinline__ float at(uint32_t i) const {
        return *(const FSA_GLOBAL(float)*)((const FSA_GLOBAL(char)*)p + (i << 2));
    }
};
struct GViewU {
    const FSA_GLOBAL(uint32_t) * p;
    __device__ __forceinline__ uint32_t operator[](uint32_t i) const { return p[i]; }
};
__device__ __forceinline__ GView4 gload_ptr(const float4* p) { return GView4{(const FSA_GLOBAL(fsa_f4)*)p}; }
__device__ __forceinline__ GViewF gload_ptr(const float* p) { return GViewF{(const FSA_GLOBAL(float)*)p}; }
__device__ __forceinline__ GViewU gload_ptr(const uint32_t* p) { return GViewU{(const FSA_GLOBAL(uint32_t)*)p}; }

// Walks a neighbor list in order, U entries at a time: the list entries and
// the per-neighbor gathers of a group are issued before any of them is used,
// then `body` runs for each entry in list order (the accumulation order of
// the reference loop is unchanged).
template <int U, class Load, class Body>
__device__ __forceinline__ void walk_neighbors(GViewU nb, uint32_t stride, uint32_t cnt, Load load, Body body) {
    using T = decltype(load(0u));
    uint32_t e = 0;
    for (; e + U <= cnt; e += U) {
        uint32_t ent[U];
        T d[U];
#pragma unroll
        for (int q = 0; q < U; q++) ent[q] = nb[(e + q) * stride];
#pragma unroll
        for (int q = 0; q < U; q++) d[q] = load(ent[q] & kIndexMask);
#pragma unroll
        for (int q = 0; q < U; q++) body(ent[q], d[q]);
    }
    for (; e < cnt; e++) {
        const uint32_t en = nb[e * stride];
        const T dd = load(en & kIndexMask);
        body(en, dd);
    }
}

// One entry of a delta-coded list from its 16-bit code c: the previous index
// plus the delta (boundary flag from bit 15), or the next escaped entry.  The
// escape queue holds two entries (qa, then qb); a list with more escapes
// (`many`) loads each further one when the queue moves up.
struct DeltaCursor {
    uint32_t prev;  // index of the previous entry
    uint32_t qa, qb;
    uint32_t next;  // escape row to load into qb when the queue moves up
    bool many;
};
__device__ __forceinline__ uint32_t delta_entry(uint32_t c, DeltaCursor& q, GViewU esc, uint32_t stride) {
    const uint32_t delta = c & kDeltaMax;
    const bool is_esc = delta == 0u;
    const uint32_t en = is_esc ? q.qa : ((q.prev + delta) | ((c & 0x8000u) << 16));
    q.prev = en & kIndexMask;
    q.qa = is_esc ? q.qb : q.qa;
    // a walked list has at most kDeltaEscRows escapes, so rows from there on are never needed
    if (is_esc && q.many && q.next < kDeltaEscRows) {
        q.qb = esc[q.next * stride];
        q.next++;
    }
    return en;
}

// walk_neighbors with the list entries addressed by 32-bit byte offsets from the
// list array's scalar base (entry e of rank k at (e * stride + k) * 4, below 4 GiB)
template <int U, class Load, class Body>
__device__ __forceinline__ void walk_neighbors32(const uint32_t* nbr, uint32_t k, uint32_t stride, uint32_t cnt,
                                                 Load load, Body body) {
    using T = decltype(load(0u));
    const FSA_GLOBAL(char)* base = (const FSA_GLOBAL(char)*)nbr;
    const uint32_t step = stride << 2;
    uint32_t off = k << 2;
    uint32_t e = 0;
    for (; e + U <= cnt; e += U) {
        uint32_t ent[U];
#pragma unroll
        for (int q = 0; q < U; q++) ent[q] = *(const FSA_GLOBAL(uint32_t)*)(base + (off + (uint32_t)q * step));
        off += (uint32_t)U * step;
        T d[U];
#pragma unroll
        for (int q = 0; q < U; q++) d[q] = load(ent[q] & kIndexMask);
#pragma unroll
        for (int q = 0; q < U; q++) body(ent[q], d[q]);
    }
    for (; e < cnt; e++) {
        const uint32_t en = *(const FSA_GLOBAL(uint32_t)*)(base + off);
        off += step;
        const T dd = load(en & kIndexMask);
        body(en, dd);
    }
}

// walk_neighbors over a delta-coded list: codes of 64 consecutive ranks are
// read two entries per 256-byte load, the entries decoded in order, and their
// gathers issued before any of them is used.  U is even.
template <int U, class Load, class Body>
__device__ __forceinline__ void walk_neighbors_delta(GViewU codes, GViewU esc, uint32_t stride, uint32_t cnt,
                                                     Load load, Body body) {
    static_assert(U % 2 == 0, "codes come in pairs");
    using T = decltype(load(0u));
    DeltaCursor q;
    q.prev = 0;
    q.qa = esc[0];
    q.qb = esc[stride];
    q.many = (q.qa & kDeltaManyBit) != 0u;
    q.next = 2;
    uint32_t e = 0;
    for (; e + U <= cnt; e += U) {
        uint32_t w[U / 2];
#pragma unroll
        for (int r = 0; r < U / 2; r++) w[r] = codes[(e / 2 + r) * stride];
        uint32_t ent[U];
#pragma unroll
        for (int t = 0; t < U; t++) ent[t] = delta_entry((t & 1) ? (w[t / 2] >> 16) : w[t / 2], q, esc, stride);
        T d[U];
#pragma unroll
        for (int t = 0; t < U; t++) d[t] = load(ent[t] & kIndexMask);
#pragma unroll
        for (int t = 0; t < U; t++) body(ent[t], d[t]);
    }
    for (; e < cnt; e++) {
        const uint32_t w = codes[(e / 2) * stride];
        const uint32_t en = delta_entry((e & 1u) ? (w >> 16) : w, q, esc, stride);
        const T dd = load(en & kIndexMask);
        body(en, dd);
    }
}

// The neighbor list of fluid rank k: delta-coded when the launch is
// specialised for it (Spec & kSpecDelta: every scene of the launch keeps delta
// lists) and the last search could code every list, otherwise the 32-bit entries.
template <uint32_t Spec, int U, class Load, class Body>
__device__ __forceinline__ void walk_list(const SceneDev& S, uint32_t k, uint32_t cnt, Load load, Body body) {
    if ((Spec & kSpecDelta) && !S.state->delta_plain) {
        walk_neighbors_delta<(U < 2 ? 2 : U)>(gload_ptr(S.dlist + k), gload_ptr(S.dlist + kDeltaCodeRows * S.list_stride + k),
                                           S.list_stride, cnt, load, body);
    } else if ((uint64_t)S.list_stride * (uint64_t)(4 * kMaxNeighbors) <= 0xffffffffull) {
        walk_neighbors32<U>(S.nbr, k, S.list_stride, cnt, load, body);
    } else {
        walk_neighbors<U>(gload_ptr(S.nbr + k), S.list_stride, cnt, load, body);
    }
}

}  // namespace fsa
