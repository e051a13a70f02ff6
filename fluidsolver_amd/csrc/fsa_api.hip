// fsa_api.hip — C ABI of the gfx950 SPH hot path (include/fluidsolver_amd.h).
//
// Host orchestration of one simulation step, as Simulator::execute_simulation_step
// drives it (Simulator.cpp:7-57): find_neighbors, then the solver step.  The
// Jacobi loop is device-gated: every pass checks the scene's `active` flag, so
// the host enqueues iterations in chunks and reads the flag once per chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <deque>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstring>
#include <exception>
#include <map>
#include <numeric>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fluidsolver_amd.h"
#include "fsa_device.hpp"
#include "fsa_host.hpp"
#include "fsa_kernels.hpp"

namespace fsa {
thread_local std::string g_error;
}  // namespace fsa

using namespace fsa;

namespace {

#define HIP_CHECK(expr)                                                                                  \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess)                                                                            \
            throw FsaError(FSA_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e));           \
    } while (0)

uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

// device-memory plan of one scene (all arrays carved from one allocation)
struct Plan {
    uint64_t n, nf, stride, nblocks, nfblocks, ntiles, dim;
    bool unsorted;  // 3D hashed / quadratic: per-particle cell of the last grid update
    bool delta;     // 3D compressed search with delta-coded lists
    uint64_t bytes;
    std::vector<std::pair<void**, uint64_t>> parts;
};

}  // namespace

struct fsa_context_s {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    std::atomic<uint64_t> in_use{0};
    std::mutex mu;
    bool timing = false;
    double timer_ms[FSA_TIMER_COUNT] = {};
    uint64_t timer_n[FSA_TIMER_COUNT] = {};
    std::vector<hipEvent_t> ev_pool;   // start/stop pairs of the phases in flight
    std::vector<int> ev_phase;
    size_t ev_used = 0;
    bool graph = false;
    uint32_t block_override = 0;
    std::vector<fsa_scene_s*> scenes;  // live scenes, destroyed with the context
    // per-scene device state lives in one pool so a batch reads it back in one copy
    static constexpr uint32_t kMaxScenes = 4096;
    SceneState* d_states = nullptr;
    SceneState* h_states = nullptr;  // pinned mirror
    std::vector<uint32_t> free_slots;
    std::map<std::vector<fsa_scene_s*>, struct BatchSet*> batches;
    uint64_t batch_clock = 0;
    std::vector<struct fsa_server_s*> servers;  // destroyed with the context
    // autotuned neighbor-walk variant per size class, with the timings behind each choice
    bool autotune = true;
    struct Tuned {
        std::array<uint8_t, 4> unroll;  // kVariant* per family
        std::array<std::array<float, kTuneVariants>, 4> us;  // per family, per kTuneVariant entry
        uint64_t fluid;
    };
    std::map<uint64_t, Tuned> tuned;
    uint8_t walk_variant = 0;  // a kVariant* forced on every family, or 0: the autotuner's choice
    void* stat_scratch = nullptr;  // this device's address of the no-allocation statistics scratch
    // pinned host staging of uploads and downloads, grown as needed (one user at a time)
    void* pin = nullptr;
    uint64_t pin_bytes = 0;
    std::mutex pin_mu;
};

namespace {
// The stream of a thread's scene work: the context stream, or a stream of the
// thread's own while another thread steps scenes on the context stream (a
// server tick uploads admitted jobs and downloads finished ones beside the step).
thread_local hipStream_t t_stream = nullptr;
inline hipStream_t stream_of(const fsa_context_s* c) { return t_stream ? t_stream : c->stream; }
}  // namespace

// Launch set over several scenes (one table entry and block-map range each).
struct BatchSet {
    std::vector<fsa_scene_s*> scenes;
    SceneDev* d_table = nullptr;
    std::vector<SceneDev> h_table;
    LaunchSet ls{};
    std::vector<BlockRef*> owned;
    uint64_t last_use = 0;
    // instantiated step heads, keyed by what the captured launches depend on:
    // sort passes (2D: entry and cell passes), Jacobi passes, spec, boundary pass
    std::map<std::array<uint32_t, 5>, hipGraphExec_t> graphs;
    bool tuned = false;
};

struct fsa_scene_s {
    fsa_context ctx = nullptr;
    fsa_scene_desc desc{};
    uint64_t total = 0;          // caller rows
    uint32_t n = 0, nf = 0, nb = 0;
    std::vector<uint32_t> inactive_rows;
    std::vector<float> inactive_movement, inactive_pdata, inactive_ext;
    std::vector<uint8_t> inactive_info;
    SceneDev hd{};
    SceneDev* d_table = nullptr;   // single-scene launch table
    void* dmem = nullptr;
    uint64_t bytes = 0;
    uint4* cells = nullptr;     // cell table, then the rank table (kCellBytes per grid cell)
    uint64_t cells_cap = 0;
    uint32_t* ranks = nullptr;
    uint32_t* bcnt = nullptr;   // counting sort: bucket counts (zero between searches), offsets, part sums
    uint32_t* boff = nullptr;
    uint32_t* bsum = nullptr;
    uint32_t radix_steps = 0;   // steps left on the radix sort after a counting-sort retry
    bool ranks_valid = false;   // the rank table holds the grid below
    int32_t rank_cmin[3] = {0, 0, 0};
    uint32_t rank_cdim[3] = {0, 0, 0}, rank_levels = 0;
    SceneState* h_state = nullptr;  // pinned mirror of this scene's slot in the context pool
    uint32_t slot = 0;
    LaunchSet single{};
    std::vector<BlockRef*> map_allocs;
    bool searched = false;
    bool boundary_dirty = true;
    uint32_t last_iterations = 2;
    // 3D IISPH: the Jacobi passes predicted for the last step, and how many steps in a
    // row the loop has ended within the prediction (the gated extra pass is left out
    // of steady scenes' steps)
    uint32_t step_chunk = 0;
    uint32_t steady_steps = 0;
    // HashedNeighborhoodSearch3D: calls since the last re-sort and the collection
    // order the reference would hold (caller rows, inactive included)
    uint32_t hashed_calls = 0;
    std::vector<uint32_t> ref_order;
    // hashed-3D re-sort on the device: scratch of the quick sort, and the
    // collection order (rows) it leaves on the device; the host copy above is
    // refreshed from it when an upload needs it
    void* qs_mem = nullptr;
    uint64_t qs_bytes = 0, qs_rows = 0;
    QsBuffers qs{};
    uint32_t* qs_ref[2] = {nullptr, nullptr};
    uint32_t* qs_inv = nullptr;
    uint32_t* qs_block_cnt = nullptr;
    uint32_t qs_cur = 0;
    bool qs_ref_valid = false;  // qs_ref[qs_cur] holds the current order
    bool qs_ref_newer = false;  // ... and the host's ref_order is out of date
    bool lists_cached = false;
    bool lists_valid = false;  // lists of the last search (valid until the next search or upload)
    std::vector<uint64_t> csr_off, csr_cols;
    float cur_v2 = 0.0f, cur_a2 = 0.0f;  // max |v|^2, |a|^2 of the active particles' current state
    // simulation entities (fsa_scene_set_entities)
    std::vector<fsa_entity> entities;
    EntityDev* d_entities = nullptr;
    uint32_t* d_tags = nullptr;          // ParticleInfo.tag per caller row, for velocity alterations
    uint64_t d_tags_cap = 0;
    std::vector<uint32_t> row_tags;      // caller rows, as last uploaded
    std::vector<uint8_t> row_types;
    bool entity_moves_boundary = false;  // a velocity alteration reaches boundary particles
    bool removal_pending = false;        // a particle remover found members in the last step
    std::vector<float> carry_err;        // IISPH predicted density errors by caller row for the next upload
    void* d_stats = nullptr;             // sensor reduction scratch
    uint64_t d_stats_bytes = 0;
    void* d_query = nullptr;             // position queries: positions, counts, offsets
    uint64_t d_query_bytes = 0;
    void* d_query_out = nullptr;         // position queries: results
    uint64_t d_query_out_bytes = 0;
};

static_assert(sizeof(EntityDev) == sizeof(fsa_entity), "EntityDev mirrors fsa_entity");

namespace {

// ---- the statistics scratch of callers that cannot allocate -------------------------------
// fsa_scene_get_statistics_noalloc keeps the sensors' per-block partials and the reduced
// result in this file-scope device array instead of a per-scene allocation.  Every GPU
// the library runs on holds its own copy, at its own address, until the process ends;
// host code reaches it only through the address hipGetSymbolAddress gives on that device.
constexpr uint32_t kNoAllocStatBlocks = 65536;  // 65,536 blocks of 256: 16,777,216 particles
static __device__ __align__(256) uint8_t scratch[(kNoAllocStatBlocks + 1) * kStatAccBytes];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {};

// address of `scratch` on `device`, which must be current; looked up once per device
void* scratch_address(int device) {
    if (device < 0 || device >= kScratchDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

// the sensors' values from the reduced accumulator (a host copy of it)
void fill_statistics(const fsa_scene_s* s, const void* res, fsa_scene_statistics* out) {
    uint64_t cnt[3];
    float mx[3], mn[3];
    double sum[5];
    read_statistics(res, cnt, mx, mn, sum);
    std::memset(out, 0, sizeof(*out));
    out->normal_particles = cnt[0];
    out->boundary_particles = cnt[1];
    out->inactive_particles = cnt[2] + s->inactive_rows.size();
    out->density_max = mx[0];
    out->density_min = mn[0];
    out->pressure_max = mx[1];
    out->pressure_min = mn[1];
    out->velocity_max = mx[2];
    out->velocity_min = mn[2];
    const double c = (double)cnt[0];
    out->density_avg = cnt[0] ? sum[0] / c : 0.0;
    out->pressure_avg = cnt[0] ? sum[1] / c : 0.0;
    out->velocity_avg = cnt[0] ? sum[2] / c : 0.0;
    out->kinetic_energy = sum[3];
    out->potential_energy = sum[4];
}

// the search a description names, with FSA_SEARCH_DEFAULT resolved per dimension
uint32_t search_of(const fsa_scene_desc* d) {
    if (d->search != FSA_SEARCH_DEFAULT) return d->search;
    return d->dimension == 2 ? FSA_SEARCH_HASHED_2D : FSA_SEARCH_COMPRESSED_3D;
}

// the scene keeps delta-coded neighbor lists (3D compressed search only)
bool delta_lists(const fsa_scene_desc& d) {
    return d.dimension == 3 && search_of(&d) == FSA_SEARCH_COMPRESSED_3D && d.neighbor_lists == FSA_LIST_DELTA;
}

void fill_params(fsa_scene_s* s) {
    const fsa_scene_desc& d = s->desc;
    SceneDev& h = s->hd;
    h.rest_density = d.rest_density;
    h.gravity = d.gravity;
    h.particle_size = d.particle_size;
    h.search_radius = d.particle_size * 2.0f;   // Math::kernel_support_factor (LibFluidMath.hpp:11)
    h.r2 = h.search_radius * h.search_radius;
    h.h = h.search_radius / 2.0f;               // CubicSplineKernel3D::initialize (:67-70)
    const float PI = 3.14159265358979323846f;
    h.dim = d.dimension;
    if (d.dimension == 3)
        h.alpha = 1.0f / (4.0f * PI * (h.h * h.h * h.h));
    else
        h.alpha = 5.0f / (14.0f * PI * (h.h * h.h));  // CubicSplineKernel::initialize (CubicSplineKernel.cpp:70-80)
    h.rho0sq = d.rest_density * d.rest_density;
    h.visc_c01 = 0.01f * d.particle_size * d.particle_size;
    const bool sesph = d.solver == FSA_SOLVER_SESPH;
    h.visc2 = 2.0f * (sesph ? d.sesph_viscosity : d.iisph_viscosity);
    h.omega = d.iisph_omega;
    h.gamma = d.iisph_gamma;
    h.gamma1 = sesph ? d.sesph_single_layer_gamma_1 : d.iisph_single_layer_gamma_1;
    h.gamma2 = d.sesph_single_layer_gamma_2;
    h.max_err = d.iisph_max_density_error_allowed;
    h.min_it = d.iisph_min_iterations;
    h.max_it = d.iisph_max_iterations;
    h.stiffness = d.sesph_stiffness_k;
    // fallback of get_non_cfl_validating_timestep: the generator's own step
    // (ConstantTimestepGenerator.cpp:28-29; DynamicCflTimestepGenerator.cpp:96-97)
    h.gen_dt = d.timestep_generator == FSA_TIMESTEP_DYNAMIC_CFL ? d.max_timestep : d.timestep;
    h.dt = d.timestep;
    h.lambda_a = d.lambda_a;
    h.lambda_v = d.lambda_v;
    uint32_t f = h.flags & (kFlagExternal | kFlagDynamicBoundary | kFlagTinyCoords);
    if (h.h == 1.0f) f |= kFlagHOne;
    if (h.rho0sq == 1.0f) f |= kFlagRho0SqOne;
    if (sesph ? d.sesph_single_layer_boundary : d.iisph_single_layer_boundary) f |= kFlagSingleLayer;
    if (sesph) f |= kFlagSesph;
    if (d.neighbor_lists == FSA_LIST_PLAIN) f |= kFlagPlainLists;
    if (d.dimension == 2 && !sesph && (d.flags & FSA_FLAG_EMULATE_INT_ABS)) f |= kFlagIntAbs;
    const uint32_t search = search_of(&d);
    if (search == FSA_SEARCH_QUADRATIC_2D || search == FSA_SEARCH_QUADRATIC_3D) f |= kFlagQuadratic;
    if (search == FSA_SEARCH_HASHED_3D || search == FSA_SEARCH_QUADRATIC_3D) f |= kFlagUnsorted;
    if (search == FSA_SEARCH_COMPRESSED_3D && d.neighbor_lists != FSA_LIST_PLAIN) f |= kFlagCodec;
    if (delta_lists(d)) f |= kFlagDeltaLists;
    h.flags = f;
}

void validate_desc(const fsa_scene_desc* d) {
    if (!d) throw FsaError(FSA_E_INVALID, "desc is NULL");
    if (d->dimension != 2 && d->dimension != 3) throw FsaError(FSA_E_INVALID, "dimension must be 2 or 3");
    const uint32_t search = search_of(d);
    if (d->dimension == 3 && search != FSA_SEARCH_COMPRESSED_3D && search != FSA_SEARCH_HASHED_3D &&
        search != FSA_SEARCH_QUADRATIC_3D)
        throw FsaError(FSA_E_INVALID, "3D scenes use the compressed-3d, hashed-3d or quadratic-3d neighbor order");
    if (d->dimension == 2 && search != FSA_SEARCH_HASHED_2D && search != FSA_SEARCH_QUADRATIC_2D)
        throw FsaError(FSA_E_INVALID, "2D scenes use the hashed or quadratic-dynamic-allocated neighbor order");
    if (d->solver != FSA_SOLVER_IISPH && d->solver != FSA_SOLVER_SESPH)
        throw FsaError(FSA_E_INVALID, "unknown solver");
    if (d->neighbor_lists > FSA_LIST_DELTA) throw FsaError(FSA_E_INVALID, "unknown neighbor-list mode");
    if (!(d->particle_size > 0.0f)) throw FsaError(FSA_E_INVALID, "particle_size must be > 0");
    if (!(d->rest_density > 0.0f)) throw FsaError(FSA_E_INVALID, "rest_density must be > 0");
    if (!(d->timestep > 0.0f)) throw FsaError(FSA_E_INVALID, "timestep must be > 0");
    if (d->timestep_generator != FSA_TIMESTEP_CONSTANT && d->timestep_generator != FSA_TIMESTEP_DYNAMIC_CFL)
        throw FsaError(FSA_E_INVALID, "unknown timestep generator");
    if (d->timestep_generator == FSA_TIMESTEP_DYNAMIC_CFL && !(d->max_timestep > 0.0f && d->min_timestep > 0.0f))
        throw FsaError(FSA_E_INVALID, "dynamic CFL time steps must be > 0");
    if (d->solver == FSA_SOLVER_IISPH) {
        if (d->iisph_max_iterations < d->iisph_min_iterations)
            throw FsaError(FSA_E_INVALID, "Max iterations are less than min number of iterations.");
        if (!(d->iisph_max_density_error_allowed > 0.0f))
            throw FsaError(FSA_E_INVALID, "MaxDensityErrorAllowed is smaller or equal to zero.");
    }
}

Plan make_plan(uint64_t n, uint64_t nf, uint32_t dim = 3, bool unsorted = false, bool delta = false) {
    Plan p{};
    p.dim = dim;
    p.unsorted = unsorted;
    p.delta = delta;
    p.n = n;
    p.nf = nf;
    p.stride = align_up(std::max<uint64_t>(nf, 1), 64);
    p.nblocks = (n + kBlock - 1) / kBlock;
    p.nfblocks = (nf + kBlock - 1) / kBlock;
    p.ntiles = (n + kSortTile - 1) / kSortTile;
    return p;
}

uint64_t plan_bytes(const Plan& p) {
    const uint64_t n = std::max<uint64_t>(p.n, 1);
    uint64_t b = 0;
    auto add = [&](uint64_t sz) { b += align_up(sz, 256); };
    for (int k = 0; k < 8; k++) add(16 * n);   // pos, vel, acc, ext x2
    for (int k = 0; k < 6; k++) add(4 * n);    // pres, err, idt x2
    add(4 * n);                                // rho_new
    add(16 * n);                               // vstar
    add(16 * n);                               // anp
    add(4 * n);                                // src
    add(4 * n);                                // diag
    add(16 * std::max<uint64_t>(p.nf, 1));     // dii
    for (int k = 0; k < 4; k++) add(4 * n);    // keys, vals x2
    add(4 * (kRadixMaxDigits * std::max<uint64_t>(p.ntiles, 1) + kRadixMaxDigits));
    add(4 * std::max<uint64_t>(p.nf, 1));      // fluid_idx
    add(4 * n);                                // rank_of
    add(4 * std::max<uint64_t>(p.nblocks, 1));
    add(4ull * kMaxNeighbors * p.stride);      // nbr
    add(4 * std::max<uint64_t>(p.nf, 1));      // nbr_cnt
    if (p.delta) add(4ull * kDeltaRows * p.stride);  // dlist
    add(sizeof(Partial) * std::max<uint64_t>(p.nblocks, 1));
    add(sizeof(BoxPartial) * std::max<uint64_t>(p.nblocks, 1));
    if (p.dim == 2) {
        add(8 * n);  // cell_prev
        add(4 * n);  // entry
    }
    if (p.unsorted) add(16 * n);  // cell_last
    return b;
}

void carve(fsa_scene_s* s, const Plan& p) {
    char* base = (char*)s->dmem;
    uint64_t off = 0;
    const uint64_t n = std::max<uint64_t>(p.n, 1);
    auto take = [&](uint64_t sz) {
        void* r = base + off;
        off += align_up(sz, 256);
        return r;
    };
    SceneDev& h = s->hd;
    for (int k = 0; k < 2; k++) h.pos[k] = (float4*)take(16 * n);
    for (int k = 0; k < 2; k++) h.vel[k] = (float4*)take(16 * n);
    for (int k = 0; k < 2; k++) h.acc[k] = (float4*)take(16 * n);
    for (int k = 0; k < 2; k++) h.ext[k] = (float4*)take(16 * n);
    for (int k = 0; k < 2; k++) h.pres[k] = (float*)take(4 * n);
    for (int k = 0; k < 2; k++) h.err[k] = (float*)take(4 * n);
    for (int k = 0; k < 2; k++) h.idt[k] = (uint32_t*)take(4 * n);
    h.rho_new = (float*)take(4 * n);
    h.vstar = (float4*)take(16 * n);
    h.anp = (float4*)take(16 * n);
    h.src = (float*)take(4 * n);
    h.diag = (float*)take(4 * n);
    h.dii = (float4*)take(16 * std::max<uint64_t>(p.nf, 1));
    for (int k = 0; k < 2; k++) h.keys[k] = (uint32_t*)take(4 * n);
    for (int k = 0; k < 2; k++) h.vals[k] = (uint32_t*)take(4 * n);
    h.hist = (uint32_t*)take(4 * (kRadixMaxDigits * std::max<uint64_t>(p.ntiles, 1) + kRadixMaxDigits));
    h.fluid_idx = (uint32_t*)take(4 * std::max<uint64_t>(p.nf, 1));
    h.rank_of = (uint32_t*)take(4 * n);
    h.block_count = (uint32_t*)take(4 * std::max<uint64_t>(p.nblocks, 1));
    h.nbr = (uint32_t*)take(4ull * kMaxNeighbors * p.stride);
    h.nbr_cnt = (uint32_t*)take(4 * std::max<uint64_t>(p.nf, 1));
    h.dlist = p.delta ? (uint32_t*)take(4ull * kDeltaRows * p.stride) : nullptr;
    h.partial = (Partial*)take(sizeof(Partial) * std::max<uint64_t>(p.nblocks, 1));
    h.box = (BoxPartial*)take(sizeof(BoxPartial) * std::max<uint64_t>(p.nblocks, 1));
    if (p.dim == 2) {
        h.cell_prev = (int2*)take(8 * n);
        h.entry = (uint32_t*)take(4 * n);
    }
    if (p.unsorted) h.cell_last = (int4*)take(16 * n);
    h.n = (uint32_t)p.n;
    h.nf = (uint32_t)p.nf;
    h.list_stride = (uint32_t)p.stride;
    h.nblocks = (uint32_t)p.nblocks;
    h.ntiles = (uint32_t)p.ntiles;
}

// ---- block maps --------------------------------------------------------------------
struct MapBuilder {
    std::vector<BlockRef> all, fluid, tiles, scan, scene, cnt;
    bool fluidless = false;  // some scene has no fluid block
    void add(uint32_t sidx, const SceneDev& h) {
        for (uint32_t b = 0; b < kCountBlocks; b++) cnt.push_back({sidx, b});
        for (uint32_t b = 0; b < h.nblocks; b++) all.push_back({sidx, b});
        const uint32_t nfb = (h.nf + kBlock - 1) / kBlock;
        for (uint32_t b = 0; b < nfb; b++) fluid.push_back({sidx, b});
        fluidless = fluidless || nfb == 0;
        for (uint32_t b = 0; b < h.ntiles; b++) tiles.push_back({sidx, b});
        for (uint32_t d = 0; d < (uint32_t)kRadix; d++) scan.push_back({sidx, d});
        scene.push_back({sidx, 0});
    }
};

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup i runs on XCD
// i mod 8), and each XCD has its own L2.  Reordering a block map so that the
// workgroups of one XCD take one contiguous range of blocks keeps the neighbor
// gathers of a block, which hit the blocks next to it in the sorted order, in
// the L2 that already holds them.  Every per-block result is indexed by the
// logical block, so the order of dispatch does not change any result.
constexpr uint32_t kXcds = 8;

// chunk = 0: each XCD takes one contiguous eighth of the map (the fluid maps,
// whose blocks all carry the same kind of work).  chunk = C: runs of C blocks
// go to the XCDs in turn, which keeps neighbors together and still spreads
// regions of uneven work (boundary-only cells) over all XCDs.
std::vector<BlockRef> xcd_order(const std::vector<BlockRef>& v, size_t chunk) {
    const size_t n = v.size();
    if (n <= kXcds) return v;
    if (chunk == 0) chunk = (n + kXcds - 1) / kXcds;
    // the logical blocks each XCD should run, in order
    std::vector<std::vector<size_t>> want(kXcds);
    for (size_t c = 0; c * chunk < n; c++)
        for (size_t b = c * chunk; b < std::min(n, (c + 1) * chunk); b++) want[c % kXcds].push_back(b);
    // workgroup i runs on XCD i mod 8; an XCD whose list runs dry takes from the longest other list
    std::vector<size_t> next(kXcds, 0);
    std::vector<BlockRef> out(n);
    for (size_t i = 0; i < n; i++) {
        size_t k = i % kXcds;
        if (next[k] == want[k].size()) {
            size_t best = k;
            for (size_t q = 0; q < kXcds; q++)
                if (want[q].size() - next[q] > want[best].size() - next[best]) best = q;
            k = best;
        }
        out[i] = v[want[k][next[k]++]];
    }
    return out;
}

// Block maps are copied on the caller's stream (a synchronous copy on the null
// stream would break a graph capture that another thread runs meanwhile).
BlockRef* upload_map(const std::vector<BlockRef>& v, std::vector<BlockRef*>& owned, hipStream_t st) {
    if (v.empty()) return nullptr;
    BlockRef* d = nullptr;
    HIP_CHECK(hipMalloc(&d, v.size() * sizeof(BlockRef)));
    owned.push_back(d);
    HIP_CHECK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(BlockRef), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
    return d;
}

void build_launch_set(const MapBuilder& mb, SceneDev* table, LaunchSet& L, std::vector<BlockRef*>& owned,
                      hipStream_t st) {
    L.scenes = table;
    L.map_all = upload_map(xcd_order(mb.all, 16), owned, st);
    L.map_fluid = upload_map(xcd_order(mb.fluid, 0), owned, st);
    L.map_tiles = upload_map(mb.tiles, owned, st);
    L.map_scan = upload_map(mb.scan, owned, st);
    L.map_scene = upload_map(mb.scene, owned, st);
    L.map_cnt = upload_map(mb.cnt, owned, st);
    L.nb_cnt = (uint32_t)mb.cnt.size();
    L.fluidless_scene = mb.fluidless;
    L.nb_all = (uint32_t)mb.all.size();
    L.nb_fluid = (uint32_t)mb.fluid.size();
    L.nb_tiles = (uint32_t)mb.tiles.size();
    L.nb_scan = (uint32_t)mb.scan.size();
    L.nb_scene = (uint32_t)mb.scene.size();
}

// ---- grid ---------------------------------------------------------------------------
inline int32_t host_cell(float x, float r) { return (int32_t)std::floor(x / r); }

// per grid cell: the cell table entry, its rank, the counting sort's bucket count and offset
constexpr uint64_t kCellBytes = sizeof(uint4) + 3 * sizeof(uint32_t);
constexpr uint64_t kCellExtra = sizeof(uint32_t) * (1 + kCountBlocks);  // the last offset, the part sums
uint64_t cell_alloc_bytes(uint64_t cap) { return cap * kCellBytes + kCellExtra; }
uint32_t bits_of(uint64_t v);
// passes of the compressed search's sort over compact keys of L levels (3L + 3 bits)
uint32_t key_passes(uint32_t L) { return (3 * L + 3 + kSearchSortBits - 1) / kSearchSortBits; }

uint32_t levels_for(const int32_t lo[3], const int32_t hi[3]) {
    uint32_t L = 1;
    for (;;) {
        const int64_t lim = 1ll << L;
        bool ok = true;
        for (int a = 0; a < 3; a++)
            if (lo[a] < -lim || hi[a] > lim - 1) ok = false;
        if (ok) return L;
        L++;
        if (L > 9) return L;
    }
}

// sets the cell grid of the coming search from the state's bounding box
void prepare_grid(fsa_scene_s* s) {
    SceneDev& h = s->hd;
    const SceneState& st = *s->h_state;
    int32_t lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = st.bbox_min[a];
        hi[a] = st.bbox_max[a];
    }
    if (s->n == 0) {
        for (int a = 0; a < 3; a++) lo[a] = hi[a] = 0;
    }
    uint32_t L = 0;
    if (h.dim == 3 && (h.flags & kFlagUnsorted)) {
        uint64_t cells = 1;
        for (int a = 0; a < 3; a++) cells *= (uint64_t)(hi[a] - (int64_t)lo[a] + 1);
        if (cells > (1ull << 28)) throw FsaError(FSA_E_INVALID, "3D scene extent exceeds 2^28 grid cells");
    } else if (h.dim == 3) {
        L = levels_for(lo, hi);
        if (L > 9)
            throw FsaError(FSA_E_INVALID,
                           "scene extent exceeds the 32-bit cell key range (cell coordinates must lie in [-512, 511])");
    } else {
        lo[2] = hi[2] = 0;
        if ((uint64_t)(hi[0] - (int64_t)lo[0] + 1) * (uint64_t)(hi[1] - (int64_t)lo[1] + 1) > (1ull << 28))
            throw FsaError(FSA_E_INVALID, "2D scene extent exceeds 2^28 grid cells");
    }
    uint64_t cells = 1;
    for (int a = 0; a < 3; a++) {
        h.cmin[a] = lo[a];
        h.cdim[a] = (uint32_t)(hi[a] - lo[a] + 1);
        cells *= h.cdim[a];
    }
    h.cdim[3] = (uint32_t)cells;
    h.levels = L;
    if (cells > s->cells_cap) {
        if (s->cells) {
            HIP_CHECK(hipFree(s->cells));
            s->ctx->in_use -= cell_alloc_bytes(s->cells_cap);
        }
        const uint64_t cap = cells + cells / 4 + 64;
        if (s->ctx->budget && s->ctx->in_use + cell_alloc_bytes(cap) > s->ctx->budget)
            throw FsaError(FSA_E_BUDGET, "cell table exceeds the HBM budget");
        HIP_CHECK(hipMalloc(&s->cells, cell_alloc_bytes(cap)));
        HIP_CHECK(hipMemsetAsync(s->cells, 0, cap * sizeof(uint4), stream_of(s->ctx)));
        s->cells_cap = cap;
        s->ranks = (uint32_t*)(s->cells + cap);
        s->bcnt = s->ranks + cap;
        s->boff = s->bcnt + cap;      // cap + 1 entries
        s->bsum = s->boff + cap + 1;  // kCountBlocks entries
        HIP_CHECK(hipMemsetAsync(s->bcnt, 0, cap * sizeof(uint32_t), stream_of(s->ctx)));
        s->ranks_valid = false;
        s->ctx->in_use += cell_alloc_bytes(cap);
    }
    h.cells = s->cells;
    // The compressed search sorts by each cell's rank among the grid's cells
    // (the compact-key order in fewer bits) when that saves a sort pass; the
    // rank table is rebuilt when the grid changes.
    if (s->radix_steps) s->radix_steps--;
    h.rank_keys = 0;
    h.ranks = s->ranks;
    h.bcnt = s->bcnt;
    h.boff = s->boff;
    h.bsum = s->bsum;
    if (h.dim == 3 && !(h.flags & kFlagUnsorted)) {
        const uint32_t rank_bits = bits_of(cells - 1);
        const uint32_t rank_passes = std::max<uint32_t>(1, (rank_bits + kSearchSortBits - 1) / kSearchSortBits);
        if (rank_passes < key_passes(L)) {
            h.rank_keys = 1;
            bool same = s->ranks_valid && s->rank_levels == L;
            for (int a = 0; a < 3; a++) same = same && s->rank_cmin[a] == h.cmin[a] && s->rank_cdim[a] == h.cdim[a];
            if (!same) {
                launch_rank_table(s->ranks, h.cmin, h.cdim, L, stream_of(s->ctx));
                for (int a = 0; a < 3; a++) {
                    s->rank_cmin[a] = h.cmin[a];
                    s->rank_cdim[a] = h.cdim[a];
                }
                s->rank_levels = L;
                s->ranks_valid = true;
            }
        }
    }
    h.stamp++;
    if (h.stamp == 0) {
        HIP_CHECK(hipMemsetAsync(s->cells, 0, s->cells_cap * sizeof(uint4), stream_of(s->ctx)));
        h.stamp = 1;
    }
}

uint32_t spec_of(const SceneDev& h) {
    uint32_t sp = 0;
    if (h.flags & kFlagHOne) sp |= kSpecHOne;
    if (h.flags & kFlagRho0SqOne) sp |= kSpecRho0SqOne;
    // the shared-reciprocal division and the short sqrt correction are exact for
    // h == 1 as long as no coordinate is tiny (2D runs its own kernels)
    if (h.dim == 3 && (h.flags & kFlagHOne) && !(h.flags & kFlagTinyCoords)) sp |= kSpecFastMath;
    if (h.dim == 3 && (h.flags & kFlagDeltaLists)) sp |= kSpecDelta;
    return sp;
}

bool tiny_coordinate(float v) {
    const float a = std::fabs(v);
    return a > 0.0f && a < kTinyCoord;
}

uint32_t sort_passes(const SceneDev& h) {
    if (h.rank_keys)
        return std::max<uint32_t>(1, (bits_of((uint64_t)h.cdim[3] - 1) + kSearchSortBits - 1) / kSearchSortBits);
    return key_passes(h.levels);
}

void read_state(fsa_scene_s* s) {
    HIP_CHECK(hipMemcpyAsync(s->h_state, s->hd.state, sizeof(SceneState), hipMemcpyDeviceToHost, stream_of(s->ctx)));
    HIP_CHECK(hipStreamSynchronize(stream_of(s->ctx)));
}

// one copy of the pool range covering every scene of the list
void read_states(fsa_context ctx, const std::vector<fsa_scene_s*>& v) {
    if (v.empty()) return;
    uint32_t lo = UINT32_MAX, hi = 0;
    for (auto* s : v) {
        lo = std::min(lo, s->slot);
        hi = std::max(hi, s->slot);
    }
    HIP_CHECK(hipMemcpyAsync(ctx->h_states + lo, ctx->d_states + lo, sizeof(SceneState) * (hi - lo + 1),
                             hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

// ---- timing ---------------------------------------------------------------------------
// Phase timing records HIP events on the context stream without waiting; the
// elapsed times are collected when the pool fills or when they are read.
void collect_timing(fsa_context ctx) {
    if (ctx->ev_used == 0) return;
    // a nested phase's pair follows its enclosing one's, so its stop event is
    // not the last one recorded: wait for every stop event
    for (size_t k = 1; k < ctx->ev_used; k += 2) HIP_CHECK(hipEventSynchronize(ctx->ev_pool[k]));
    for (size_t k = 0; k + 1 < ctx->ev_used; k += 2) {
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev_pool[k], ctx->ev_pool[k + 1]));
        ctx->timer_ms[ctx->ev_phase[k / 2]] += ms;
        ctx->timer_n[ctx->ev_phase[k / 2]] += 1;
    }
    ctx->ev_used = 0;
}

// A phase takes its start/stop pair when it opens, so one phase may enclose
// another (the list kernel inside the search).  Only an outermost phase may
// collect the pool, and it leaves room for one nested pair.
struct PhaseTimer {
    fsa_context ctx;
    int id;
    size_t slot = 0;
    bool on;
    PhaseTimer(fsa_context c, int i, bool nested = false) : ctx(c), id(i), on(c->timing) {
        if (!on) return;
        if (!nested && ctx->ev_used + 4 > ctx->ev_pool.size()) collect_timing(ctx);
        slot = ctx->ev_used;
        ctx->ev_used += 2;
        ctx->ev_phase[slot / 2] = id;
        HIP_CHECK(hipEventRecord(ctx->ev_pool[slot], ctx->stream));
    }
    ~PhaseTimer() noexcept(false) {
        if (!on) return;
        HIP_CHECK(hipEventRecord(ctx->ev_pool[slot + 1], ctx->stream));
    }
};

// ---- the step, for any set of scenes sharing a launch set --------------------------------
void search_set(fsa_context ctx, const LaunchSet& L, uint32_t passes) {
    PhaseTimer t(ctx, FSA_TIMER_SEARCH);
    uint32_t sbuf = passes & 1u;
    if (L.counting) {
        launch_count_sort(L, ctx->stream);
        sbuf = 1u;
    } else {
        launch_keys_sort(L, passes, ctx->stream);
    }
    launch_reorder(L, sbuf, ctx->stream);
    {
        PhaseTimer tl(ctx, FSA_TIMER_LISTS, true);
        launch_lists(L, sbuf, ctx->stream);
    }
    if (L.codec) launch_codec(L, ctx->stream, false);
}

void simulate_set(fsa_context ctx, const LaunchSet& L, std::vector<fsa_scene_s*>& scenes, bool sesph,
                  bool boundary_pass) {
    hipStream_t st = ctx->stream;
    const bool d2 = !scenes.empty() && scenes[0]->hd.dim == 2;
    if (L.entities_before) launch_entities_before(L, st);
    if (sesph) {
        if (d2) {
            PhaseTimer t(ctx, FSA_TIMER_DENSITY);
            launch2_sesph(L, st);
        } else {
            {
                PhaseTimer t(ctx, FSA_TIMER_DENSITY);
                launch_sesph_density(L, st);
            }
            {
                PhaseTimer t(ctx, FSA_TIMER_DIAGONAL);
                launch_sesph_accel(L, st);
            }
        }
    } else {
        {
            PhaseTimer t(ctx, FSA_TIMER_DENSITY);
            if (d2)
                launch2_iisph_prepare(L, st);
            else
                launch_iisph_prepare(L, boundary_pass, st);
        }
        // 3D: the loop state is set before phase 2 (by phase 1, and for scenes
        // without fluid here), which performs the first Jacobi iteration's
        // pressure update on the way
        if (!d2) launch_iisph_begin_loop(L, st);
        {
            PhaseTimer t(ctx, FSA_TIMER_DIAGONAL);
            if (d2)
                launch2_iisph_diagonal(L, st);
            else
                launch_iisph_diagonal(L, st);
        }
        if (d2) launch2_iisph_begin_loop(L, st);
        uint32_t enq = 0, max_it = 0, guess = 0;
        for (auto* s : scenes) {
            max_it = std::max(max_it, s->hd.max_it);
            guess = std::max(guess, std::max(s->hd.min_it, s->last_iterations));
        }
        uint32_t chunk = std::min(max_it, std::max<uint32_t>(guess, 1));
        while (enq < max_it) {
            for (uint32_t k = 0; k < chunk && enq < max_it; k++, enq++) {
                if (!d2 && enq == 0) {
                    // every pressure is zero: phase 2 did the first iteration's update
                    PhaseTimer t(ctx, FSA_TIMER_PRESSURE_FIRST);
                    launch_iisph_first_converge(L, st);
                    continue;
                }
                {
                    PhaseTimer t(ctx, FSA_TIMER_PRESSURE_A);
                    if (d2)
                        launch2_iisph_pass_a(L, st);
                    else
                        launch_iisph_pass_a(L, st);
                }
                {
                    PhaseTimer t(ctx, FSA_TIMER_PRESSURE_B);
                    if (d2)
                        launch2_iisph_pass_b(L, st);
                    else
                        launch_iisph_pass_b(L, st);
                }
            }
            if (enq >= max_it) break;
            bool any_active = false;
            read_states(ctx, scenes);
            for (auto* s : scenes)
                if (s->h_state->active) any_active = true;
            if (!any_active) break;
            chunk = 4;
        }
    }
    {
        PhaseTimer t(ctx, FSA_TIMER_INTEGRATE);
        if (d2)
            launch2_integrate(L, st);
        else
            launch_integrate(L, st);
    }
    if (L.remover) launch_remover_check(L, st);
    HIP_CHECK(hipStreamSynchronize(st));
}

void fill_stats(fsa_scene_s* s, fsa_step_stats* out) {
    const SceneState& st = *s->h_state;
    if (!out) return;
    out->desired_time_step = s->hd.dt;
    out->actual_time_step = st.dt_actual;
    out->iterations = s->desc.solver == FSA_SOLVER_SESPH ? 0u : st.iterations;
    out->status = FSA_OK;
    if (st.status & kStatusCapacity) out->status = FSA_E_CAPACITY;
    out->average_density_error = st.avg_err;
    out->max_velocity = st.max_v;
    out->max_acceleration = st.max_a;
}

void check_status(fsa_scene_s* s) {
    const uint32_t status = s->h_state->status;
    if (status & kStatusCapacity)
        throw FsaError(FSA_E_CAPACITY, "a particle has more neighbors than a list holds (97)");
    if (status & kStatusGrid) throw FsaError(FSA_E_STATE, "a particle left the cell grid (internal error)");
}

// neighbor-list byte offsets need more than 32 bits
bool wide_lists(const SceneDev& h) { return 4ull * kMaxNeighbors * h.list_stride > 0xffffffffull; }

// the counting sort serves a launch set whose scenes all sort rank keys and
// none is back on the radix sort after a retry
bool counting_sort_ok(const fsa_scene_s* s) { return s->hd.rank_keys && s->radix_steps == 0; }

// the first Jacobi iteration cannot end the loop: its check can be left out
bool first_check_needless(const SceneDev& h) { return h.min_it >= 2 && h.max_it >= 2; }

void upload_scene_table(fsa_scene_s* s) {
    s->single.spec = spec_of(s->hd);
    s->single.counting = counting_sort_ok(s);
    s->single.skip_first_check = first_check_needless(s->hd);
    s->single.codec = (s->hd.flags & kFlagCodec) != 0;
    s->single.wide_lists = wide_lists(s->hd);
    s->single.entities_before = (s->hd.entity_bits & (kEntityBitBoundary | kEntityBitVelocity)) != 0;
    s->single.remover = (s->hd.entity_bits & kEntityBitRemover) != 0;
    HIP_CHECK(hipMemcpyAsync(s->d_table, &s->hd, sizeof(SceneDev), hipMemcpyHostToDevice, stream_of(s->ctx)));
}

uint32_t bits_of(uint64_t v) {
    uint32_t b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

// 2D: radix passes of the entry-count sort and of the cell sort
uint32_t entry_passes(const SceneDev& h) {
    if (h.flags & kFlagQuadratic) return 0;
    return (bits_of(h.search_count) + kRadixBits - 1) / kRadixBits;
}
uint32_t cell_passes(const SceneDev& h) {
    return std::max<uint32_t>(1, (bits_of(h.cdim[3]) + kRadixBits - 1) / kRadixBits);
}

// 3D hashed / quadratic: radix passes of the cell sort
uint32_t unsorted_passes(const SceneDev& h) {
    return std::max<uint32_t>(1, (bits_of(h.cdim[3]) + kRadixBits - 1) / kRadixBits);
}

void search3u_set(fsa_context ctx, const LaunchSet& L, uint32_t passes) {
    PhaseTimer t(ctx, FSA_TIMER_SEARCH);
    launch3u_search(L, passes, ctx->stream);
}

// libmorton morton3D_64_encode of the cell key (CompressedNeighbors.cpp:43-56,
// HashedNeighborhoodSearch3D.cpp:236-244): coordinate + 2^31, x bits at 3k
// (22 kept), y at 3k + 1 and z at 3k + 2 (21 kept)

// table sizes of the device quick sort for r rows
struct QsSizes {
    uint64_t nseg, nchunk, nsmall, nblk;
    explicit QsSizes(uint64_t r)
        : nseg(r / (kQsSmall + 1) + 2), nchunk(r / kQsChunk + r / (kQsSmall + 1) + 3), nsmall(r / 2 + 1),
          nblk((r + 255) / 256 + 1) {}
};

// device bytes of the hashed search's re-sort scratch for `rows` collection rows
uint64_t quick_sort_bytes(uint64_t rows) {
    const uint64_t r = std::max<uint64_t>(rows, 1);
    const QsSizes z(r);
    uint64_t bytes = 0;
    auto add = [&](uint64_t sz) { bytes += align_up(sz, 256); };
    add(8 * r);                              // key
    for (int k = 0; k < 6; k++) add(4 * r);  // pay, apos, bpos, ref x2, inv
    add(4 * z.nblk);                         // block counts
    for (int k = 0; k < 2; k++) add(sizeof(QsSeg) * z.nseg);
    for (int k = 0; k < 2; k++) add(sizeof(QsChunk) * z.nchunk);
    add(sizeof(uint2) * z.nsmall);
    add(sizeof(QsCtr) * 3);
    return bytes;
}

// Scratch of the device quick sort for `rows` collection rows (kept for the scene's life).
void ensure_quick_sort(fsa_scene_s* s, uint64_t rows) {
    if (s->qs_mem && s->qs_rows >= rows) return;
    const uint64_t r = std::max<uint64_t>(rows, 1);
    const QsSizes z(r);
    const uint64_t nseg = z.nseg, nchunk = z.nchunk, nsmall = z.nsmall, nblk = z.nblk;
    const uint64_t bytes = quick_sort_bytes(r);
    if (s->ctx->budget && s->ctx->in_use - s->qs_bytes + bytes > s->ctx->budget)
        throw FsaError(FSA_E_BUDGET, "the hashed search's re-sort does not fit the HBM budget");
    void* m = nullptr;
    if (hipMalloc(&m, bytes) != hipSuccess) throw FsaError(FSA_E_NOMEM, "device allocation failed");
    if (s->qs_mem) {
        // the order on the device moves to the new scratch
        if (s->qs_ref_newer) {
            s->ref_order.resize(s->qs_rows);
            HIP_CHECK(hipMemcpyAsync(s->ref_order.data(), s->qs_ref[s->qs_cur], 4 * s->qs_rows, hipMemcpyDeviceToHost,
                                     stream_of(s->ctx)));
            HIP_CHECK(hipStreamSynchronize(stream_of(s->ctx)));
            s->qs_ref_newer = false;
        }
        (void)hipFree(s->qs_mem);
        s->ctx->in_use -= s->qs_bytes;
    }
    s->qs_mem = m;
    s->qs_bytes = bytes;
    s->qs_rows = r;
    s->ctx->in_use += bytes;
    char* base = (char*)m;
    uint64_t off = 0;
    auto take = [&](uint64_t sz) {
        void* p = base + off;
        off += align_up(sz, 256);
        return p;
    };
    s->qs.key = (uint64_t*)take(8 * r);
    s->qs.pay = (uint32_t*)take(4 * r);
    s->qs.apos = (uint32_t*)take(4 * r);
    s->qs.bpos = (uint32_t*)take(4 * r);
    s->qs_ref[0] = (uint32_t*)take(4 * r);
    s->qs_ref[1] = (uint32_t*)take(4 * r);
    s->qs_inv = (uint32_t*)take(4 * r);
    s->qs_block_cnt = (uint32_t*)take(4 * nblk);
    for (int k = 0; k < 2; k++) s->qs.seg[k] = (QsSeg*)take(sizeof(QsSeg) * nseg);
    for (int k = 0; k < 2; k++) s->qs.chunk[k] = (QsChunk*)take(sizeof(QsChunk) * nchunk);
    s->qs.small = (uint2*)take(sizeof(uint2) * nsmall);
    s->qs.ctr = (QsCtr*)take(sizeof(QsCtr) * 3);
    s->qs.seg_cap = (uint32_t)nseg;
    s->qs.chunk_cap = (uint32_t)nchunk;
    s->qs.small_cap = (uint32_t)nsmall;
    s->qs_cur = 0;
    s->qs_ref_valid = false;
}

// Sort::quick_sort (ParticleCollectionAlgorithm.hpp:28-127) of qs.key / qs.pay
// [0, rows): parts above kQsSmall elements are partitioned level by level on
// the device, each level's split writing the next level's tables; the host
// checks every kQsLevelsPerCheck levels whether parts are left, then the small
// parts are sorted one per lane.
constexpr uint32_t kQsLevelsPerCheck = 4;

void device_quick_sort(fsa_scene_s* s, uint64_t rows) {
    hipStream_t st = stream_of(s->ctx);
    const uint32_t hi = (uint32_t)(rows - 1);
    QsCtr ctr[3] = {{0u, 0u}, {0u, 0u}, {0u, 0u}};
    std::vector<QsChunk> chunks;
    QsSeg top{0u, hi, 0u, 0u, 0u, 0u, 0u, 0u};
    uint2 part = make_uint2(0u, hi);
    if (rows > kQsSmall) {
        for (uint64_t lo = 0; lo <= hi; lo += kQsChunk)
            chunks.push_back(QsChunk{0u, (uint32_t)lo, (uint32_t)std::min<uint64_t>(hi, lo + kQsChunk - 1), 0u, 0u, 0u,
                                     0u, 0u});
        top.c1 = (uint32_t)chunks.size();
        ctr[0] = QsCtr{1u, (uint32_t)chunks.size()};
        HIP_CHECK(hipMemcpyAsync(s->qs.seg[0], &top, sizeof(QsSeg), hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(s->qs.chunk[0], chunks.data(), sizeof(QsChunk) * chunks.size(), hipMemcpyHostToDevice,
                                 st));
    } else {
        ctr[2] = QsCtr{1u, 0u};
        HIP_CHECK(hipMemcpyAsync(s->qs.small, &part, sizeof(uint2), hipMemcpyHostToDevice, st));
    }
    HIP_CHECK(hipMemcpyAsync(s->qs.ctr, ctr, sizeof(ctr), hipMemcpyHostToDevice, st));
    uint32_t cur = 0;
    uint64_t levels = 0;
    while (ctr[cur].nseg != 0) {
        for (uint32_t k = 0; k < kQsLevelsPerCheck; k++, cur ^= 1u) launch_qs_level(s->qs, cur, st);
        levels += kQsLevelsPerCheck;
        HIP_CHECK(hipMemcpyAsync(ctr, s->qs.ctr, sizeof(ctr), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        // every level splits each part, so rows levels always suffice
        if (ctr[2].nchunk != 0 || levels > rows + kQsLevelsPerCheck)
            throw FsaError(FSA_E_DEVICE, "the hashed search's re-sort failed (internal error)");
    }
    launch_qs_small(s->qs, ctr[2].nseg, st);
}

// HashedNeighborhoodSearch3D::improve_cache_efficiency (HashedNeighborhoodSearch3D.cpp:219-247):
// on every 101st call the collection, inactive particles included, is re-sorted
// by the Morton key of each particle's cell from the previous grid update
// (GridCellState.current; INT_MIN, key 0, for particles never updated).  The
// device computes the reference quick sort's permutation over every row, keeps
// the new collection order, and applies the resident rows' share of it.
void hashed_resort_if_due(fsa_scene_s* s) {
    if (search_of(&s->desc) != FSA_SEARCH_HASHED_3D) return;
    s->hashed_calls++;
    if (s->hashed_calls <= 100) return;
    s->hashed_calls = 0;
    const uint32_t n = s->n;
    const uint64_t N = s->total;
    if (N < 2) return;
    hipStream_t st = stream_of(s->ctx);
    ensure_quick_sort(s, N);
    read_state(s);
    if (!s->qs_ref_valid) {
        HIP_CHECK(hipMemcpyAsync(s->qs_ref[s->qs_cur], s->ref_order.data(), 4 * N, hipMemcpyHostToDevice, st));
        s->qs_ref_valid = true;
    }
    launch_qs_keys(s->hd.idt[s->h_state->cur], n, s->hd.cell_last, s->qs_ref[s->qs_cur], (uint32_t)N, s->qs_inv,
                   s->qs.key, s->qs.pay, st);
    device_quick_sort(s, N);
    launch_qs_apply(s->qs, s->qs_ref[s->qs_cur], (uint32_t)N, s->qs_inv, s->qs_ref[s->qs_cur ^ 1u], s->qs_block_cnt,
                    s->hd.vals[0], st);
    s->qs_cur ^= 1u;
    s->qs_ref_newer = true;
    if (n) {
        upload_scene_table(s);
        launch_permute_only(s->single, 0u, st);
    }
    HIP_CHECK(hipStreamSynchronize(st));
}

void search2d_set(fsa_context ctx, const LaunchSet& L, uint32_t epasses, uint32_t cpasses) {
    PhaseTimer t(ctx, FSA_TIMER_SEARCH);
    launch2_search(L, epasses, cpasses, ctx->stream);
}

// A counting sort that met a bucket it does not order left the scene as it was
// before the search (kStatusSortRetry): the flag is cleared and the scene takes
// the radix sort for kRadixRetrySteps steps.  True when the search must be redone.
constexpr uint32_t kRadixRetrySteps = 64;
bool take_sort_retry(fsa_scene_s* s) {
    if (!(s->h_state->status & kStatusSortRetry)) return false;
    s->h_state->status &= ~kStatusSortRetry;
    HIP_CHECK(hipMemcpyAsync(&s->hd.state->status, &s->h_state->status, sizeof(uint32_t), hipMemcpyHostToDevice,
                             stream_of(s->ctx)));
    HIP_CHECK(hipStreamSynchronize(stream_of(s->ctx)));
    s->radix_steps = kRadixRetrySteps;
    return true;
}

void do_find_neighbors(fsa_scene_s* s) {
    hashed_resort_if_due(s);
    prepare_grid(s);
    upload_scene_table(s);
    if (s->hd.dim == 2) {
        search2d_set(s->ctx, s->single, entry_passes(s->hd), cell_passes(s->hd));
        s->hd.search_count++;
    } else if (s->hd.flags & kFlagUnsorted) {
        s->hd.vals_buf = unsorted_passes(s->hd) & 1u;
        search3u_set(s->ctx, s->single, unsorted_passes(s->hd));
    } else {
        search_set(s->ctx, s->single, sort_passes(s->hd));
    }
    read_state(s);
    if (take_sort_retry(s)) {
        do_find_neighbors(s);  // the radix sort now
        return;
    }
    check_status(s);
    s->searched = true;
    s->lists_valid = true;
    s->lists_cached = false;
}

void set_step_dt(fsa_scene_s* s, float dt);
void apply_removal(fsa_scene_s* s);

void pre_simulate(fsa_scene_s* s, float dt) {
    if (!s->searched) throw FsaError(FSA_E_STATE, "simulate called before find_neighbors");
    if (!(dt > 0.0f)) throw FsaError(FSA_E_INVALID, "desired time step must be > 0");
    set_step_dt(s, dt);
}

// The desired time step of the next step (Simulator.cpp:17-24):
// ConstantTimestepGenerator: desc.timestep.  DynamicCflTimestepGenerator::generate_next_timestep
// (DynamicCflTimestepGenerator.cpp:46-62): from the largest |v| and |a| of the
// active particles, with get_non_cfl_validating_timestep (:84-109) falling
// back to max_timestep; sqrt of a float there is ::sqrt(double).
float generated_timestep(const fsa_scene_s* s) {
    const fsa_scene_desc& d = s->desc;
    if (d.timestep_generator != FSA_TIMESTEP_DYNAMIC_CFL) return d.timestep;
    const float mv = std::sqrt(s->cur_v2), ma = std::sqrt(s->cur_a2);
    float ts = d.min_timestep;
    const float eps = std::numeric_limits<float>::epsilon();
    if (mv > eps && ma > eps) {
        float dta = d.max_timestep, dtv = d.max_timestep;
        if (ma > 0.0f) dta = (float)(std::sqrt((double)(d.particle_size / ma)) * (double)d.lambda_a);
        if (mv > 0.0f) dtv = d.particle_size / mv * d.lambda_v;
        const float allowed = std::fmax(std::fmin(dta, dtv), 0.000001f);  // MIN_ALLOWED_TIMESTEP
        ts = std::max(ts, std::min(d.max_timestep, allowed));
    }
    return ts;
}

// time step of the coming solver step, and the SESPH CFL share of inactive particles
void set_step_dt(fsa_scene_s* s, float dt) {
    s->hd.dt = dt;
    if (s->desc.solver == FSA_SOLVER_SESPH) {
        // SESPH's CFL scan covers inactive particles too (SESPHFluidSolver3D.hpp:96-107)
        float a2 = 0.0f, v2 = 0.0f;
        for (size_t k = 0; k < s->inactive_rows.size(); k++) {
            const float* m = &s->inactive_movement[9 * k];
            const float ax = m[6], ay = m[7], az = m[8];
            a2 = std::fmax(a2, (ax * ax + ay * ay) + az * az);
            const float vx = m[3] + dt * ax, vy = m[4] + dt * ay, vz = m[5] + dt * az;
            v2 = std::fmax(v2, (vx * vx + vy * vy) + vz * vz);
        }
        s->hd.inactive_a2 = a2;
        s->hd.inactive_v2 = v2;
    }
}

// steps in a row whose Jacobi loop ended within the predicted passes before the
// gated extra pass is left out
constexpr uint32_t kSteadySteps = 4;

// after the step's final state read
void post_simulate(fsa_scene_s* s, fsa_step_stats* stats) {
    s->boundary_dirty = false;
    if (s->h_state->status & kStatusRemoval) s->removal_pending = true;
    // velocity alterations reach inactive particles too; they are held on the host
    for (const auto& e : s->entities)
        if (e.type == FSA_ENTITY_VELOCITY_ALTERATION_BY_TAG)
            for (size_t k = 0; k < s->inactive_rows.size(); k++)
                if (s->row_tags[s->inactive_rows[k]] == e.tag)
                    for (int a = 0; a < 3; a++) s->inactive_movement[9 * k + 3 + a] = e.velocity[a];
    s->cur_v2 = s->h_state->next_v2;
    s->cur_a2 = s->h_state->next_a2;
    if (s->h_state->status & kStatusTinyCoord) s->hd.flags |= kFlagTinyCoords;
    if (s->desc.solver == FSA_SOLVER_IISPH) {
        s->last_iterations = std::max<uint32_t>(s->h_state->iterations, 1);
        s->steady_steps = s->h_state->iterations <= s->step_chunk ? s->steady_steps + 1 : 0;
    }
    fill_stats(s, stats);
    s->searched = false;
    check_status(s);
}

void do_simulate(fsa_scene_s* s, float dt, fsa_step_stats* stats) {
    pre_simulate(s, dt);
    upload_scene_table(s);
    std::vector<fsa_scene_s*> one{s};
    const bool bpass = (s->hd.flags & kFlagDynamicBoundary) || s->boundary_dirty;
    simulate_set(s->ctx, s->single, one, s->desc.solver == FSA_SOLVER_SESPH, bpass);
    read_state(s);
    post_simulate(s, stats);
    if (s->removal_pending) apply_removal(s);
}

// ---- batching -------------------------------------------------------------------------
void free_batch(BatchSet* b);

// Launch tables are cached per scene set; a serving loop changes the set as
// jobs come and go, so the cache keeps the most recently used sets only.
constexpr size_t kBatchCacheSize = 16;

BatchSet* batch_for(fsa_context ctx, const std::vector<fsa_scene_s*>& v) {
    // the cache is also pruned by scene destruction, which a server tick may run
    // on another thread while this one steps
    std::lock_guard<std::mutex> g(ctx->mu);
    auto it = ctx->batches.find(v);
    if (it != ctx->batches.end()) {
        it->second->last_use = ++ctx->batch_clock;
        return it->second;
    }
    while (ctx->batches.size() >= kBatchCacheSize) {
        auto lru = ctx->batches.begin();
        for (auto jt = ctx->batches.begin(); jt != ctx->batches.end(); ++jt)
            if (jt->second->last_use < lru->second->last_use) lru = jt;
        free_batch(lru->second);
        ctx->batches.erase(lru);
    }
    auto* b = new BatchSet();
    b->last_use = ++ctx->batch_clock;
    b->scenes = v;
    b->h_table.resize(v.size());
    HIP_CHECK(hipMalloc(&b->d_table, sizeof(SceneDev) * v.size()));
    MapBuilder mb;
    for (uint32_t q = 0; q < v.size(); q++) mb.add(q, v[q]->hd);
    build_launch_set(mb, b->d_table, b->ls, b->owned, ctx->stream);
    ctx->batches[v] = b;
    return b;
}

void free_batch(BatchSet* b) {
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    (void)hipFree(b->d_table);
    for (auto* m : b->owned) (void)hipFree(m);
    delete b;
}

void drop_batches_of(fsa_context ctx, fsa_scene_s* s) {
    for (auto it = ctx->batches.begin(); it != ctx->batches.end();) {
        if (std::find(it->first.begin(), it->first.end(), s) != it->first.end()) {
            free_batch(it->second);
            it = ctx->batches.erase(it);
        } else {
            ++it;
        }
    }
}

void upload_batch_table(fsa_context ctx, BatchSet* b) {
    uint32_t spec = kSpecHOne | kSpecRho0SqOne | kSpecFastMath | kSpecDelta;  // the specialisations every scene allows
    bool codec = false, wide = false, ent = false, rem = false, counting = true, skip_first = true;
    for (uint32_t q = 0; q < b->scenes.size(); q++) {
        counting = counting && counting_sort_ok(b->scenes[q]);
        skip_first = skip_first && first_check_needless(b->scenes[q]->hd);
        b->h_table[q] = b->scenes[q]->hd;
        spec &= spec_of(b->scenes[q]->hd);
        codec = codec || (b->scenes[q]->hd.flags & kFlagCodec);
        wide = wide || wide_lists(b->scenes[q]->hd);
        ent = ent || (b->scenes[q]->hd.entity_bits & (kEntityBitBoundary | kEntityBitVelocity));
        rem = rem || (b->scenes[q]->hd.entity_bits & kEntityBitRemover);
    }
    b->ls.entities_before = ent;
    b->ls.remover = rem;
    b->ls.spec = spec;
    b->ls.codec = codec;
    b->ls.wide_lists = wide;
    b->ls.counting = counting;
    b->ls.skip_first_check = skip_first;
    if (ctx->walk_variant)
        for (int f = 0; f < kTuneFamilies; f++) b->ls.unroll[f] = ctx->walk_variant;
    HIP_CHECK(hipMemcpyAsync(b->d_table, b->h_table.data(), sizeof(SceneDev) * b->h_table.size(),
                             hipMemcpyHostToDevice, ctx->stream));
}

// One step of every scene in v (same solver and dimension) in shared launches.
// The head of the step (search, solver phases, the predicted number of gated
// Jacobi iterations) is one launch sequence, replayed from a hipGraph when
// graphs are enabled; the host then checks the scenes once, runs further
// iterations for scenes still converging, integrates and reads the states.
// A scene whose search fails (neighbor capacity) is skipped by the solver
// kernels: its stats carry the error, the other scenes complete the step, and
// the first error is returned.
// ---- autotuning ----------------------------------------------------------------------------
// The neighbor-walk unroll of the four IISPH kernel families is chosen per size
// class (kernel spec, total fluid particles rounded to a power of two) by
// timing every candidate on the live neighbor lists of the first step of that
// class.  The timed launches read the real particle arrays and lists but write
// to a shadow scene table whose written arrays and state are scratch copies,
// so the simulation state is untouched.
uint64_t tune_class(const BatchSet* b, const std::vector<fsa_scene_s*>& v) {
    uint64_t nf = 0;
    for (auto* s : v) nf += s->nf;
    return (uint64_t)b->ls.spec << 32 | bits_of(nf);
}

void autotune(fsa_context ctx, BatchSet* b, const std::vector<fsa_scene_s*>& v) {
    const uint64_t cls = tune_class(b, v);
    auto hit = ctx->tuned.find(cls);
    if (hit != ctx->tuned.end()) {
        for (int f = 0; f < kTuneFamilies; f++) b->ls.unroll[f] = hit->second.unroll[f];
        return;
    }
    hipStream_t st = ctx->stream;
    // scratch per scene: acc, vstar, anp, ext, dii (16 B) and pres, err, rho_new, src, diag (4 B)
    // per particle, partials per block, one state
    std::vector<uint64_t> offs(v.size());
    uint64_t total = 0;
    for (size_t q = 0; q < v.size(); q++) {
        offs[q] = total;
        const uint64_t n = std::max<uint32_t>(v[q]->n, 1);
        total += align_up(4 * 16 * n, 256) + align_up(5 * 4 * n, 256) + align_up(16 * n, 256) +
                 align_up(sizeof(Partial) * std::max<uint32_t>(v[q]->hd.nblocks, 1), 256) +
                 align_up(sizeof(SceneState), 256);
    }
    if (ctx->budget && ctx->in_use.load() + total > ctx->budget) return;  // no room: keep the defaults
    char* scratch = nullptr;
    if (hipMalloc(&scratch, total) != hipSuccess) return;
    SceneDev* table = nullptr;
    if (hipMalloc(&table, sizeof(SceneDev) * v.size()) != hipSuccess) {
        (void)hipFree(scratch);
        return;
    }
    std::vector<SceneDev> sh(v.size());
    for (size_t q = 0; q < v.size(); q++) {
        const SceneDev& h = v[q]->hd;
        const uint32_t c = v[q]->h_state->cur;
        const uint64_t n = std::max<uint32_t>(v[q]->n, 1);
        char* p = scratch + offs[q];
        auto take = [&](uint64_t sz) {
            char* r = p;
            p += align_up(sz, 256);
            return r;
        };
        SceneDev& d = sh[q];
        d = h;
        float4* acc = (float4*)take(4 * 16 * n);
        float4* vstar = acc + n;
        float4* anp = acc + 2 * n;
        float4* ext = acc + 3 * n;
        float* f4 = (float*)take(5 * 4 * n);
        d.acc[0] = d.acc[1] = acc;
        d.vstar = vstar;
        d.anp = anp;
        d.ext[0] = d.ext[1] = ext;
        d.pres[0] = d.pres[1] = f4;
        d.err[0] = d.err[1] = f4 + n;
        d.rho_new = f4 + 2 * n;
        d.src = f4 + 3 * n;
        d.diag = f4 + 4 * n;
        d.dii = (float4*)take(16 * n);
        d.partial = (Partial*)take(sizeof(Partial) * std::max<uint32_t>(h.nblocks, 1));
        d.state = (SceneState*)take(sizeof(SceneState));
        // realistic inputs: copies of the step's arrays, a state that keeps the loop active
        const uint64_t nn = v[q]->n;
        if (nn) {
            HIP_CHECK(hipMemcpyAsync(acc, h.acc[c], 16 * nn, hipMemcpyDeviceToDevice, st));
            HIP_CHECK(hipMemcpyAsync(vstar, h.vstar, 16 * nn, hipMemcpyDeviceToDevice, st));
            HIP_CHECK(hipMemcpyAsync(anp, h.anp, 16 * nn, hipMemcpyDeviceToDevice, st));
            HIP_CHECK(hipMemcpyAsync(ext, h.ext[c], 16 * nn, hipMemcpyDeviceToDevice, st));
            HIP_CHECK(hipMemcpyAsync(d.pres[0], h.pres[c], 4 * nn, hipMemcpyDeviceToDevice, st));
            HIP_CHECK(hipMemcpyAsync(d.err[0], h.err[c], 4 * nn, hipMemcpyDeviceToDevice, st));
            HIP_CHECK(hipMemcpyAsync(d.rho_new, h.rho_new, 4 * nn, hipMemcpyDeviceToDevice, st));
            HIP_CHECK(hipMemcpyAsync(d.src, h.src, 4 * nn, hipMemcpyDeviceToDevice, st));
            HIP_CHECK(hipMemcpyAsync(d.diag, h.diag, 4 * nn, hipMemcpyDeviceToDevice, st));
            if (v[q]->nf) HIP_CHECK(hipMemcpyAsync(d.dii, h.dii, 16ull * v[q]->nf, hipMemcpyDeviceToDevice, st));
        }
        SceneState ss = *v[q]->h_state;
        ss.active = 1;
        ss.status = 0;
        HIP_CHECK(hipMemcpyAsync(d.state, &ss, sizeof(SceneState), hipMemcpyHostToDevice, st));
        HIP_CHECK(hipStreamSynchronize(st));  // ss is a stack copy
    }
    HIP_CHECK(hipMemcpyAsync(table, sh.data(), sizeof(SceneDev) * v.size(), hipMemcpyHostToDevice, st));
    LaunchSet L = b->ls;
    L.scenes = table;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    fsa_context_s::Tuned t{};
    t.fluid = 0;
    for (auto* s : v) t.fluid += s->nf;
    for (int f = 0; f < kTuneFamilies; f++) {
        float best = 0.0f;
        uint8_t pick = kVariantU4;
        for (int c = 0; c < kTuneVariants; c++) {
            L.unroll[f] = kTuneVariant[c];
            launch_tuned_kernel(L, f, st);  // warm-up
            HIP_CHECK(hipEventRecord(e0, st));
            for (int r = 0; r < 3; r++) launch_tuned_kernel(L, f, st);
            HIP_CHECK(hipEventRecord(e1, st));
            HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            t.us[f][c] = 1000.0f * ms / 3.0f;
            if (c == 0 || t.us[f][c] < best) {
                best = t.us[f][c];
                pick = kTuneVariant[c];
            }
        }
        L.unroll[f] = pick;
        t.unroll[f] = pick;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(table);
    (void)hipFree(scratch);
    for (int f = 0; f < kTuneFamilies; f++) b->ls.unroll[f] = t.unroll[f];
    ctx->tuned[cls] = t;
}

// 3D: integration gated on the loop state (it acts only on scenes whose Jacobi loop has
// ended and that have not integrated this step)
void launch_gated_integrate(fsa_context ctx, const LaunchSet& L) {
    PhaseTimer t(ctx, FSA_TIMER_INTEGRATE);
    launch_integrate(L, ctx->stream);
}

// The launch sequence of a step up to the first host check.  3D scenes take the
// predicted Jacobi passes, `tail` more passes that act only on scenes still
// iterating, and the gated integration, so a step whose loops end within them
// needs a single host synchronisation.
void step_head(fsa_context ctx, const LaunchSet& L, bool d2, bool unsorted, uint32_t passes, uint32_t ep,
               uint32_t cp, bool sesph, bool bpass, uint32_t chunk, uint32_t tail) {
    hipStream_t st = ctx->stream;
    if (d2)
        search2d_set(ctx, L, ep, cp);
    else if (unsorted)
        search3u_set(ctx, L, passes);
    else
        search_set(ctx, L, passes);
    if (L.entities_before) launch_entities_before(L, ctx->stream);
    if (sesph) {
        PhaseTimer t(ctx, FSA_TIMER_DENSITY);
        if (d2) {
            launch2_sesph(L, st);
        } else {
            launch_sesph_density(L, st);
            launch_sesph_accel(L, st);
        }
        if (!d2) launch_gated_integrate(ctx, L);
        return;
    }
    {
        PhaseTimer t(ctx, FSA_TIMER_DENSITY);
        if (d2)
            launch2_iisph_prepare(L, st);
        else
            launch_iisph_prepare(L, bpass, st);
    }
    // 3D: the loop state is set before phase 2 (by phase 1, and for scenes
    // without fluid here), which performs the first Jacobi iteration's pressure
    // update on the way
    if (!d2) launch_iisph_begin_loop(L, st);
    {
        PhaseTimer t(ctx, FSA_TIMER_DIAGONAL);
        if (d2)
            launch2_iisph_diagonal(L, st);
        else
            launch_iisph_diagonal(L, st);
    }
    if (d2) launch2_iisph_begin_loop(L, st);
    for (uint32_t k = 0; k < chunk + (d2 ? 0u : tail); k++) {
        if (!d2 && k == 0) {
            // every pressure is zero: phase 2 did the first iteration's update
            PhaseTimer t(ctx, FSA_TIMER_PRESSURE_FIRST);
            launch_iisph_first_converge(L, st);
            continue;
        }
        if (k >= chunk) {
            // the gated pass: timed apart, since it is a no-op when the prediction held
            PhaseTimer t(ctx, FSA_TIMER_PRESSURE_TAIL);
            launch_iisph_pass_a(L, st);
            launch_iisph_pass_b(L, st);
            continue;
        }
        {
            PhaseTimer t(ctx, FSA_TIMER_PRESSURE_A);
            if (d2)
                launch2_iisph_pass_a(L, st);
            else
                launch_iisph_pass_a(L, st);
        }
        {
            PhaseTimer t(ctx, FSA_TIMER_PRESSURE_B);
            if (d2)
                launch2_iisph_pass_b(L, st);
            else
                launch_iisph_pass_b(L, st);
        }
    }
    if (!d2) launch_gated_integrate(ctx, L);
}

int32_t batch_step_group(fsa_context ctx, const std::vector<fsa_scene_s*>& all, std::vector<fsa_step_stats*>& all_out) {
    // Per-scene set-up before any launch (the hashed re-sort, the cell grid of the
    // coming search): a scene that fails here (its cells leave the key range, the
    // cell table exceeds the budget) fails alone with its status in its stats,
    // and the other scenes of the group take the step.
    int32_t first_error = FSA_OK;
    std::vector<fsa_scene_s*> v;
    std::vector<fsa_step_stats*> out;
    for (uint32_t q = 0; q < all.size(); q++) {
        fsa_scene_s* s = all[q];
        try {
            hashed_resort_if_due(s);
            prepare_grid(s);
        } catch (const FsaError& e) {
            s->searched = false;
            s->lists_valid = false;
            if (all_out[q]) {
                fill_stats(s, all_out[q]);
                all_out[q]->status = (uint32_t)e.code;
            }
            if (first_error == FSA_OK) {
                first_error = e.code;
                g_error = e.what();
            }
            continue;
        }
        v.push_back(s);
        out.push_back(all_out[q]);
    }
    if (v.empty()) return first_error;
    BatchSet* b = batch_for(ctx, v);
    hipStream_t st = ctx->stream;
    const bool d2 = v[0]->hd.dim == 2;
    const bool unsorted = (v[0]->hd.flags & kFlagUnsorted) != 0;
    const bool sesph = v[0]->desc.solver == FSA_SOLVER_SESPH;
    uint32_t passes = 1, epasses = 0, cpasses = 1, max_it = 0, guess = 0;
    bool bpass = false;
    for (auto* s : v) {
        if (d2) {
            epasses = std::max(epasses, entry_passes(s->hd));
            cpasses = std::max(cpasses, cell_passes(s->hd));
        } else if (unsorted) {
            passes = std::max(passes, unsorted_passes(s->hd));
        } else {
            passes = std::max(passes, sort_passes(s->hd));
        }
        set_step_dt(s, generated_timestep(s));
        bpass = bpass || (s->hd.flags & kFlagDynamicBoundary) || s->boundary_dirty;
        max_it = std::max(max_it, s->hd.max_it);
        guess = std::max(guess, std::max(s->hd.min_it, s->last_iterations));
    }
    const uint32_t chunk = sesph ? 0u : std::min(max_it, std::max<uint32_t>(guess, 1));
    // one gated pass beyond the prediction (3D IISPH), when the iteration limit allows
    // it, unless every scene's loop has ended within the prediction for kSteadySteps
    // steps in a row (a loop that runs longer then takes the host's further passes)
    bool steady = true;
    for (auto* s : v) {
        s->step_chunk = chunk;
        if (s->steady_steps < kSteadySteps) steady = false;
    }
    const uint32_t tail = (sesph || d2 || steady) ? 0u : std::min<uint32_t>(1u, max_it - chunk);
    for (auto* s : v) s->hd.vals_buf = passes & 1u;  // the unsorted searches' final bucket buffer
    upload_batch_table(ctx, b);
    if (ctx->graph && !ctx->timing) {
        const std::array<uint32_t, 5> key = {
            d2 ? (epasses << 8 | cpasses) : passes, chunk | tail << 16, b->ls.spec,
            (bpass ? 1u : 0u) | (sesph ? 2u : 0u) | (unsorted ? 4u : 0u) | (b->ls.entities_before ? 8u : 0u) |
                (b->ls.codec ? 16u : 0u) | (b->ls.wide_lists ? 32u : 0u) | (b->ls.remover ? 64u : 0u) |
                (b->ls.counting ? 128u : 0u) | (b->ls.skip_first_check ? 256u : 0u),
            (uint32_t)b->ls.unroll[0] | (uint32_t)b->ls.unroll[1] << 8 | (uint32_t)b->ls.unroll[2] << 16 |
                (uint32_t)b->ls.unroll[3] << 24};
        auto it = b->graphs.find(key);
        if (it == b->graphs.end()) {
            hipGraph_t g = nullptr;
            hipGraphExec_t exec = nullptr;
            HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            step_head(ctx, b->ls, d2, unsorted, passes, epasses, cpasses, sesph, bpass, chunk, tail);
            HIP_CHECK(hipStreamEndCapture(st, &g));
            const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) throw FsaError(FSA_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            it = b->graphs.emplace(key, exec).first;
        }
        HIP_CHECK(hipGraphLaunch(it->second, st));
    } else {
        step_head(ctx, b->ls, d2, unsorted, passes, epasses, cpasses, sesph, bpass, chunk, tail);
    }
    if (d2)
        for (auto* s : v) s->hd.search_count++;
    read_states(ctx, v);
    // scenes whose counting sort asked for a retry took no part in the step (every
    // kernel after the sort skipped them): they step again below with the radix sort
    std::vector<bool> retry(v.size(), false);
    bool any_retry = false;
    for (uint32_t q = 0; q < v.size(); q++)
        if (v[q]->h_state->status & kStatusSortRetry) retry[q] = any_retry = true;
    if (!b->tuned && !any_retry) {
        if (!d2 && !sesph && ctx->autotune && !ctx->walk_variant) autotune(ctx, b, v);
        b->tuned = true;
    }
    if (!sesph) {
        // scenes still converging: further iterations, four between checks (3D: then the
        // gated integration of the scenes whose loops have ended)
        uint32_t enq = chunk + tail;
        while (enq < max_it) {
            bool any_active = false;
            for (auto* s : v)
                if (s->h_state->active) any_active = true;
            if (!any_active) break;
            for (uint32_t k = 0; k < 4 && enq < max_it; k++, enq++) {
                {
                    PhaseTimer t(ctx, FSA_TIMER_PRESSURE_A);
                    if (d2)
                        launch2_iisph_pass_a(b->ls, st);
                    else
                        launch_iisph_pass_a(b->ls, st);
                }
                {
                    PhaseTimer t(ctx, FSA_TIMER_PRESSURE_B);
                    if (d2)
                        launch2_iisph_pass_b(b->ls, st);
                    else
                        launch_iisph_pass_b(b->ls, st);
                }
            }
            if (!d2) launch_gated_integrate(ctx, b->ls);
            read_states(ctx, v);
        }
    }
    if (d2) {
        PhaseTimer t(ctx, FSA_TIMER_INTEGRATE);
        launch2_integrate(b->ls, st);
    }
    // the remover checks the integrated positions; 3D steps otherwise end at the check above
    if (b->ls.remover) launch_remover_check(b->ls, st);
    if (d2 || b->ls.remover) read_states(ctx, v);
    for (uint32_t q = 0; q < v.size(); q++) {
        fsa_scene_s* s = v[q];
        if (retry[q]) continue;
        try {
            check_status(s);
        } catch (const FsaError& e) {
            s->searched = false;
            s->lists_valid = false;
            if (out[q]) {
                fill_stats(s, out[q]);
                out[q]->status = (uint32_t)e.code;
            }
            if (first_error == FSA_OK) {
                first_error = e.code;
                g_error = e.what();
            }
            continue;
        }
        s->searched = true;
        s->lists_valid = true;
        s->lists_cached = false;
        post_simulate(s, out[q]);
    }
    // particle removers act after the step; the launch tables of this group are not used again
    for (uint32_t q = 0; q < v.size(); q++)
        if (!retry[q] && v[q]->removal_pending) apply_removal(v[q]);
    if (any_retry) {
        std::vector<fsa_scene_s*> again;
        std::vector<fsa_step_stats*> again_out;
        for (uint32_t q = 0; q < v.size(); q++) {
            if (!retry[q]) continue;
            (void)take_sort_retry(v[q]);
            again.push_back(v[q]);
            again_out.push_back(out[q]);
        }
        const int32_t rc = batch_step_group(ctx, again, again_out);
        if (first_error == FSA_OK) first_error = rc;
    }
    return first_error;
}

// ---- upload / download ------------------------------------------------------------------
// 2D: every particle on the device in the caller's order, z = 0
// (MovementData {vec2 position, velocity, acceleration}, ExternalForces {vec2})
// The context's pinned staging buffer of at least `bytes` (the caller holds pin_mu).
char* pinned_staging(fsa_context ctx, uint64_t bytes) {
    if (ctx->pin_bytes < bytes) {
        const uint64_t want = std::max<uint64_t>(bytes, ctx->pin_bytes + ctx->pin_bytes / 2);
        if (ctx->pin) HIP_CHECK(hipHostFree(ctx->pin));
        ctx->pin = nullptr;
        ctx->pin_bytes = 0;
        if (hipHostMalloc(&ctx->pin, want, hipHostMallocDefault) != hipSuccess) {
            ctx->pin = nullptr;
            throw FsaError(FSA_E_NOMEM, "pinned host allocation failed");
        }
        ctx->pin_bytes = want;
    }
    return (char*)ctx->pin;
}

// f(lo, hi, part) over [0, n) in up to kHostParts contiguous parts on host
// threads (one part below kHostSerial items); parts write disjoint outputs.
constexpr unsigned kHostParts = 8;
constexpr uint64_t kHostSerial = 1u << 16;
template <class F>
void host_parallel(uint64_t n, F&& f) {
    unsigned parts = std::min(kHostParts, std::max(1u, std::thread::hardware_concurrency()));
    parts = (unsigned)std::min<uint64_t>(parts, std::max<uint64_t>(1, n / (kHostSerial / 2)));
    if (n < kHostSerial || parts <= 1) {
        f((uint64_t)0, n, 0u);
        return;
    }
    std::vector<std::thread> th;
    for (unsigned t = 1; t < parts; t++) th.emplace_back([&, t] { f(n * t / parts, n * (t + 1) / parts, t); });
    f((uint64_t)0, n / parts, 0u);
    for (auto& x : th) x.join();
}

void do_upload2d(fsa_scene_s* s, const fsa_particle_view* v, bool first) {
    const uint64_t N = v->count;
    if (N >= (1ull << kTypeShift)) throw FsaError(FSA_E_INVALID, "too many particles (max 2^30)");
    std::vector<uint32_t> fidx;
    uint32_t nb = 0;
    for (uint64_t i = 0; i < N; i++) {
        const uint8_t t = ((const uint8_t*)v->info + i * v->info_stride)[4];
        if (t > 2) throw FsaError(FSA_E_INVALID, "unknown particle type");
        if (t == FSA_TYPE_NORMAL) fidx.push_back((uint32_t)i);
        if (t == FSA_TYPE_BOUNDARY) nb++;
    }
    const uint32_t nf = (uint32_t)fidx.size();
    if (!first && (N != s->total || nf != s->nf || nb != s->nb))
        throw FsaError(FSA_E_INVALID, "upload must keep the particle count and types of the scene");
    s->total = N;
    s->n = (uint32_t)N;
    s->nf = nf;
    s->nb = nb;
    const uint32_t n = s->n;
    std::vector<float4> pos(n), vel(n), acc(n), ext(n);
    std::vector<float> pres(n), zero(n, 0.0f);
    std::vector<uint32_t> idt(n), rank(n, 0);
    int32_t lo[2] = {INT32_MAX, INT32_MAX}, hi[2] = {INT32_MIN, INT32_MIN};
    const float r = s->desc.particle_size * 2.0f;
    bool any_ext = false;
    for (uint32_t i = 0; i < n; i++) {
        const float* m = (const float*)((const char*)v->movement + (uint64_t)i * v->movement_stride);
        const float* pd = (const float*)((const char*)v->particle_data + (uint64_t)i * v->particle_data_stride);
        const uint8_t t = ((const uint8_t*)v->info + (uint64_t)i * v->info_stride)[4];
        float e[2] = {0.0f, 0.0f};
        if (v->external) {
            const float* ep = (const float*)((const char*)v->external + (uint64_t)i * v->external_stride);
            e[0] = ep[0];
            e[1] = ep[1];
        }
        pos[i] = make_float4(m[0], m[1], 0.0f, pd[0]);
        vel[i] = make_float4(m[2], m[3], 0.0f, pd[2]);
        acc[i] = make_float4(m[4], m[5], 0.0f, 0.0f);
        ext[i] = make_float4(e[0], e[1], 0.0f, 0.0f);
        pres[i] = pd[1];
        idt[i] = i | ((uint32_t)t << kTypeShift);
        if (e[0] != 0.0f || e[1] != 0.0f) any_ext = true;
        // GetGridCellByPosition: truncation toward zero
        const int32_t c[2] = {(int32_t)(m[0] / r), (int32_t)(m[1] / r)};
        for (int a = 0; a < 2; a++) {
            lo[a] = std::min(lo[a], c[a]);
            hi[a] = std::max(hi[a], c[a]);
        }
    }
    for (uint32_t k = 0; k < nf; k++) rank[fidx[k]] = k;
    s->inactive_rows.clear();
    SceneDev& h = s->hd;
    h.flags &= ~(kFlagExternal | kFlagDynamicBoundary);
    if (any_ext) h.flags |= kFlagExternal;
    h.grid_count = (h.flags & kFlagQuadratic) ? n : (uint32_t)(uint16_t)n;  // particleAmount_t (uint16) count
    hipStream_t st = stream_of(s->ctx);
    SceneState init{};
    init.dt_actual = s->desc.timestep;
    for (int a = 0; a < 2; a++) {
        init.bbox_min[a] = n ? lo[a] : 0;
        init.bbox_max[a] = n ? hi[a] : 0;
    }
    if (n) {
        HIP_CHECK(hipMemcpyAsync(h.pos[0], pos.data(), 16ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.vel[0], vel.data(), 16ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.acc[0], acc.data(), 16ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.ext[0], ext.data(), 16ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.pres[0], pres.data(), 4ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.idt[0], idt.data(), 4ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.rank_of, rank.data(), 4ull * n, hipMemcpyHostToDevice, st));
        // rho_new holds the densities of the step; start from the uploaded ones
        std::vector<float> dens(n);
        for (uint32_t i = 0; i < n; i++) dens[i] = vel[i].w;
        HIP_CHECK(hipMemcpyAsync(h.rho_new, dens.data(), 4ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.src, zero.data(), 4ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.diag, zero.data(), 4ull * n, hipMemcpyHostToDevice, st));
        if (first) {
            // the grid starts empty: every particle's previous cell is the reference's
            // {INT32_MAX, INT32_MAX} sentinel (HashedNeighborhoodSearch.cpp:122-126)
            std::vector<int2> prev(n, make_int2(INT32_MAX, INT32_MAX));
            std::vector<uint32_t> ent(n, 0u);
            HIP_CHECK(hipMemcpyAsync(h.cell_prev, prev.data(), 8ull * n, hipMemcpyHostToDevice, st));
            HIP_CHECK(hipMemcpyAsync(h.entry, ent.data(), 4ull * n, hipMemcpyHostToDevice, st));
            h.search_count = 0;
            HIP_CHECK(hipStreamSynchronize(st));
        }
    }
    if (nf) HIP_CHECK(hipMemcpyAsync(h.fluid_idx, fidx.data(), 4ull * nf, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(h.state, &init, sizeof(SceneState), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
    *s->h_state = init;
    s->searched = false;
    s->boundary_dirty = true;
    s->lists_cached = false;
    s->lists_valid = false;
}

// DynamicCflTimestepGenerator::calculate_maximum_velocity_and_acceleration
// (DynamicCflTimestepGenerator.cpp:6-43) over an uploaded view, as squared
// lengths: sqrt is correctly rounded and monotone, so the square root of the
// largest squared length is the largest (float)glm::length.
void upload_maxima(fsa_scene_s* s, const fsa_particle_view* v) {
    float v2 = 0.0f, a2 = 0.0f;
    const bool d3 = s->desc.dimension == 3;
    s->row_tags.resize(v->count);
    s->row_types.resize(v->count);
    for (uint64_t i = 0; i < v->count; i++) {
        const uint8_t* info = (const uint8_t*)v->info + i * v->info_stride;
        std::memcpy(&s->row_tags[i], info, 4);
        s->row_types[i] = info[4];
    }
    for (uint64_t i = 0; i < v->count; i++) {
        if (((const uint8_t*)v->info + i * v->info_stride)[4] == FSA_TYPE_INACTIVE) continue;
        const float* m = (const float*)((const char*)v->movement + i * v->movement_stride);
        const float vv = d3 ? (m[3] * m[3] + m[4] * m[4]) + m[5] * m[5] : m[2] * m[2] + m[3] * m[3];
        const float aa = d3 ? (m[6] * m[6] + m[7] * m[7]) + m[8] * m[8] : m[4] * m[4] + m[5] * m[5];
        v2 = std::fmax(v2, vv);
        a2 = std::fmax(a2, aa);
    }
    s->cur_v2 = v2;
    s->cur_a2 = a2;
}

// CubicSplineKernel3D::GetKernelValue from |x| on the host, with the device's
// operation order (the scene's h and alpha)
float host_kernel_w(const SceneDev& h, float len) {
    const float q = (h.flags & kFlagHOne) ? len : len / h.h;
    const float a = 2.0f - q, b = 1.0f - q;
    const float a3 = a * a * a;
    const float inner = a3 - 4.0f * (b * b * b);
    const float ret = q < 1.0f ? inner : (q < 2.0f ? a3 : 0.0f);
    return h.alpha * ret;
}

// Device copies and bits of the scene's entity list: tags by caller row for
// velocity alterations, BoundaryPreprocessor::initialize's gamma_1 and mass
// cap (BoundaryPreprocessor.cpp:50-62, 96-117: a perfectly sampled plane of
// 5 x 5 particles), and the dynamic-boundary flag when a velocity alteration
// reaches boundary particles.
void refresh_entities(fsa_scene_s* s) {
    SceneDev& h = s->hd;
    uint32_t bits = 0;
    for (const auto& e : s->entities) {
        if (e.type == FSA_ENTITY_BOUNDARY_PREPROCESSOR) bits |= kEntityBitBoundary;
        if (e.type == FSA_ENTITY_VELOCITY_ALTERATION_BY_TAG) bits |= kEntityBitVelocity;
        if (e.type == FSA_ENTITY_PARTICLE_REMOVER_3D) bits |= kEntityBitRemover;
    }
    h.entity_bits = bits;
    h.n_entities = (uint32_t)s->entities.size();
    h.entities = s->d_entities;
    if (bits & kEntityBitVelocity) {
        const uint64_t rows = std::max<size_t>(s->row_tags.size(), 1);
        if (!s->d_tags || s->d_tags_cap < rows) {
            if (s->d_tags) (void)hipFree(s->d_tags);
            s->d_tags = nullptr;
            if (hipMalloc(&s->d_tags, 4 * rows) != hipSuccess) {
                s->d_tags = nullptr;
                throw FsaError(FSA_E_NOMEM, "device allocation failed");
            }
            s->d_tags_cap = rows;
        }
        if (!s->row_tags.empty())
            HIP_CHECK(hipMemcpyAsync(s->d_tags, s->row_tags.data(), 4 * s->row_tags.size(), hipMemcpyHostToDevice,
                                     stream_of(s->ctx)));
    }
    h.tag_of_row = s->d_tags;
    if (bits & kEntityBitBoundary) {
        const float ps = s->desc.particle_size;
        float ideal = 0.0f;
        for (int x = 0; x <= 0; x++)
            for (int y = -2; y <= 2; y++)
                for (int z = -2; z <= 2; z++) {
                    const float px = (float)x * ps, py = (float)y * ps, pz = (float)z * ps;
                    ideal += host_kernel_w(h, std::sqrt((px * px + py * py) + pz * pz));
                }
        h.bp_gamma1 = ((ps * ps) * ps) * ideal;  // Math::pow3(particle_size) * ideal_volume_reciprocal
        h.bp_max_mass = s->desc.rest_density * ((ps * ps) * ps);
    }
    bool moves = false;
    for (const auto& e : s->entities)
        if (e.type == FSA_ENTITY_VELOCITY_ALTERATION_BY_TAG)
            for (size_t r = 0; r < s->row_tags.size() && !moves; r++)
                if (s->row_types[r] == FSA_TYPE_BOUNDARY && s->row_tags[r] == e.tag) moves = true;
    s->entity_moves_boundary = moves;
    if (moves) h.flags |= kFlagDynamicBoundary;
}

// given_order: caller rows in the device order to keep (a relayout after
// particles turned inactive); the particle counts may then change.
void do_upload(fsa_scene_s* s, const fsa_particle_view* v, bool first,
               const std::vector<uint32_t>* given_order = nullptr) {
    if (!v || !v->movement || !v->particle_data || !v->info) throw FsaError(FSA_E_INVALID, "incomplete particle view");
    upload_maxima(s, v);
    if (s->desc.dimension == 2) {
        do_upload2d(s, v, first);
        return;
    }
    const uint64_t N = v->count;
    if (N >= (1ull << kTypeShift)) throw FsaError(FSA_E_INVALID, "too many particles (max 2^30)");
    {
        // neighbor-list entries are addressed with 32-bit indices (row * stride + rank)
        uint64_t nf_all = 0;
        for (uint64_t i = 0; i < N; i++) nf_all += ((const uint8_t*)v->info + i * v->info_stride)[4] == FSA_TYPE_NORMAL;
        if ((uint64_t)kMaxNeighbors * make_plan(N, nf_all).stride > 0xffffffffull)
            throw FsaError(FSA_E_INVALID, "too many fluid particles in one scene (max 44,000,000)");
    }
    std::vector<uint32_t> active;
    active.reserve(N);
    std::vector<uint32_t> inactive;
    uint32_t nf = 0, nb = 0;
    for (uint64_t i = 0; i < N; i++) {
        const uint8_t t = ((const uint8_t*)v->info + i * v->info_stride)[4];
        if (t == FSA_TYPE_INACTIVE) {
            inactive.push_back((uint32_t)i);
            continue;
        }
        if (t > 2) throw FsaError(FSA_E_INVALID, "unknown particle type");
        active.push_back((uint32_t)i);
        if (t == FSA_TYPE_NORMAL)
            nf++;
        else
            nb++;
    }
    if (!first && !given_order && (N != s->total || nf != s->nf || nb != s->nb))
        throw FsaError(FSA_E_INVALID, "upload must keep the particle count and types of the scene");
    // Device order of the active particles: the caller's order on the first upload;
    // afterwards the order the reference collection holds by now (its searches
    // re-sort the collection in place), so later sorts start from the same order.
    const bool hashed = search_of(&s->desc) == FSA_SEARCH_HASHED_3D;
    if (first && hashed) {
        s->ref_order.resize(N);
        std::iota(s->ref_order.begin(), s->ref_order.end(), 0u);
        s->hashed_calls = 0;
        s->qs_ref_valid = false;
        s->qs_ref_newer = false;
        ensure_quick_sort(s, N);  // the re-sort's scratch, allocated up front
    }
    if (!first && hashed && s->qs_ref_newer) {
        // the last re-sort left the collection order on the device
        s->ref_order.resize(s->total);
        HIP_CHECK(hipMemcpyAsync(s->ref_order.data(), s->qs_ref[s->qs_cur], 4ull * s->total, hipMemcpyDeviceToHost,
                                 stream_of(s->ctx)));
        HIP_CHECK(hipStreamSynchronize(stream_of(s->ctx)));
        s->qs_ref_newer = false;
    }
    if (!first) {
        std::vector<uint32_t> order;
        if (given_order) {
            for (uint32_t r : *given_order)
                if (r < N && ((const uint8_t*)v->info + (uint64_t)r * v->info_stride)[4] != FSA_TYPE_INACTIVE)
                    order.push_back(r);
        } else if (hashed && s->ref_order.size() == N) {
            for (uint32_t r : s->ref_order)
                if (((const uint8_t*)v->info + (uint64_t)r * v->info_stride)[4] != FSA_TYPE_INACTIVE) order.push_back(r);
        } else if (!(s->hd.flags & kFlagUnsorted) && s->n) {
            read_state(s);
            std::vector<uint32_t> idt(s->n);
            HIP_CHECK(hipMemcpyAsync(idt.data(), s->hd.idt[s->h_state->cur], 4ull * s->n, hipMemcpyDeviceToHost,
                                     stream_of(s->ctx)));
            HIP_CHECK(hipStreamSynchronize(stream_of(s->ctx)));
            for (uint32_t w : idt) order.push_back(w & kIdMask);
        }
        if (order.size() == active.size()) {
            std::vector<uint32_t> a = active, b = order;
            std::sort(b.begin(), b.end());
            if (a == b) active.swap(order);
        }
    }
    s->total = N;
    s->n = (uint32_t)active.size();
    s->nf = nf;
    s->nb = nb;
    const uint32_t n = s->n;
    // The device arrays are packed in the context's pinned staging (pos, vel, acc,
    // ext, pres, err, idt), the rows split over host threads, each with its own
    // flags and cell bounds.
    std::unique_lock<std::mutex> pin_lock(s->ctx->pin_mu);
    // 4 x 16 B + 3 x 4 B per particle
    char* stg = pinned_staging(s->ctx, 76ull * std::max<uint32_t>(n, 1));
    float4* const pos = (float4*)stg;
    float4* const vel = pos + n;
    float4* const acc = vel + n;
    float4* const ext = acc + n;
    float* const pres = (float*)(ext + n);
    float* const errs = pres + n;
    uint32_t* const idt = (uint32_t*)(errs + n);
    struct Part {
        bool any_ext = false, dyn_boundary = false, any_tiny = false;
        int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    };
    Part parts[kHostParts];
    const float r = s->desc.particle_size * 2.0f;
    const bool carry = s->carry_err.size() == N;
    host_parallel(n, [&](uint64_t k0, uint64_t k1, unsigned part) {
        Part P;
        for (uint64_t k = k0; k < k1; k++) {
            const uint64_t i = active[k];
            const float* m = (const float*)((const char*)v->movement + i * v->movement_stride);
            const float* pd = (const float*)((const char*)v->particle_data + i * v->particle_data_stride);
            const uint8_t t = ((const uint8_t*)v->info + i * v->info_stride)[4];
            float e[3] = {0.0f, 0.0f, 0.0f};
            if (v->external) {
                const float* ep = (const float*)((const char*)v->external + i * v->external_stride);
                e[0] = ep[0];
                e[1] = ep[1];
                e[2] = ep[2];
            }
            pos[k] = make_float4(m[0], m[1], m[2], pd[0]);
            vel[k] = make_float4(m[3], m[4], m[5], pd[2]);
            acc[k] = make_float4(m[6], m[7], m[8], 0.0f);
            ext[k] = make_float4(e[0], e[1], e[2], 0.0f);
            pres[k] = pd[1];
            errs[k] = carry ? s->carry_err[i] : 0.0f;
            idt[k] = (uint32_t)i | ((uint32_t)t << kTypeShift);
            if (e[0] != 0.0f || e[1] != 0.0f || e[2] != 0.0f) P.any_ext = true;
            if (tiny_coordinate(m[0]) || tiny_coordinate(m[1]) || tiny_coordinate(m[2])) P.any_tiny = true;
            if (t == FSA_TYPE_BOUNDARY &&
                (m[3] != 0.0f || m[4] != 0.0f || m[5] != 0.0f || m[6] != 0.0f || m[7] != 0.0f || m[8] != 0.0f ||
                 e[0] != 0.0f || e[1] != 0.0f || e[2] != 0.0f))
                P.dyn_boundary = true;
            const int32_t c[3] = {host_cell(m[0], r), host_cell(m[1], r), host_cell(m[2], r)};
            for (int a = 0; a < 3; a++) {
                P.lo[a] = std::min(P.lo[a], c[a]);
                P.hi[a] = std::max(P.hi[a], c[a]);
            }
        }
        parts[part] = P;
    });
    bool any_ext = false, dyn_boundary = false, any_tiny = false;
    int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (const Part& P : parts) {
        any_ext = any_ext || P.any_ext;
        dyn_boundary = dyn_boundary || P.dyn_boundary;
        any_tiny = any_tiny || P.any_tiny;
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], P.lo[a]);
            hi[a] = std::max(hi[a], P.hi[a]);
        }
    }
    // inactive rows are carried on the host unchanged (every solver loop skips them)
    s->inactive_rows = inactive;
    s->inactive_movement.assign(9 * inactive.size(), 0.0f);
    s->inactive_pdata.assign(3 * inactive.size(), 0.0f);
    s->inactive_ext.assign(3 * inactive.size(), 0.0f);
    s->inactive_info.assign(8 * inactive.size(), 0);
    for (size_t k = 0; k < inactive.size(); k++) {
        const uint64_t i = inactive[k];
        std::memcpy(&s->inactive_movement[9 * k], (const char*)v->movement + i * v->movement_stride, 36);
        std::memcpy(&s->inactive_pdata[3 * k], (const char*)v->particle_data + i * v->particle_data_stride, 12);
        std::memcpy(&s->inactive_info[8 * k], (const char*)v->info + i * v->info_stride, 8);
        if (v->external)
            std::memcpy(&s->inactive_ext[3 * k], (const char*)v->external + i * v->external_stride, 12);
    }
    SceneDev& h = s->hd;
    h.flags &= ~(kFlagExternal | kFlagDynamicBoundary | kFlagTinyCoords);
    if (any_ext) h.flags |= kFlagExternal;
    if (dyn_boundary) h.flags |= kFlagDynamicBoundary;
    if (any_tiny) h.flags |= kFlagTinyCoords;
    hipStream_t st = stream_of(s->ctx);
    if ((h.flags & kFlagUnsorted) && n) {
        // the collection keeps its order between searches: fluid ranks are fixed here
        std::vector<uint32_t> fidx, rank(n, 0);
        for (uint32_t k = 0; k < n; k++)
            if ((idt[k] >> kTypeShift) == 0u) {
                rank[k] = (uint32_t)fidx.size();
                fidx.push_back(k);
            }
        if (!fidx.empty())
            HIP_CHECK(hipMemcpyAsync(h.fluid_idx, fidx.data(), 4ull * fidx.size(), hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.rank_of, rank.data(), 4ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipStreamSynchronize(st));
    }
    SceneState init{};
    init.cur = 0;
    init.dt_actual = s->desc.timestep;
    for (int a = 0; a < 3; a++) {
        init.bbox_min[a] = n ? lo[a] : 0;
        init.bbox_max[a] = n ? hi[a] : 0;
    }
    if (n) {
        HIP_CHECK(hipMemcpyAsync(h.pos[0], pos, 16ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.vel[0], vel, 16ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.acc[0], acc, 16ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.ext[0], ext, 16ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.pres[0], pres, 4ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.err[0], errs, 4ull * n, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(h.idt[0], idt, 4ull * n, hipMemcpyHostToDevice, st));
    }
    s->carry_err.clear();
    HIP_CHECK(hipMemcpyAsync(h.state, &init, sizeof(SceneState), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
    pin_lock.unlock();  // the staging copies are done
    *s->h_state = init;
    s->searched = false;
    s->boundary_dirty = true;
    s->lists_cached = false;
    s->lists_valid = false;
    refresh_entities(s);
}

struct HostCopy {
    std::vector<float4> pos, vel, acc, ext;
    std::vector<float> pres, err;
    std::vector<uint32_t> idt;
};

void fetch(fsa_scene_s* s, HostCopy& hc, bool with_err) {
    read_state(s);
    const uint32_t c = s->h_state->cur;
    const uint32_t n = s->n;
    hc.pos.resize(n);
    hc.vel.resize(n);
    hc.acc.resize(n);
    hc.ext.resize(n);
    hc.pres.resize(n);
    hc.idt.resize(n);
    hipStream_t st = stream_of(s->ctx);
    const SceneDev& h = s->hd;
    if (n) {
        HIP_CHECK(hipMemcpyAsync(hc.pos.data(), h.pos[c], 16ull * n, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(hc.vel.data(), h.vel[c], 16ull * n, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(hc.acc.data(), h.acc[c], 16ull * n, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(hc.ext.data(), h.ext[c], 16ull * n, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(hc.pres.data(), h.pres[c], 4ull * n, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(hc.idt.data(), h.idt[c], 4ull * n, hipMemcpyDeviceToHost, st));
        if (with_err) {
            hc.err.resize(n);
            HIP_CHECK(hipMemcpyAsync(hc.err.data(), h.err[c], 4ull * n, hipMemcpyDeviceToHost, st));
        }
    }
    HIP_CHECK(hipStreamSynchronize(st));
}

// 3D download: the live arrays into the context's pinned staging, then the
// rows written in caller order on host threads (each device particle owns its row)
void do_download3d(fsa_scene_s* s, const fsa_particle_view* v) {
    read_state(s);
    const uint32_t c = s->h_state->cur;
    const uint32_t n = s->n;
    const bool ext_live = (s->hd.flags & kFlagExternal) != 0;
    const SceneDev& h = s->hd;
    hipStream_t st = stream_of(s->ctx);
    std::unique_lock<std::mutex> pin_lock(s->ctx->pin_mu);
    // 4 x 16 B + 2 x 4 B per particle
    char* stg = pinned_staging(s->ctx, 72ull * std::max<uint32_t>(n, 1));
    float4* const pos = (float4*)stg;
    float4* const vel = pos + n;
    float4* const acc = vel + n;
    float4* const ext = acc + n;
    float* const pres = (float*)(ext + n);
    uint32_t* const idt = (uint32_t*)(pres + n);
    if (n) {
        HIP_CHECK(hipMemcpyAsync(pos, h.pos[c], 16ull * n, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(vel, h.vel[c], 16ull * n, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(acc, h.acc[c], 16ull * n, hipMemcpyDeviceToHost, st));
        if (ext_live && v->external) HIP_CHECK(hipMemcpyAsync(ext, h.ext[c], 16ull * n, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(pres, h.pres[c], 4ull * n, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(idt, h.idt[c], 4ull * n, hipMemcpyDeviceToHost, st));
    }
    HIP_CHECK(hipStreamSynchronize(st));
    host_parallel(n, [&](uint64_t k0, uint64_t k1, unsigned) {
        for (uint64_t k = k0; k < k1; k++) {
            const uint64_t i = idt[k] & kIdMask;
            const uint8_t t = (uint8_t)(idt[k] >> kTypeShift);
            if (v->movement) {
                float* m = (float*)((char*)v->movement + i * v->movement_stride);
                m[0] = pos[k].x;
                m[1] = pos[k].y;
                m[2] = pos[k].z;
                m[3] = vel[k].x;
                m[4] = vel[k].y;
                m[5] = vel[k].z;
                m[6] = acc[k].x;
                m[7] = acc[k].y;
                m[8] = acc[k].z;
            }
            if (v->particle_data) {
                float* pd = (float*)((char*)v->particle_data + i * v->particle_data_stride);
                pd[0] = pos[k].w;
                pd[1] = pres[k];
                pd[2] = vel[k].w;
            }
            if (v->info) ((uint8_t*)v->info + i * v->info_stride)[4] = t;
            if (v->external) {
                float* e = (float*)((char*)v->external + i * v->external_stride);
                e[0] = ext_live ? ext[k].x : 0.0f;
                e[1] = ext_live ? ext[k].y : 0.0f;
                e[2] = ext_live ? ext[k].z : 0.0f;
            }
        }
    });
    pin_lock.unlock();
    for (size_t k = 0; k < s->inactive_rows.size(); k++) {
        const uint64_t i = s->inactive_rows[k];
        if (v->movement) std::memcpy((char*)v->movement + i * v->movement_stride, &s->inactive_movement[9 * k], 36);
        if (v->particle_data)
            std::memcpy((char*)v->particle_data + i * v->particle_data_stride, &s->inactive_pdata[3 * k], 12);
        if (v->info) std::memcpy((char*)v->info + i * v->info_stride, &s->inactive_info[8 * k], 8);
        if (v->external)
            std::memcpy((char*)v->external + i * v->external_stride, &s->inactive_ext[3 * k], 12);
    }
}

void do_download(fsa_scene_s* s, const fsa_particle_view* v) {
    if (!v || v->count != s->total) throw FsaError(FSA_E_INVALID, "download view count mismatch");
    if (s->desc.dimension == 3) {
        do_download3d(s, v);
        return;
    }
    HostCopy hc;
    fetch(s, hc, false);
    if (s->desc.dimension == 2) {
        for (uint32_t i = 0; i < s->n; i++) {
            if (v->movement) {
                float* m = (float*)((char*)v->movement + (uint64_t)i * v->movement_stride);
                m[0] = hc.pos[i].x;
                m[1] = hc.pos[i].y;
                m[2] = hc.vel[i].x;
                m[3] = hc.vel[i].y;
                m[4] = hc.acc[i].x;
                m[5] = hc.acc[i].y;
            }
            if (v->particle_data) {
                float* pd = (float*)((char*)v->particle_data + (uint64_t)i * v->particle_data_stride);
                pd[0] = hc.pos[i].w;
                pd[1] = hc.pres[i];
                pd[2] = hc.vel[i].w;
            }
            if (v->info) ((uint8_t*)v->info + (uint64_t)i * v->info_stride)[4] = (uint8_t)(hc.idt[i] >> kTypeShift);
            if (v->external) {
                float* e = (float*)((char*)v->external + (uint64_t)i * v->external_stride);
                e[0] = hc.ext[i].x;
                e[1] = hc.ext[i].y;
            }
        }
        return;
    }
    const bool ext_live = (s->hd.flags & kFlagExternal) != 0;
    for (uint32_t k = 0; k < s->n; k++) {
        const uint64_t i = hc.idt[k] & kIdMask;
        const uint8_t t = (uint8_t)(hc.idt[k] >> kTypeShift);
        if (v->movement) {
            float* m = (float*)((char*)v->movement + i * v->movement_stride);
            m[0] = hc.pos[k].x;
            m[1] = hc.pos[k].y;
            m[2] = hc.pos[k].z;
            m[3] = hc.vel[k].x;
            m[4] = hc.vel[k].y;
            m[5] = hc.vel[k].z;
            m[6] = hc.acc[k].x;
            m[7] = hc.acc[k].y;
            m[8] = hc.acc[k].z;
        }
        if (v->particle_data) {
            float* pd = (float*)((char*)v->particle_data + i * v->particle_data_stride);
            pd[0] = hc.pos[k].w;
            pd[1] = hc.pres[k];
            pd[2] = hc.vel[k].w;
        }
        if (v->info) ((uint8_t*)v->info + i * v->info_stride)[4] = t;
        if (v->external) {
            float* e = (float*)((char*)v->external + i * v->external_stride);
            e[0] = ext_live ? hc.ext[k].x : 0.0f;
            e[1] = ext_live ? hc.ext[k].y : 0.0f;
            e[2] = ext_live ? hc.ext[k].z : 0.0f;
        }
    }
    for (size_t k = 0; k < s->inactive_rows.size(); k++) {
        const uint64_t i = s->inactive_rows[k];
        if (v->movement) std::memcpy((char*)v->movement + i * v->movement_stride, &s->inactive_movement[9 * k], 36);
        if (v->particle_data)
            std::memcpy((char*)v->particle_data + i * v->particle_data_stride, &s->inactive_pdata[3 * k], 12);
        if (v->info) std::memcpy((char*)v->info + i * v->info_stride, &s->inactive_info[8 * k], 8);
        if (v->external)
            std::memcpy((char*)v->external + i * v->external_stride, &s->inactive_ext[3 * k], 12);
    }
}

// Re-lays the scene out for a collection whose active set changed: a new plan
// in the same allocation when it fits, new launch maps, and an upload that
// keeps the given device order of the particles that stay active.
void relayout(fsa_scene_s* s, const fsa_particle_view* v, const std::vector<uint32_t>& order) {
    uint64_t nf = 0, nb = 0;
    for (uint64_t i = 0; i < v->count; i++) {
        const uint8_t t = ((const uint8_t*)v->info + i * v->info_stride)[4];
        nf += t == FSA_TYPE_NORMAL;
        nb += t == FSA_TYPE_BOUNDARY;
    }
    const Plan p = make_plan(nf + nb, nf, 3, (s->hd.flags & kFlagUnsorted) != 0, (s->hd.flags & kFlagDeltaLists) != 0);
    const uint64_t bytes = plan_bytes(p);
    HIP_CHECK(hipStreamSynchronize(stream_of(s->ctx)));
    if (bytes > s->bytes) {
        void* m = nullptr;
        if (hipMalloc(&m, bytes) != hipSuccess) throw FsaError(FSA_E_NOMEM, "device allocation failed");
        (void)hipFree(s->dmem);
        s->ctx->in_use -= s->bytes;
        s->ctx->in_use += bytes;
        s->dmem = m;
        s->bytes = bytes;
    }
    carve(s, p);
    for (auto* m : s->map_allocs) (void)hipFree(m);
    s->map_allocs.clear();
    MapBuilder mb;
    mb.add(0, s->hd);
    build_launch_set(mb, s->d_table, s->single, s->map_allocs, stream_of(s->ctx));
    {
        std::lock_guard<std::mutex> g(s->ctx->mu);
        drop_batches_of(s->ctx, s);
    }
    do_upload(s, v, false, &order);
    upload_scene_table(s);
}

// GroupActions::remove over each ParticleRemover3D's VolumeGroup
// (ParticleRemover3D.cpp:8-16; VolumeGroup.cpp:43-58) on the host copy of the
// collection, then a relayout that keeps the device order and the IISPH
// predicted density errors (IISPHParticleData3D persists in the collection).
void apply_removal(fsa_scene_s* s) {
    s->removal_pending = false;
    const uint64_t N = s->total;
    std::vector<float> mv(9 * N), pd(3 * N), ex(3 * N);
    std::vector<uint8_t> info(8 * N, 0);
    for (uint64_t r = 0; r < N && r < s->row_tags.size(); r++) std::memcpy(&info[8 * r], &s->row_tags[r], 4);
    fsa_particle_view v{N, mv.data(), 36, pd.data(), 12, info.data(), 8, ex.data(), 12};
    do_download(s, &v);
    HostCopy hc;
    fetch(s, hc, true);
    std::vector<uint32_t> order(hc.idt.size());
    s->carry_err.assign(N, 0.0f);
    for (size_t k = 0; k < hc.idt.size(); k++) {
        const uint32_t row = hc.idt[k] & kIdMask;
        order[k] = row;
        s->carry_err[row] = hc.err[k];
    }
    for (const auto& e : s->entities) {
        if (e.type != FSA_ENTITY_PARTICLE_REMOVER_3D) continue;
        float lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            hi[a] = e.center[a] + e.distance_from_center[a];
            lo[a] = e.center[a] - e.distance_from_center[a];
        }
        for (uint64_t r = 0; r < N; r++) {
            const float* m = &mv[9 * r];
            const bool inside = !(m[0] < lo[0] || m[0] > hi[0] || m[1] > hi[1] || m[1] < lo[1] || m[2] < lo[2] ||
                                  m[2] > hi[2]);
            if (inside == (e.remove_if_outside == 0u)) info[8 * r + 4] = FSA_TYPE_INACTIVE;
        }
    }
    relayout(s, &v, order);
}

void do_download_iisph(fsa_scene_s* s, void* rec, uint64_t stride) {
    if (!rec) throw FsaError(FSA_E_INVALID, "records is NULL");
    HostCopy hc;
    fetch(s, hc, true);
    const uint32_t n = s->n;
    std::vector<float4> vs(n), an(n);
    std::vector<float> src(n), dg(n);
    hipStream_t st = stream_of(s->ctx);
    if (n) {
        HIP_CHECK(hipMemcpyAsync(vs.data(), s->hd.vstar, 16ull * n, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(an.data(), s->hd.anp, 16ull * n, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(src.data(), s->hd.src, 4ull * n, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(dg.data(), s->hd.diag, 4ull * n, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
    }
    if (s->desc.dimension == 2) {
        // IISPHParticleData (IISPHFluidSolver.hpp:14-18): predicted velocity, source term, diagonal
        for (uint32_t i = 0; i < n; i++) {
            float* r = (float*)((char*)rec + (uint64_t)i * stride);
            r[0] = vs[i].x;
            r[1] = vs[i].y;
            r[2] = src[i];
            r[3] = dg[i];
        }
        return;
    }
    for (uint32_t k = 0; k < n; k++) {
        const uint64_t i = hc.idt[k] & kIdMask;
        float* r = (float*)((char*)rec + i * stride);
        r[0] = vs[k].x;
        r[1] = vs[k].y;
        r[2] = vs[k].z;
        r[3] = an[k].x;
        r[4] = an[k].y;
        r[5] = an[k].z;
        r[6] = src[k];
        r[7] = dg[k];
        r[8] = hc.err[k];
    }
}

void build_csr(fsa_scene_s* s) {
    if (s->lists_cached) return;
    if (!s->lists_valid) throw FsaError(FSA_E_STATE, "no neighbor lists: call find_neighbors first");
    read_state(s);
    const uint32_t c = s->h_state->cur;
    const uint32_t n = s->n, nf = s->nf;
    std::vector<uint32_t> idt(n), fidx(nf), cnt(nf), nbr((size_t)kMaxNeighbors * s->hd.list_stride);
    hipStream_t st = stream_of(s->ctx);
    if (n) HIP_CHECK(hipMemcpyAsync(idt.data(), s->hd.idt[c], 4ull * n, hipMemcpyDeviceToHost, st));
    if (nf) {
        HIP_CHECK(hipMemcpyAsync(fidx.data(), s->hd.fluid_idx, 4ull * nf, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(cnt.data(), s->hd.nbr_cnt, 4ull * nf, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(nbr.data(), s->hd.nbr, 4ull * nbr.size(), hipMemcpyDeviceToHost, st));
    }
    HIP_CHECK(hipStreamSynchronize(st));
    std::vector<std::vector<uint64_t>> rows(s->total);
    for (uint32_t k = 0; k < nf; k++) {
        const uint64_t row = idt[fidx[k]] & kIdMask;
        auto& r = rows[row];
        r.reserve(cnt[k]);
        for (uint32_t e = 0; e < cnt[k]; e++) {
            const uint32_t j = nbr[(size_t)e * s->hd.list_stride + k] & kIndexMask;
            r.push_back(idt[j] & kIdMask);
        }
        std::sort(r.begin(), r.end());
    }
    s->csr_off.assign(s->total + 1, 0);
    s->csr_cols.clear();
    for (uint64_t i = 0; i < s->total; i++) {
        s->csr_off[i] = s->csr_cols.size();
        s->csr_cols.insert(s->csr_cols.end(), rows[i].begin(), rows[i].end());
    }
    s->csr_off[s->total] = s->csr_cols.size();
    s->lists_cached = true;
}

void destroy_scene(fsa_scene_s* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->ctx->mu);
        auto& v = s->ctx->scenes;
        v.erase(std::remove(v.begin(), v.end(), s), v.end());
    }
    if (s->dmem) {
        (void)hipFree(s->dmem);
        s->ctx->in_use -= s->bytes;
    }
    if (s->cells) {
        (void)hipFree(s->cells);
        s->ctx->in_use -= cell_alloc_bytes(s->cells_cap);
    }
    if (s->qs_mem) {
        (void)hipFree(s->qs_mem);
        s->ctx->in_use -= s->qs_bytes;
    }
    if (s->d_table) (void)hipFree(s->d_table);
    for (auto* m : s->map_allocs) (void)hipFree(m);
    if (s->d_entities) (void)hipFree(s->d_entities);
    if (s->d_tags) (void)hipFree(s->d_tags);
    if (s->d_stats) (void)hipFree(s->d_stats);
    if (s->d_query) (void)hipFree(s->d_query);
    if (s->d_query_out) (void)hipFree(s->d_query_out);
    {
        std::lock_guard<std::mutex> g(s->ctx->mu);
        if (s->h_state) s->ctx->free_slots.push_back(s->slot);
        drop_batches_of(s->ctx, s);
    }
    delete s;
}


// ---- serving ----------------------------------------------------------------------------
struct Job {
    uint64_t id = 0;
    fsa_scene_desc desc{};
    uint32_t steps_total = 0, steps_done = 0;
    uint32_t state = FSA_JOB_QUEUED, status = FSA_OK;
    uint64_t fluid = 0;
    double submit_s = 0, admit_s = 0, finish_s = 0;
    fsa_step_stats last{};
    // host copy of the particles (upload source, then the final state)
    uint64_t count = 0;
    std::vector<float> movement, pdata, external;
    std::vector<uint8_t> info;
    fsa_scene_s* scene = nullptr;
    fsa_particle_view view() {
        fsa_particle_view v{};
        v.count = count;
        v.movement = movement.data();
        v.movement_stride = 36;
        v.particle_data = pdata.data();
        v.particle_data_stride = 12;
        v.info = info.data();
        v.info_stride = 8;
        v.external = external.data();
        v.external_stride = 12;
        return v;
    }
};

}  // namespace

struct fsa_server_s {
    fsa_context ctx = nullptr;
    fsa_server_config cfg{};
    uint64_t next_id = 1;
    std::map<uint64_t, Job*> jobs;
    std::deque<uint64_t> queue;
    std::vector<uint64_t> running;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    uint64_t ticks = 0, scene_steps = 0, particle_updates = 0, submitted = 0, done = 0, failed = 0;
    double busy_s = 0.0;
    double admit_s = 0.0, step_s = 0.0, retire_s = 0.0;  // parts of busy_s and of retirement
    std::vector<std::pair<double, uint32_t>> step_ms;  // tick latency, scenes stepped
    hipStream_t xfer = nullptr;  // uploads and downloads beside a tick's step (default priority)
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

uint64_t server_budget(fsa_server srv) {
    if (srv->cfg.hbm_budget_bytes) return srv->cfg.hbm_budget_bytes;
    return srv->ctx->budget;
}

void admit(fsa_server srv) {
    while (!srv->queue.empty()) {
        if (srv->cfg.max_running && srv->running.size() >= srv->cfg.max_running) return;
        Job* j = srv->jobs[srv->queue.front()];
        uint64_t nb = 0;
        for (uint64_t i = 0; i < j->count; i++)
            if (j->info[8 * i + 4] == FSA_TYPE_BOUNDARY) nb++;
        // scene plan + cell table allowance (the grid of a dam scene is about the particle count)
        const uint64_t need = plan_bytes(make_plan(j->fluid + nb, j->fluid, 3, false, delta_lists(j->desc))) +
                              kCellBytes * (j->fluid + nb);
        const uint64_t budget = server_budget(srv);
        if (budget && srv->ctx->in_use.load() + need > budget) {
            if (!srv->running.empty()) return;  // FIFO: wait until running jobs free memory
            // nothing will free memory: the job can never be admitted
            srv->queue.pop_front();
            j->state = FSA_JOB_FAILED;
            j->status = FSA_E_BUDGET;
            j->finish_s = srv->now();
            srv->failed++;
            continue;
        }
        srv->queue.pop_front();
        fsa_particle_view v = j->view();
        fsa_scene sc = nullptr;
        const int32_t rc = fsa_scene_create(srv->ctx, &j->desc, &v, &sc);
        if (rc == FSA_E_BUDGET && !srv->running.empty()) {
            srv->queue.push_front(j->id);  // the estimate was short: wait for running jobs to free memory
            return;
        }
        if (rc != FSA_OK) {
            j->state = FSA_JOB_FAILED;
            j->status = (uint32_t)rc;
            j->finish_s = srv->now();
            srv->failed++;
            continue;
        }
        j->scene = sc;
        j->state = FSA_JOB_RUNNING;
        j->admit_s = srv->now();
        srv->running.push_back(j->id);
    }
}

void destroy_server(fsa_server srv) {
    for (auto& kv : srv->jobs) {
        if (kv.second->scene) destroy_scene(kv.second->scene);
        delete kv.second;
    }
    if (srv->xfer) {
        (void)hipStreamSynchronize(srv->xfer);
        (void)hipStreamDestroy(srv->xfer);
    }
    auto& v = srv->ctx->servers;
    v.erase(std::remove(v.begin(), v.end(), srv), v.end());
    delete srv;
}

// Ends a running job: a DONE job's final state is copied into the job's host
// arrays first (unless `fetched`: copied already), then the scene is destroyed.
void retire(fsa_server srv, Job* j, uint32_t state, uint32_t status, bool fetched = false) {
    const double t0 = srv->now();
    if (j->scene) {
        if (state == FSA_JOB_DONE && !fetched) {
            fsa_particle_view v = j->view();
            do_download(j->scene, &v);
        }
        destroy_scene(j->scene);
        j->scene = nullptr;
    }
    j->state = state;
    j->status = status;
    j->finish_s = srv->now();
    srv->running.erase(std::remove(srv->running.begin(), srv->running.end(), j->id), srv->running.end());
    if (state == FSA_JOB_DONE) srv->done++;
    if (state == FSA_JOB_FAILED) srv->failed++;
    if (state != FSA_JOB_DONE) {
        j->movement.clear();
        j->pdata.clear();
        j->external.clear();
        j->info.clear();
    }
    srv->retire_s += srv->now() - t0;
}

constexpr uint32_t kNotStepped = 0xffffffffu;

// One step of the jobs' scenes in shared launches; st receives each scene's
// stats (status kNotStepped for a scene the batch did not reach).
int32_t step_jobs(fsa_server srv, const std::vector<Job*>& js, std::vector<fsa_step_stats>& st) {
    std::vector<fsa_scene> scenes;
    for (Job* j : js) scenes.push_back(j->scene);
    st.assign(js.size(), fsa_step_stats{});
    for (auto& e : st) e.status = kNotStepped;
    return fsa_batch_step(srv->ctx, scenes.data(), (uint32_t)scenes.size(), 1, st.data());
}

// Books a step of the jobs: a scene that failed fails its job alone (a failed
// batch reports per scene; scenes with status OK completed the step).  Jobs
// whose last step this was are retired now with `retire_done`, otherwise in the
// next tick.
void book_step(fsa_server srv, const std::vector<Job*>& js, const std::vector<fsa_step_stats>& st, int32_t rc,
               bool retire_done) {
    for (size_t q = 0; q < js.size(); q++) {
        Job* j = js[q];
        const uint32_t status = rc == FSA_OK ? FSA_OK : (st[q].status == kNotStepped ? (uint32_t)rc : st[q].status);
        if (status != FSA_OK) {
            retire(srv, j, FSA_JOB_FAILED, status);
            continue;
        }
        j->last = st[q];
        j->steps_done++;
        srv->particle_updates += j->fluid;
        if (retire_done && j->steps_done >= j->steps_total) retire(srv, j, FSA_JOB_DONE, FSA_OK);
    }
}

// While alive, the calling thread's scene work goes to `st` and its capture
// mode is relaxed, so that its allocations and copies are not refused while
// another thread captures the step's launches into a graph.
struct StreamScope {
    hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
    explicit StreamScope(hipStream_t st) {
        HIP_CHECK(hipThreadExchangeStreamCaptureMode(&mode));
        t_stream = st;
    }
    ~StreamScope() {
        t_stream = nullptr;
        (void)hipThreadExchangeStreamCaptureMode(&mode);
    }
};

// An overlapped tick: the stepping jobs take their step on a worker thread
// (context stream) while this thread copies the finished jobs' final states to
// host memory and admits queued jobs (uploads on the server's stream).  Device
// memory is freed after the step; admitted jobs take their first step in the
// next tick.
void overlapped_tick(fsa_server srv, const std::vector<Job*>& stepping, const std::vector<Job*>& finished) {
    fsa_context c = srv->ctx;
    // the pinned staging is grown here, not while the step runs
    uint64_t stage = 0;
    for (Job* j : finished) stage = std::max<uint64_t>(stage, 72ull * j->scene->n);
    for (uint64_t id : srv->queue) stage = std::max<uint64_t>(stage, 76ull * srv->jobs[id]->count);
    if (stage) {
        std::lock_guard<std::mutex> g(c->pin_mu);
        pinned_staging(c, stage);
    }
    std::vector<fsa_step_stats> st;
    int32_t rc = FSA_OK;
    std::string err;
    double step_s = 0.0;
    std::thread worker([&] {
        const double ts = srv->now();
        try {
            rc = step_jobs(srv, stepping, st);
            if (rc != FSA_OK) err = g_error;
        } catch (const std::exception& e) {  // host allocation failure before the batch ran
            rc = FSA_E_NOMEM;
            err = e.what();
        }
        step_s = srv->now() - ts;
    });
    std::vector<uint32_t> fetched(finished.size(), FSA_OK);
    std::vector<uint64_t> admitted;
    std::exception_ptr failure;
    try {
        StreamScope scope(srv->xfer);
        const double tr = srv->now();
        for (size_t q = 0; q < finished.size(); q++) {
            try {
                fsa_particle_view v = finished[q]->view();
                do_download(finished[q]->scene, &v);
            } catch (const FsaError& e) {
                fetched[q] = (uint32_t)e.code;
            }
        }
        srv->retire_s += srv->now() - tr;
        const size_t before = srv->running.size();
        const double ta = srv->now();
        admit(srv);
        srv->admit_s += srv->now() - ta;
        admitted.assign(srv->running.begin() + (std::ptrdiff_t)before, srv->running.end());
    } catch (...) {
        failure = std::current_exception();
    }
    worker.join();
    srv->step_s += step_s;
    if (st.size() != stepping.size()) {  // the batch never started
        st.assign(stepping.size(), fsa_step_stats{});
        for (auto& e : st) e.status = kNotStepped;
    }
    book_step(srv, stepping, st, rc, false);
    for (size_t q = 0; q < finished.size(); q++)
        retire(srv, finished[q], fetched[q] == FSA_OK ? FSA_JOB_DONE : FSA_JOB_FAILED, fetched[q], true);
    // the step's state reads copy a range of slots and may have covered an
    // admitted scene's slot while it was written: read those states again
    for (uint64_t id : admitted) {
        Job* j = srv->jobs[id];
        try {
            read_state(j->scene);
        } catch (const FsaError& e) {
            retire(srv, j, FSA_JOB_FAILED, (uint32_t)e.code);
        }
    }
    if (failure) std::rethrow_exception(failure);
    if (rc != FSA_OK) g_error = err;
}

}  // namespace

// =====================================================================================
extern "C" {

FSA_API int32_t fsa_version(void) { return 1; }

FSA_API const char* fsa_last_error(void) { return fsa::g_error.c_str(); }

FSA_API void fsa_scene_desc_defaults(fsa_scene_desc* d, uint32_t dimension, uint32_t solver) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->dimension = dimension;
    d->solver = solver;
    d->neighbor_lists = FSA_LIST_COMPRESSED;
    d->flags = dimension == 2 ? FSA_FLAG_EMULATE_INT_ABS : 0u;
    d->rest_density = 1.0f;
    d->gravity = 9.81f;
    d->particle_size = 1.0f;
    d->iisph_max_density_error_allowed = 0.001f;
    d->iisph_min_iterations = 2;
    d->iisph_max_iterations = 100;
    d->iisph_omega = 0.5f;
    d->iisph_gamma = 0.7f;
    d->iisph_viscosity = 5.0f;
    d->iisph_single_layer_boundary = 0;
    d->iisph_single_layer_gamma_1 = 1.1f;
    d->sesph_stiffness_k = 100000.0f;
    d->sesph_viscosity = 5.0f;
    d->sesph_single_layer_boundary = 0;
    d->sesph_single_layer_gamma_1 = 1.1f;
    d->sesph_single_layer_gamma_2 = 1.1f;
    d->timestep = 0.001f;
    d->lambda_v = 0.4f;
    d->lambda_a = 0.25f;
    d->search = dimension == 2 ? FSA_SEARCH_HASHED_2D : FSA_SEARCH_COMPRESSED_3D;
    d->timestep_generator = FSA_TIMESTEP_CONSTANT;
    // DynamicCflTimestepGenerator::DynamicCflTimestepGeneratorSettings (DynamicCflTimestepGenerator.hpp:10-17)
    d->max_timestep = 0.02f;
    d->min_timestep = 0.0001f;
    d->cfl_number = 0.8f;
}

FSA_API int32_t fsa_context_create(int32_t device, uint64_t budget, fsa_context* out) {
    return guarded([&] {
        if (!out) throw FsaError(FSA_E_INVALID, "out is NULL");
        int count = 0;
        HIP_CHECK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw FsaError(FSA_E_INVALID, "no such device");
        HIP_CHECK(hipSetDevice(device));
        auto* c = new fsa_context_s();
        c->device = device;
        c->budget = budget;
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            throw FsaError(FSA_E_DEVICE, "hipStreamCreate failed");
        }
        HIP_CHECK(hipMalloc(&c->d_states, sizeof(SceneState) * fsa_context_s::kMaxScenes));
        HIP_CHECK(hipMemset(c->d_states, 0, sizeof(SceneState) * fsa_context_s::kMaxScenes));
        HIP_CHECK(hipHostMalloc((void**)&c->h_states, sizeof(SceneState) * fsa_context_s::kMaxScenes,
                                hipHostMallocDefault));
        for (uint32_t k = fsa_context_s::kMaxScenes; k > 0; k--) c->free_slots.push_back(k - 1);
        c->ev_pool.resize(4096);
        c->ev_phase.resize(2048);
        for (auto& e : c->ev_pool) HIP_CHECK(hipEventCreate(&e));
        c->stat_scratch = scratch_address(device);  // device is current
        *out = c;
    });
}

FSA_API int32_t fsa_context_destroy(fsa_context c) {
    return guarded([&] {
        if (!c) return;
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        for (auto* srv : std::vector<fsa_server_s*>(c->servers)) destroy_server(srv);
        std::vector<fsa_scene_s*> live;
        {
            std::lock_guard<std::mutex> g(c->mu);
            live = c->scenes;
        }
        for (auto* s : live) destroy_scene(s);
        for (auto& kv : c->batches) free_batch(kv.second);
        c->batches.clear();
        if (c->d_states) (void)hipFree(c->d_states);
        if (c->h_states) (void)hipHostFree(c->h_states);
        if (c->pin) (void)hipHostFree(c->pin);
        for (auto& e : c->ev_pool) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(c->stream);
        delete c;
    });
}

FSA_API int32_t fsa_context_synchronize(fsa_context c) {
    return guarded([&] {
        if (!c) throw FsaError(FSA_E_INVALID, "context is NULL");
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

FSA_API int32_t fsa_context_memory_in_use(fsa_context c, uint64_t* bytes) {
    return guarded([&] {
        if (!c || !bytes) throw FsaError(FSA_E_INVALID, "NULL argument");
        *bytes = c->in_use.load();
    });
}

FSA_API int32_t fsa_scene_memory_estimate(const fsa_scene_desc* desc, uint64_t n_fluid, uint64_t n_boundary,
                                          uint64_t* bytes) {
    return guarded([&] {
        validate_desc(desc);
        if (!bytes) throw FsaError(FSA_E_INVALID, "bytes is NULL");
        const uint32_t search = search_of(desc);
        *bytes = plan_bytes(make_plan(n_fluid + n_boundary, n_fluid, desc->dimension,
                                      search == FSA_SEARCH_HASHED_3D || search == FSA_SEARCH_QUADRATIC_3D,
                                      delta_lists(*desc))) +
                 sizeof(SceneDev);
        // the hashed search's re-sort scratch (for the active rows; inactive rows add to it)
        if (search == FSA_SEARCH_HASHED_3D) *bytes += quick_sort_bytes(n_fluid + n_boundary);
    });
}

FSA_API int32_t fsa_scene_create(fsa_context c, const fsa_scene_desc* desc, const fsa_particle_view* view,
                                 fsa_scene* out) {
    fsa_scene_s* s = nullptr;
    int32_t rc = guarded([&] {
        if (!c || !out) throw FsaError(FSA_E_INVALID, "NULL argument");
        validate_desc(desc);
        if (!view) throw FsaError(FSA_E_INVALID, "view is NULL");
        HIP_CHECK(hipSetDevice(c->device));
        // count the particles the device holds first to size the allocation
        // (3D: active ones; 2D: all, the 2D searches see inactive particles)
        uint64_t nf = 0, nb = 0;
        for (uint64_t i = 0; i < view->count; i++) {
            const uint8_t t = ((const uint8_t*)view->info + i * view->info_stride)[4];
            if (t == FSA_TYPE_NORMAL) nf++;
            if (t == FSA_TYPE_BOUNDARY) nb++;
        }
        const uint64_t ndev = desc->dimension == 2 ? view->count : nf + nb;
        const uint32_t search = search_of(desc);
        const Plan p = make_plan(ndev, nf, desc->dimension,
                                 search == FSA_SEARCH_HASHED_3D || search == FSA_SEARCH_QUADRATIC_3D,
                                 delta_lists(*desc));
        if (p.n >= (1ull << 27)) throw FsaError(FSA_E_INVALID, "scene too large for 27-bit run starts");
        const uint64_t bytes = plan_bytes(p);
        if (c->budget && c->in_use.load() + bytes > c->budget)
            throw FsaError(FSA_E_BUDGET, "scene does not fit the HBM budget");
        s = new fsa_scene_s();
        s->ctx = c;
        s->desc = *desc;
        {
            std::lock_guard<std::mutex> g(c->mu);
            c->scenes.push_back(s);
        }
        if (hipMalloc(&s->dmem, bytes) != hipSuccess) {
            s->dmem = nullptr;
            throw FsaError(FSA_E_NOMEM, "device allocation failed");
        }
        s->bytes = bytes;
        c->in_use += bytes;
        {
            std::lock_guard<std::mutex> g(c->mu);
            if (c->free_slots.empty()) throw FsaError(FSA_E_BUDGET, "too many scenes on this context");
            s->slot = c->free_slots.back();
            c->free_slots.pop_back();
        }
        s->h_state = &c->h_states[s->slot];
        std::memset(s->h_state, 0, sizeof(SceneState));
        carve(s, p);
        s->hd.state = &c->d_states[s->slot];
        fill_params(s);
        HIP_CHECK(hipMalloc(&s->d_table, sizeof(SceneDev)));
        MapBuilder mb;
        mb.add(0, s->hd);
        build_launch_set(mb, s->d_table, s->single, s->map_allocs, stream_of(s->ctx));
        do_upload(s, view, true);
        upload_scene_table(s);
        HIP_CHECK(hipStreamSynchronize(stream_of(c)));
        *out = s;
    });
    if (rc != FSA_OK && s) destroy_scene(s);
    return rc;
}

FSA_API int32_t fsa_scene_destroy(fsa_scene s) {
    return guarded([&] {
        if (!s) return;
        (void)hipStreamSynchronize(stream_of(s->ctx));
        destroy_scene(s);
    });
}

FSA_API int32_t fsa_scene_upload(fsa_scene s, const fsa_particle_view* v) {
    return guarded([&] {
        if (!s) throw FsaError(FSA_E_INVALID, "scene is NULL");
        do_upload(s, v, false);
    });
}

FSA_API int32_t fsa_scene_download(fsa_scene s, const fsa_particle_view* v) {
    return guarded([&] {
        if (!s) throw FsaError(FSA_E_INVALID, "scene is NULL");
        do_download(s, v);
    });
}

FSA_API int32_t fsa_scene_download_iisph(fsa_scene s, void* rec, uint64_t stride) {
    return guarded([&] {
        if (!s) throw FsaError(FSA_E_INVALID, "scene is NULL");
        do_download_iisph(s, rec, stride);
    });
}

FSA_API int32_t fsa_scene_set_desc(fsa_scene s, const fsa_scene_desc* d) {
    return guarded([&] {
        if (!s) throw FsaError(FSA_E_INVALID, "scene is NULL");
        validate_desc(d);
        if (d->dimension != s->desc.dimension) throw FsaError(FSA_E_INVALID, "dimension cannot change");
        // the search fixes the scene's buffer plan (3D sorted vs unsorted orders carry a
        // grid-state array; the 2D searches keep their own grid state) and its order
        // state, as the reference fixes it with the solver's template arguments
        if (search_of(d) != search_of(&s->desc))
            throw FsaError(FSA_E_INVALID, "the neighborhood search of a scene cannot change");
        // delta-coded lists take their own buffer
        if (delta_lists(*d) != delta_lists(s->desc))
            throw FsaError(FSA_E_INVALID, "a scene cannot switch to or from delta-coded neighbor lists");
        s->desc = *d;
        const uint32_t stamp = s->hd.stamp;
        fill_params(s);
        s->hd.stamp = stamp;
        upload_scene_table(s);
    });
}

FSA_API int32_t fsa_scene_find_neighbors(fsa_scene s) {
    return guarded([&] {
        if (!s) throw FsaError(FSA_E_INVALID, "scene is NULL");
        do_find_neighbors(s);
    });
}

FSA_API int32_t fsa_scene_get_neighbors(fsa_scene s, uint64_t index, uint64_t* out, uint64_t cap, uint64_t* count) {
    return guarded([&] {
        if (!s || !count) throw FsaError(FSA_E_INVALID, "NULL argument");
        if (index >= s->total) throw FsaError(FSA_E_INVALID, "index out of range");
        build_csr(s);
        const uint64_t b = s->csr_off[index], e = s->csr_off[index + 1];
        *count = e - b;
        if (out)
            for (uint64_t k = 0; k < std::min(cap, e - b); k++) out[k] = s->csr_cols[b + k];
        if (e - b > cap && out) throw FsaError(FSA_E_INVALID, "output capacity too small");
    });
}

FSA_API int32_t fsa_scene_export_neighbors(fsa_scene s, uint64_t* off, uint64_t* cols, uint64_t cap, uint64_t* nnz) {
    return guarded([&] {
        if (!s || !nnz) throw FsaError(FSA_E_INVALID, "NULL argument");
        build_csr(s);
        *nnz = s->csr_cols.size();
        if (off) std::memcpy(off, s->csr_off.data(), s->csr_off.size() * sizeof(uint64_t));
        if (cols) {
            if (cap < s->csr_cols.size()) throw FsaError(FSA_E_INVALID, "output capacity too small");
            std::memcpy(cols, s->csr_cols.data(), s->csr_cols.size() * sizeof(uint64_t));
        }
    });
}

FSA_API int32_t fsa_scene_neighbor_stats(fsa_scene s, uint64_t* total, uint64_t* dev_bytes, uint64_t* ref_bytes) {
    return guarded([&] {
        if (!s) throw FsaError(FSA_E_INVALID, "scene is NULL");
        if (!s->lists_valid) throw FsaError(FSA_E_STATE, "no neighbor lists: call find_neighbors first");
        // the NeighborStorage sizes of the current lists, counted on the device
        upload_scene_table(s);
        LaunchSet L = s->single;
        launch_codec(L, stream_of(s->ctx), true);
        read_state(s);
        if (total) *total = s->h_state->total_neighbors;
        if (dev_bytes)
            *dev_bytes = (s->hd.flags & kFlagDeltaLists) && !s->h_state->delta_plain ? s->h_state->delta_bytes
                                                                                     : 4 * s->h_state->total_neighbors;
        if (ref_bytes) *ref_bytes = s->h_state->codec_bytes;
    });
}

FSA_API int32_t fsa_scene_simulate(fsa_scene s, float dt, fsa_step_stats* stats) {
    return guarded([&] {
        if (!s) throw FsaError(FSA_E_INVALID, "scene is NULL");
        do_simulate(s, dt, stats);
    });
}

FSA_API int32_t fsa_scene_set_entities(fsa_scene s, const fsa_entity* e, uint32_t count) {
    return guarded([&] {
        if (!s || (count && !e)) throw FsaError(FSA_E_INVALID, "NULL argument");
        for (uint32_t k = 0; k < count; k++) {
            if (e[k].type != FSA_ENTITY_BOUNDARY_PREPROCESSOR && e[k].type != FSA_ENTITY_VELOCITY_ALTERATION_BY_TAG &&
                e[k].type != FSA_ENTITY_PARTICLE_REMOVER_3D)
                throw FsaError(FSA_E_INVALID, "unknown entity type");
        }
        if (count && s->desc.dimension != 3)
            throw FsaError(FSA_E_INVALID, "the entities act on 3D collections (MovementData3D)");
        HIP_CHECK(hipSetDevice(s->ctx->device));
        HIP_CHECK(hipStreamSynchronize(stream_of(s->ctx)));
        s->entities.assign(e, e + count);
        if (s->d_entities) (void)hipFree(s->d_entities);
        s->d_entities = nullptr;
        if (count) {
            if (hipMalloc(&s->d_entities, sizeof(EntityDev) * count) != hipSuccess) {
                s->d_entities = nullptr;
                throw FsaError(FSA_E_NOMEM, "device allocation failed");
            }
            HIP_CHECK(hipMemcpyAsync(s->d_entities, e, sizeof(EntityDev) * count, hipMemcpyHostToDevice,
                                     stream_of(s->ctx)));
        }
        refresh_entities(s);
        s->boundary_dirty = true;
        upload_scene_table(s);
        HIP_CHECK(hipStreamSynchronize(stream_of(s->ctx)));
    });
}

FSA_API int32_t fsa_scene_get_statistics(fsa_scene s, float zero_height, fsa_scene_statistics* out) {
    return guarded([&] {
        if (!s || !out) throw FsaError(FSA_E_INVALID, "NULL argument");
        HIP_CHECK(hipSetDevice(s->ctx->device));
        hipStream_t st = stream_of(s->ctx);
        const uint64_t need = statistics_scratch_bytes(std::max<uint32_t>(s->hd.nblocks, 1));
        if (s->d_stats_bytes < need) {
            if (s->d_stats) (void)hipFree(s->d_stats);
            s->d_stats = nullptr;
            if (hipMalloc(&s->d_stats, need) != hipSuccess) {
                s->d_stats = nullptr;
                throw FsaError(FSA_E_NOMEM, "device allocation failed");
            }
            s->d_stats_bytes = need;
        }
        upload_scene_table(s);
        launch_statistics(s->d_table, s->hd.nblocks, zero_height, s->d_stats, st);
        const uint64_t rb = statistics_result_bytes();
        alignas(16) uint8_t res[kStatAccBytes];
        HIP_CHECK(hipMemcpyAsync(res, (char*)s->d_stats + need - rb, rb, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        fill_statistics(s, res, out);
    });
}

FSA_API int32_t fsa_scene_get_statistics_noalloc(fsa_scene s, float zero_height, fsa_scene_statistics* out) {
    return guarded([&] {
        if (!s || !out) throw FsaError(FSA_E_INVALID, "NULL argument");
        const uint32_t nb = std::max<uint32_t>(s->hd.nblocks, 1);
        if (nb > kNoAllocStatBlocks)
            throw FsaError(FSA_E_INVALID, "scene exceeds the no-allocation statistics limit of 16,777,216 "
                                          "resident particles: use fsa_scene_get_statistics");
        if (!s->ctx->stat_scratch)
            throw FsaError(FSA_E_DEVICE, "the statistics scratch of this device has no address");
        HIP_CHECK(hipSetDevice(s->ctx->device));
        hipStream_t st = stream_of(s->ctx);
        upload_scene_table(s);
        // every block writes its partial and the finalize block writes the result: nothing
        // in the scratch needs clearing between calls
        launch_statistics(s->d_table, s->hd.nblocks, zero_height, s->ctx->stat_scratch, st);
        alignas(16) uint8_t res[kStatAccBytes];
        HIP_CHECK(hipMemcpyAsync(res, (char*)s->ctx->stat_scratch + (uint64_t)nb * kStatAccBytes, kStatAccBytes,
                                 hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        fill_statistics(s, res, out);
    });
}

// ---- position queries -------------------------------------------------------------------
// scratch that grows to the largest request and is kept with the scene
void grow(void*& p, uint64_t& have, uint64_t need) {
    if (have >= need) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    have = 0;
    if (hipMalloc(&p, need) != hipSuccess) {
        p = nullptr;
        throw FsaError(FSA_E_NOMEM, "device allocation failed");
    }
    have = need;
}

// position queries read the cell structure of the last search over the present layout
void check_queryable(fsa_scene_s* s) {
    if (s->hd.dim != 3) throw FsaError(FSA_E_INVALID, "position queries are defined for 3D scenes");
    if (!s->lists_valid) throw FsaError(FSA_E_STATE, "no search over the present layout: call find_neighbors first");
}

// positions (count x 3 floats) to float4 scratch at the start of d_query
float4* upload_queries(fsa_scene_s* s, const float* positions, uint64_t count, uint64_t extra) {
    if (count >= (1ull << 31)) throw FsaError(FSA_E_INVALID, "too many query positions");
    if (count && !positions) throw FsaError(FSA_E_INVALID, "positions is NULL");
    std::vector<float4> q(count);
    for (uint64_t k = 0; k < count; k++) {
        const float x = positions[3 * k], y = positions[3 * k + 1], z = positions[3 * k + 2];
        if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z))
            throw FsaError(FSA_E_INVALID, "query position is not finite");
        q[k] = make_float4(x, y, z, 0.0f);
    }
    grow(s->d_query, s->d_query_bytes, std::max<uint64_t>(16 * count + extra, 256));
    if (count)
        HIP_CHECK(hipMemcpyAsync(s->d_query, q.data(), 16 * count, hipMemcpyHostToDevice, stream_of(s->ctx)));
    return (float4*)s->d_query;
}

FSA_API int32_t fsa_scene_query_neighbors(fsa_scene s, const float* positions, uint64_t count, uint64_t* offsets,
                                          uint64_t* rows, uint64_t capacity, uint64_t* nnz) {
    return guarded([&] {
        if (!s || !nnz) throw FsaError(FSA_E_INVALID, "NULL argument");
        HIP_CHECK(hipSetDevice(s->ctx->device));
        check_queryable(s);
        hipStream_t st = stream_of(s->ctx);
        const uint64_t cbytes = (4 * count + 15) & ~15ull;
        float4* q = upload_queries(s, positions, count, cbytes + 8 * count);
        uint32_t* d_counts = (uint32_t*)((char*)q + 16 * count);
        uint64_t* d_off = (uint64_t*)((char*)d_counts + cbytes);
        upload_scene_table(s);
        launch_query_count(s->d_table, q, (uint32_t)count, d_counts, st);
        std::vector<uint32_t> cnt(count);
        if (count) HIP_CHECK(hipMemcpyAsync(cnt.data(), d_counts, 4 * count, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        std::vector<uint64_t> off(count + 1, 0);
        for (uint64_t k = 0; k < count; k++) off[k + 1] = off[k] + cnt[k];
        *nnz = off[count];
        if (offsets) std::memcpy(offsets, off.data(), 8 * (count + 1));
        if (!rows) return;
        if (capacity < off[count]) throw FsaError(FSA_E_INVALID, "output capacity too small");
        if (!off[count]) return;
        grow(s->d_query_out, s->d_query_out_bytes, 8 * off[count]);
        HIP_CHECK(hipMemcpyAsync(d_off, off.data(), 8 * count, hipMemcpyHostToDevice, st));
        launch_query_fill(s->d_table, q, (uint32_t)count, d_off, (uint64_t*)s->d_query_out, st);
        HIP_CHECK(hipMemcpyAsync(rows, s->d_query_out, 8 * off[count], hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
    });
}

FSA_API void fsa_sensor_plane_defaults(fsa_sensor_plane* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->span_x[0] = 1.0f;
    p->span_y[1] = 1.0f;
    p->width = 30.0f;
    p->height = 10.0f;
    p->number_of_samples_x = 300;
    p->number_of_samples_y = 100;
    p->max_image_value = 1.02f;
    p->min_image_value = 0.98f;
    p->sensor_type = FSA_SENSOR_PLANE_DENSITY;
    p->sub_sample_grid_size = 2;
    p->calculate_plane_every_nth_step = 1;
}

FSA_API int32_t fsa_scene_sensor_plane(fsa_scene s, const fsa_sensor_plane* p, float* values) {
    return guarded([&] {
        if (!s || !p || !values) throw FsaError(FSA_E_INVALID, "NULL argument");
        HIP_CHECK(hipSetDevice(s->ctx->device));
        check_queryable(s);
        if (p->sensor_type > FSA_SENSOR_PLANE_PRESSURE) throw FsaError(FSA_E_INVALID, "unknown sensor plane type");
        const uint64_t samples = p->number_of_samples_x * p->number_of_samples_y;
        if (p->number_of_samples_x == 0 || p->number_of_samples_y == 0 || p->number_of_samples_x >= (1ull << 24) ||
            p->number_of_samples_y >= (1ull << 24) || samples >= (1ull << 31))
            throw FsaError(FSA_E_INVALID, "sample counts must be positive and below 2^24 each, 2^31 in all");
        if (p->sub_sample_grid_size >= 4096) throw FsaError(FSA_E_INVALID, "sub_sample_grid_size must be below 4096");
        for (int k = 0; k < 3; k++)
            if (!std::isfinite(p->origin[k]) || !std::isfinite(p->span_x[k]) || !std::isfinite(p->span_y[k]))
                throw FsaError(FSA_E_INVALID, "sensor plane geometry is not finite");
        if (!std::isfinite(p->width) || !std::isfinite(p->height))
            throw FsaError(FSA_E_INVALID, "sensor plane geometry is not finite");
        // glm::normalize(v) = v * (1 / sqrt(dot(v, v))); x_step = width / number_of_samples_x
        PlaneParams P{};
        const float x_step = p->width / (float)p->number_of_samples_x;
        const float y_step = p->height / (float)p->number_of_samples_y;
        const float* sp[2] = {p->span_x, p->span_y};
        float* dst[2] = {P.ax, P.ay};
        const float step[2] = {x_step, y_step};
        for (int a = 0; a < 2; a++) {
            const float* v = sp[a];
            const float d = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
            const float inv = 1.0f / std::sqrt(d);
            for (int k = 0; k < 3; k++) dst[a][k] = (v[k] * inv) * step[a];
        }
        for (int k = 0; k < 3; k++) P.origin[k] = p->origin[k];
        P.nx = (uint32_t)p->number_of_samples_x;
        P.ny = (uint32_t)p->number_of_samples_y;
        P.sub = (uint32_t)p->sub_sample_grid_size;
        P.type = p->sensor_type;
        P.sub_den = (float)p->sub_sample_grid_size + 1.0f;
        const float g = (float)(p->sub_sample_grid_size + 1);
        P.div = g * g;
        hipStream_t st = stream_of(s->ctx);
        grow(s->d_query_out, s->d_query_out_bytes, 12 * samples);
        upload_scene_table(s);
        launch_sensor_plane(s->d_table, P, (float*)s->d_query_out, st);
        HIP_CHECK(hipMemcpyAsync(values, s->d_query_out, 12 * samples, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
    });
}

FSA_API int32_t fsa_scene_evaluate_volume(fsa_scene s, const float* positions, uint64_t count,
                                          fsa_volume_sample* out) {
    return guarded([&] {
        if (!s || (count && !out)) throw FsaError(FSA_E_INVALID, "NULL argument");
        HIP_CHECK(hipSetDevice(s->ctx->device));
        check_queryable(s);
        hipStream_t st = stream_of(s->ctx);
        float4* q = upload_queries(s, positions, count, 0);
        if (!count) return;
        grow(s->d_query_out, s->d_query_out_bytes, 36 * count);
        upload_scene_table(s);
        launch_volume(s->d_table, q, (uint32_t)count, (float*)s->d_query_out, st);
        HIP_CHECK(hipMemcpyAsync(out, s->d_query_out, 36 * count, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
    });
}

FSA_API int32_t fsa_scene_next_timestep(fsa_scene s, float* dt) {
    return guarded([&] {
        if (!s || !dt) throw FsaError(FSA_E_INVALID, "NULL argument");
        *dt = generated_timestep(s);
    });
}

FSA_API int32_t fsa_scene_step(fsa_scene s, uint32_t steps, fsa_step_stats* stats) {
    return guarded([&] {
        if (!s) throw FsaError(FSA_E_INVALID, "scene is NULL");
        HIP_CHECK(hipSetDevice(s->ctx->device));
        // a step of one scene is a batch of one: search and solve without a host check in between
        std::vector<fsa_scene_s*> one{s};
        std::vector<fsa_step_stats*> out{stats};
        for (uint32_t k = 0; k < steps; k++) {
            const int32_t rc = batch_step_group(s->ctx, one, out);
            if (rc != FSA_OK) throw FsaError(rc, g_error);
        }
    });
}

FSA_API int32_t fsa_batch_step(fsa_context c, const fsa_scene* scenes, uint32_t n, uint32_t steps,
                               fsa_step_stats* stats) {
    return guarded([&] {
        if (!c || (!scenes && n)) throw FsaError(FSA_E_INVALID, "NULL argument");
        std::vector<fsa_scene_s*> v(scenes, scenes + n);
        for (auto* s : v)
            if (!s || s->ctx != c) throw FsaError(FSA_E_INVALID, "scene of another context");
        // scenes of one solver, dimension and search family share launches
        std::vector<fsa_scene_s*> groups[8];
        std::vector<fsa_step_stats*> gstats[8];
        for (uint32_t q = 0; q < n; q++) {
            const int g = (v[q]->desc.solver == FSA_SOLVER_SESPH ? 1 : 0) + (v[q]->desc.dimension == 2 ? 2 : 0) +
                          ((v[q]->hd.flags & kFlagUnsorted) ? 4 : 0);
            if (std::find(groups[g].begin(), groups[g].end(), v[q]) != groups[g].end())
                throw FsaError(FSA_E_INVALID, "a scene appears twice in the batch");
            groups[g].push_back(v[q]);
            gstats[g].push_back(stats ? &stats[q] : nullptr);
        }
        HIP_CHECK(hipSetDevice(c->device));
        int32_t first_error = FSA_OK;
        for (uint32_t k = 0; k < steps && first_error == FSA_OK; k++)
            for (int g = 0; g < 8; g++)
                if (!groups[g].empty()) {
                    const int32_t rc = batch_step_group(c, groups[g], gstats[g]);
                    if (first_error == FSA_OK) first_error = rc;
                }
        if (first_error != FSA_OK) throw FsaError(first_error, g_error);
    });
}

FSA_API int32_t fsa_batch_set_graph(fsa_context c, uint32_t enabled) {
    return guarded([&] {
        if (!c) throw FsaError(FSA_E_INVALID, "context is NULL");
        c->graph = enabled != 0;
    });
}

FSA_API int32_t fsa_context_set_autotune(fsa_context c, uint32_t enabled) {
    return guarded([&] {
        if (!c) throw FsaError(FSA_E_INVALID, "context is NULL");
        c->autotune = enabled != 0;
    });
}

FSA_API int32_t fsa_context_set_walk_variant(fsa_context c, uint32_t variant) {
    return guarded([&] {
        if (!c) throw FsaError(FSA_E_INVALID, "context is NULL");
        if (variant && std::find(kTuneVariant, kTuneVariant + kTuneVariants, variant) == kTuneVariant + kTuneVariants)
            throw FsaError(FSA_E_INVALID, "walk variant is none of the autotuner's candidates");
        c->walk_variant = (uint8_t)variant;
        // cached batches follow: a forced variant at their next step, else a fresh choice
        for (auto& kv : c->batches) {
            if (variant) continue;
            for (int f = 0; f < kTuneFamilies; f++) kv.second->ls.unroll[f] = kVariantU4;
            kv.second->tuned = false;
        }
    });
}

FSA_API int32_t fsa_context_get_tuning(fsa_context c, char* buf, uint64_t cap, uint64_t* len) {
    return guarded([&] {
        if (!c || !len) throw FsaError(FSA_E_INVALID, "NULL argument");
        static const char* names[kTuneFamilies] = {"prepare", "diagonal", "pass_a", "pass_b"};
        std::string out;
        char line[256];
        for (auto& kv : c->tuned) {
            snprintf(line, sizeof(line), "class spec=%u fluid<2^%u (first seen with %llu fluid):",
                     (unsigned)(kv.first >> 32), (unsigned)(kv.first & 0xffffffffu),
                     (unsigned long long)kv.second.fluid);
            out += line;
            for (int f = 0; f < kTuneFamilies; f++) {
                // a variant reads U<unroll> or U<unroll>W<waves per SIMD hint>
                auto name = [](uint8_t v) {
                    std::string r = "U" + std::to_string(v & 15u);
                    if (v >> 4) r += "W" + std::to_string(v >> 4);
                    return r;
                };
                out += " " + std::string(names[f]) + "=" + name(kv.second.unroll[f]) + " [";
                for (int c = 0; c < kTuneVariants; c++) {
                    snprintf(line, sizeof(line), "%s%s %.1f us", c ? ", " : "", name(kTuneVariant[c]).c_str(),
                             kv.second.us[f][c]);
                    out += line;
                }
                out += "]";
            }
            out += "\n";
        }
        *len = out.size();
        if (buf && cap) {
            const uint64_t m = std::min<uint64_t>(cap - 1, out.size());
            std::memcpy(buf, out.data(), m);
            buf[m] = 0;
        }
    });
}

FSA_API int32_t fsa_context_set_block_size(fsa_context c, uint32_t threads) {
    return guarded([&] {
        if (!c) throw FsaError(FSA_E_INVALID, "context is NULL");
        if (threads != 0 && threads != 256) throw FsaError(FSA_E_INVALID, "unsupported block size");
        c->block_override = threads;
    });
}

FSA_API int32_t fsa_context_set_timing(fsa_context c, uint32_t enabled) {
    return guarded([&] {
        if (!c) throw FsaError(FSA_E_INVALID, "context is NULL");
        collect_timing(c);
        c->timing = enabled != 0;
        for (int k = 0; k < FSA_TIMER_COUNT; k++) {
            c->timer_ms[k] = 0.0;
            c->timer_n[k] = 0;
        }
    });
}

FSA_API int32_t fsa_context_get_timing(fsa_context c, double* total_ms, uint64_t* launches, uint32_t n) {
    return guarded([&] {
        if (!c) throw FsaError(FSA_E_INVALID, "context is NULL");
        collect_timing(c);
        for (uint32_t k = 0; k < n && k < (uint32_t)FSA_TIMER_COUNT; k++) {
            if (total_ms) total_ms[k] = c->timer_ms[k];
            if (launches) launches[k] = c->timer_n[k];
        }
    });
}

FSA_API int32_t fsa_server_create(fsa_context c, const fsa_server_config* cfg, fsa_server* out) {
    return guarded([&] {
        if (!c || !out) throw FsaError(FSA_E_INVALID, "NULL argument");
        HIP_CHECK(hipSetDevice(c->device));
        hipStream_t xfer = nullptr;
        if (hipStreamCreateWithFlags(&xfer, hipStreamNonBlocking) != hipSuccess)
            throw FsaError(FSA_E_DEVICE, "hipStreamCreate failed");
        auto* s = new fsa_server_s();
        s->ctx = c;
        s->xfer = xfer;
        if (cfg) s->cfg = *cfg;
        c->servers.push_back(s);
        *out = s;
    });
}

FSA_API int32_t fsa_server_destroy(fsa_server srv) {
    return guarded([&] {
        if (srv) destroy_server(srv);
    });
}

FSA_API int32_t fsa_server_submit(fsa_server srv, const fsa_scene_desc* desc, const fsa_particle_view* v,
                                  uint32_t steps, uint64_t* job_id) {
    return guarded([&] {
        if (!srv || !desc || !v || !job_id) throw FsaError(FSA_E_INVALID, "NULL argument");
        validate_desc(desc);
        if (!v->movement || !v->particle_data || !v->info) throw FsaError(FSA_E_INVALID, "incomplete particle view");
        auto* j = new Job();
        j->id = srv->next_id++;
        j->desc = *desc;
        j->steps_total = steps;
        j->count = v->count;
        j->movement.resize(9 * v->count);
        j->pdata.resize(3 * v->count);
        j->external.assign(3 * v->count, 0.0f);
        j->info.resize(8 * v->count);
        for (uint64_t i = 0; i < v->count; i++) {
            std::memcpy(&j->movement[9 * i], (const char*)v->movement + i * v->movement_stride, 36);
            std::memcpy(&j->pdata[3 * i], (const char*)v->particle_data + i * v->particle_data_stride, 12);
            std::memcpy(&j->info[8 * i], (const char*)v->info + i * v->info_stride, 8);
            if (v->external)
                std::memcpy(&j->external[3 * i], (const char*)v->external + i * v->external_stride, 12);
            if (j->info[8 * i + 4] == FSA_TYPE_NORMAL) j->fluid++;
        }
        j->submit_s = srv->now();
        srv->jobs[j->id] = j;
        srv->queue.push_back(j->id);
        srv->submitted++;
        *job_id = j->id;
    });
}

FSA_API int32_t fsa_server_cancel(fsa_server srv, uint64_t id) {
    return guarded([&] {
        if (!srv) throw FsaError(FSA_E_INVALID, "server is NULL");
        auto it = srv->jobs.find(id);
        if (it == srv->jobs.end()) throw FsaError(FSA_E_INVALID, "unknown job");
        Job* j = it->second;
        if (j->state == FSA_JOB_QUEUED) {
            srv->queue.erase(std::remove(srv->queue.begin(), srv->queue.end(), id), srv->queue.end());
            j->state = FSA_JOB_CANCELLED;
            j->finish_s = srv->now();
        } else if (j->state == FSA_JOB_RUNNING) {
            // a job whose steps have all run (retired by the next tick) is done, not cancelled
            if (j->steps_done >= j->steps_total)
                retire(srv, j, FSA_JOB_DONE, FSA_OK);
            else
                retire(srv, j, FSA_JOB_CANCELLED, FSA_OK);
        }
    });
}

FSA_API int32_t fsa_server_tick(fsa_server srv, uint32_t* running) {
    return guarded([&] {
        if (!srv) throw FsaError(FSA_E_INVALID, "server is NULL");
        if (running) *running = 0;
        const double t0 = srv->now();
        // running jobs with steps left, and jobs whose steps ended in an earlier tick
        std::vector<Job*> stepping, finished;
        for (uint64_t id : srv->running) {
            Job* j = srv->jobs[id];
            (j->steps_done >= j->steps_total ? finished : stepping).push_back(j);
        }
        if (!srv->cfg.serial_ticks && !stepping.empty() && !srv->ctx->timing) {
            if (running) *running = (uint32_t)stepping.size();
            overlapped_tick(srv, stepping, finished);
            const double t1 = srv->now();
            srv->ticks++;
            srv->busy_s += t1 - t0;
            srv->step_ms.push_back({1000.0 * (t1 - t0), (uint32_t)stepping.size()});
            srv->scene_steps += stepping.size();
            return;
        }
        // one after another: retire, admit, step every running job
        for (Job* j : finished) retire(srv, j, FSA_JOB_DONE, FSA_OK);
        const double ta = srv->now();
        admit(srv);
        srv->admit_s += srv->now() - ta;
        stepping.clear();
        for (uint64_t id : std::vector<uint64_t>(srv->running)) {
            Job* j = srv->jobs[id];
            if (j->steps_done >= j->steps_total)
                retire(srv, j, FSA_JOB_DONE, FSA_OK);  // admitted with no steps
            else
                stepping.push_back(j);
        }
        if (running) *running = (uint32_t)stepping.size();
        if (stepping.empty()) return;
        std::vector<fsa_step_stats> st;
        const double ts = srv->now();
        const int32_t rc = step_jobs(srv, stepping, st);
        const double t1 = srv->now();
        srv->step_s += t1 - ts;
        srv->ticks++;
        srv->busy_s += t1 - t0;
        srv->step_ms.push_back({1000.0 * (t1 - t0), (uint32_t)stepping.size()});
        srv->scene_steps += stepping.size();
        book_step(srv, stepping, st, rc, srv->cfg.serial_ticks != 0);
    });
}

FSA_API int32_t fsa_server_job(fsa_server srv, uint64_t id, fsa_job_info* info) {
    return guarded([&] {
        if (!srv || !info) throw FsaError(FSA_E_INVALID, "NULL argument");
        auto it = srv->jobs.find(id);
        if (it == srv->jobs.end()) throw FsaError(FSA_E_INVALID, "unknown job");
        const Job* j = it->second;
        info->job_id = j->id;
        info->state = j->state;
        info->status = j->status;
        info->steps_done = j->steps_done;
        info->steps_total = j->steps_total;
        info->fluid_particles = j->fluid;
        info->submit_s = j->submit_s;
        info->admit_s = j->admit_s;
        info->finish_s = j->finish_s;
        info->last = j->last;
    });
}

FSA_API int32_t fsa_server_collect(fsa_server srv, uint64_t id, const fsa_particle_view* v) {
    return guarded([&] {
        if (!srv) throw FsaError(FSA_E_INVALID, "server is NULL");
        auto it = srv->jobs.find(id);
        if (it == srv->jobs.end()) throw FsaError(FSA_E_INVALID, "unknown job");
        Job* j = it->second;
        if (j->state == FSA_JOB_QUEUED || j->state == FSA_JOB_RUNNING)
            throw FsaError(FSA_E_STATE, "job has not finished");
        // a failed or cancelled job has no final state: it is forgotten without a copy
        if (v && j->state == FSA_JOB_DONE) {
            if (v->count != j->count) throw FsaError(FSA_E_INVALID, "view count mismatch");
            for (uint64_t i = 0; i < j->count; i++) {
                if (v->movement) std::memcpy((char*)v->movement + i * v->movement_stride, &j->movement[9 * i], 36);
                if (v->particle_data)
                    std::memcpy((char*)v->particle_data + i * v->particle_data_stride, &j->pdata[3 * i], 12);
                if (v->info) std::memcpy((char*)v->info + i * v->info_stride, &j->info[8 * i], 8);
                if (v->external)
                    std::memcpy((char*)v->external + i * v->external_stride, &j->external[3 * i], 12);
            }
        }
        delete j;
        srv->jobs.erase(it);
    });
}

FSA_API int32_t fsa_server_get_stats(fsa_server srv, fsa_server_stats* out) {
    return guarded([&] {
        if (!srv || !out) throw FsaError(FSA_E_INVALID, "NULL argument");
        std::memset(out, 0, sizeof(*out));
        out->ticks = srv->ticks;
        out->scene_steps = srv->scene_steps;
        out->particle_updates = srv->particle_updates;
        out->jobs_submitted = srv->submitted;
        out->jobs_done = srv->done;
        out->jobs_failed = srv->failed;
        out->jobs_queued = srv->queue.size();
        out->jobs_running = srv->running.size();
        out->hbm_in_use = srv->ctx->in_use.load();
        out->busy_s = srv->busy_s;
        out->admit_s = srv->admit_s;
        out->step_s = srv->step_s;
        out->retire_s = srv->retire_s;
        // percentiles over scene-steps: a tick's latency counts once per scene it stepped
        std::vector<std::pair<double, uint32_t>> v = srv->step_ms;
        std::sort(v.begin(), v.end());
        uint64_t total = 0;
        double sum = 0.0;
        for (auto& e : v) {
            total += e.second;
            sum += e.first * e.second;
        }
        if (total) {
            auto pct = [&](double q) {
                const uint64_t target = (uint64_t)std::ceil(q * (double)total);
                uint64_t acc = 0;
                for (auto& e : v) {
                    acc += e.second;
                    if (acc >= std::max<uint64_t>(target, 1)) return e.first;
                }
                return v.back().first;
            };
            out->p50_step_ms = pct(0.50);
            out->p99_step_ms = pct(0.99);
            out->mean_step_ms = sum / (double)total;
            out->max_step_ms = v.back().first;
        }
    });
}

}  // extern "C"
