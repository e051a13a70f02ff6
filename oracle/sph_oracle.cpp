// sph_oracle.cpp — TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of libFluid's per-timestep hot path (neighborhood search +
// IISPH/SESPH particle updates), written from the reference's source text.
// It is the parity oracle for the HIP library and the timed CPU baseline of
// bench.py.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it, and only as the checker; the product path never does.
//
// Two arithmetic modes share one template:
//   * T = float : reference evaluation order (glm: dot = (x*x'+y*y')+z*z',
//                 length = sqrt(dot), no FMA contraction: built with
//                 -ffp-contract=off), FP32 serial convergence sums in index
//                 order as IISPHFluidSolver3D.hpp:467-509 does.
//   * T = double: the same algorithm in FP64, the accuracy reference.
//
// Build decisions where the reference is nondeterministic (documented in the
// design notes, applied identically on the GPU):
//   * viscosity reads neighbor densities from a start-of-step snapshot
//     (the reference reads them while the same loop writes them:
//     IISPHFluidSolver3D.hpp:61 vs 210; IISPHFluidSolver.hpp:160-167, 261);
//   * the 2D IISPH density-error sum is a serial sum (the reference sums it
//     from several threads unsynchronised: IISPHFluidSolver.hpp:508-509).
//
// Pinning of every function (tests/test_oracle_*.py): the reference's own
// known answers (NeighborhoodSearchTests.cpp:98-366, NeighborhoodStorageTests.cpp,
// PCQuickSortStableTest.cpp), brute-force neighbor sets, kernel closed forms
// and quadrature normalisation, finite-difference gradients, momentum
// conservation of the pair forces, free fall and FP32-vs-FP64 agreement.
#include "../include/fluidsolver_amd.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <bitset>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <list>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace orc {

static thread_local std::string g_err;
static int g_threads = 0;
// Planted mistakes for the tests that show a solver check detects them (0 = none):
// 1 boundary factor 2 dropped from d_ii, 2 gradW_ij for gradW_ji in a_ii's second
// term, 3 boundary factor 2 dropped from pass A, 4 the 3D boundary source term with
// v*_i - v*_b (2D: with v*_i alone), 5 the Jacobi step's sign flipped, 6 pass A's
// fluid term with p_i for p_j, 7 SESPH's fluid pressure term with p_i/rho_i^2 for
// p_j/rho_j^2, 8 SESPH's pressure without the clamp at 0.  Set only by tests
// (orc_set_mutation).
static int g_mutation = 0;

static int thread_count() {
    if (g_threads > 0) return g_threads;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

// StdParallelForEach::loop_for (parallelization/StdParallelForEach.cpp:64-84):
// every index is processed independently, so the split does not change results.
template <class F>
static void loop_for(size_t n, F&& fn) {
    int nt = thread_count();
    if (nt <= 1 || n < 4096) {
        for (size_t i = 0; i < n; i++) fn(i);
        return;
    }
    // chunks of 512 indices handed out dynamically (fluid and boundary interleave
    // unevenly in the sorted order, so static halves would idle most threads)
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++) {
        pool.emplace_back([&fn, &next, n]() {
            for (;;) {
                const size_t b = next.fetch_add(512);
                if (b >= n) break;
                const size_t e = std::min(n, b + 512);
                for (size_t i = b; i < e; i++) fn(i);
            }
        });
    }
    for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------
// glm-order small vectors
template <class T, int D>
struct Vec {
    T v[D];
};
template <class T, int D>
static inline Vec<T, D> vzero() {
    Vec<T, D> r;
    for (int k = 0; k < D; k++) r.v[k] = T(0);
    return r;
}
template <class T, int D>
static inline Vec<T, D> operator+(const Vec<T, D>& a, const Vec<T, D>& b) {
    Vec<T, D> r;
    for (int k = 0; k < D; k++) r.v[k] = a.v[k] + b.v[k];
    return r;
}
template <class T, int D>
static inline Vec<T, D> operator-(const Vec<T, D>& a, const Vec<T, D>& b) {
    Vec<T, D> r;
    for (int k = 0; k < D; k++) r.v[k] = a.v[k] - b.v[k];
    return r;
}
template <class T, int D>
static inline Vec<T, D> operator*(T s, const Vec<T, D>& a) {
    Vec<T, D> r;
    for (int k = 0; k < D; k++) r.v[k] = s * a.v[k];
    return r;
}
template <class T, int D>
static inline Vec<T, D> operator*(const Vec<T, D>& a, T s) {
    Vec<T, D> r;
    for (int k = 0; k < D; k++) r.v[k] = a.v[k] * s;
    return r;
}
template <class T, int D>
static inline Vec<T, D> operator/(const Vec<T, D>& a, T s) {
    Vec<T, D> r;
    for (int k = 0; k < D; k++) r.v[k] = a.v[k] / s;
    return r;
}
template <class T, int D>
static inline Vec<T, D> operator-(const Vec<T, D>& a) {
    Vec<T, D> r;
    for (int k = 0; k < D; k++) r.v[k] = -a.v[k];
    return r;
}
// glm::dot: ((x*x') + (y*y')) + (z*z')
template <class T, int D>
static inline T dot(const Vec<T, D>& a, const Vec<T, D>& b) {
    T s = a.v[0] * b.v[0];
    for (int k = 1; k < D; k++) s = s + a.v[k] * b.v[k];
    return s;
}
template <class T, int D>
static inline T length(const Vec<T, D>& a) {
    return std::sqrt(dot(a, a));
}

template <class T>
struct Consts;
template <>
struct Consts<float> {
    static constexpr float PI = 3.14159265358979323846f;
    static float eps() { return std::numeric_limits<float>::epsilon(); }
};
template <>
struct Consts<double> {
    static constexpr double PI = 3.14159265358979323846;
    // the FP64 mode keeps the reference's FP32 thresholds
    static double eps() { return (double)std::numeric_limits<float>::epsilon(); }
};

// ---------------------------------------------------------------------------
// CubicSplineKernel3D (kernel/CubicSplineKernel3D.cpp:7-71) and
// CubicSplineKernel (kernel/CubicSplineKernel.cpp:9-80).
template <class T, int D>
struct CubicSpline {
    T h = 0, alpha = 0;
    void initialize(T kernel_support) {
        h = kernel_support / T(2);
        if (D == 3)
            alpha = T(1) / (T(4) * Consts<T>::PI * (h * h * h));
        else
            alpha = T(5) / (T(14) * Consts<T>::PI * (h * h));
    }
    T value(const Vec<T, D>& position) const {
        T q = length(position) / h;
        T ret = T(0);
        if (q < T(1)) {
            T a = T(2) - q, b = T(1) - q;
            ret = a * a * a - T(4) * (b * b * b);
        } else if (q < T(2)) {
            T a = T(2) - q;
            ret = a * a * a;
        }
        return alpha * ret;
    }
    // GetKernelDerivativeValue(position)
    Vec<T, D> derivative(const Vec<T, D>& position) const {
        T len = length(position);
        T q = len / h;
        Vec<T, D> pre = vzero<T, D>();
        if (len > Consts<T>::eps()) pre = (h * position) / len;
        Vec<T, D> ret = vzero<T, D>();
        if (q < T(1)) {
            T a = T(2) - q, b = T(1) - q;
            ret = pre * (T(-3) * (a * a) + T(12) * (b * b));
        } else if (q < T(2)) {
            T a = T(2) - q;
            ret = pre * (T(-3) * (a * a));
        }
        return (alpha * ret) * T(-1);
    }
    // GetKernelValue(neighborPosition, position) = value(position - neighborPosition)
    T value(const Vec<T, D>& neighbor, const Vec<T, D>& position) const { return value(position - neighbor); }
    // GetKernelDerivativeReversedValue(neighborPosition, position)
    Vec<T, D> reversed(const Vec<T, D>& neighbor, const Vec<T, D>& position) const {
        return derivative(position - neighbor) * T(-1);
    }
};

// ---------------------------------------------------------------------------
// libmorton morton3D_64_encode as used at CompressedNeighbors.cpp:52-54: x
// bits land at 3k (22 bits kept), y at 3k+1 and z at 3k+2 (21 bits kept);
// higher coordinate bits fall off the 64-bit key.
// The low 21 bits of v spread to every third bit (bit b to bit 3b).
static inline uint64_t spread3_21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}
static inline uint64_t morton3_64(uint32_t x, uint32_t y, uint32_t z) {
    // bits 0..20 of each axis, and x's bit 21 at bit 63
    return spread3_21(x) | spread3_21(y) << 1 | spread3_21(z) << 2 | (uint64_t)((x >> 21) & 1u) << 63;
}
static inline uint32_t to_unsigned(int32_t v) {
    int64_t v2 = (int64_t)v;
    v2 -= std::numeric_limits<int32_t>::min();
    return (uint32_t)v2;
}

// CompressedNeighborhoodSearch::NeighborStorage (CompressedNeighbors.hpp:26-60,
// CompressedNeighbors.cpp:260-379): first neighbor, 2-bit control per delta
// (00 -> 1, 01 -> 2, 10 -> one byte d-3, 11 -> four bytes LE d-259).
struct NeighborStorage {
    static constexpr size_t MAX_DELTAS = 96, DELTAS_SIZE = 96, MAX_CONTROLS = 192;
    uint64_t first_neighbor = (uint64_t)-1;
    std::bitset<MAX_CONTROLS> control;
    uint8_t deltas[DELTAS_SIZE] = {};
    size_t size_value = 0, bytes = 0;
    bool overflow = false;

    void clear() {
        control.reset();
        bytes = 0;
        size_value = 0;
        first_neighbor = (uint64_t)-1;
        std::memset(deltas, 0, sizeof(deltas));
        overflow = false;
    }
    void set_first(uint64_t index) {
        first_neighbor = index;
        size_value = 1;
        bytes = 0;
    }
    // returns false where the reference would assert (capacity exceeded)
    bool set_next(uint64_t delta) {
        if (size_value - 1 >= MAX_DELTAS || delta == 0) {
            overflow = true;
            return false;
        }
        size_t bp = size_value - 1;
        if (delta == 1) {
            control.set(bp * 2 + 0, false);
            control.set(bp * 2 + 1, false);
        } else if (delta == 2) {
            control.set(bp * 2 + 0, false);
            control.set(bp * 2 + 1, true);
        } else if (delta <= 256 + 2) {
            if (bytes >= DELTAS_SIZE) {
                overflow = true;
                return false;
            }
            control.set(bp * 2 + 0, true);
            control.set(bp * 2 + 1, false);
            deltas[bytes] = (uint8_t)(delta - 3);
            bytes++;
        } else {
            if (bytes + 3 >= DELTAS_SIZE) {
                overflow = true;
                return false;
            }
            control.set(bp * 2 + 0, true);
            control.set(bp * 2 + 1, true);
            uint32_t r = (uint32_t)(delta - (256 + 3));
            deltas[bytes + 0] = r & 0xff;
            deltas[bytes + 1] = (r >> 8) & 0xff;
            deltas[bytes + 2] = (r >> 16) & 0xff;
            deltas[bytes + 3] = (r >> 24) & 0xff;
            bytes += 4;
        }
        size_value++;
        return true;
    }
    size_t size() const { return size_value; }
    // get_delta rescans the controls before k to find the payload offset, as
    // the reference does (CompressedNeighbors.cpp:335-379).
    uint32_t get_delta(size_t k) const {
        if (!control[k * 2 + 0]) return control[k * 2 + 1] ? 2u : 1u;
        size_t start = 0;
        for (size_t i = 0; i < k; i++) {
            if (!control[i * 2 + 0]) continue;
            start += control[i * 2 + 1] ? 4 : 1;
        }
        if (!control[k * 2 + 1]) return (uint32_t)deltas[start] + 3u;
        uint32_t v = (uint32_t)deltas[start + 3] << 24 | (uint32_t)deltas[start + 2] << 16 |
                     (uint32_t)deltas[start + 1] << 8 | (uint32_t)deltas[start + 0];
        return v + 3u + 256u;
    }
    // delta k given the payload offset of delta k (advanced past it)
    uint32_t next_delta(size_t k, size_t& off) const {
        if (!control[k * 2 + 0]) return control[k * 2 + 1] ? 2u : 1u;
        if (!control[k * 2 + 1]) return (uint32_t)deltas[off++] + 3u;
        uint32_t v = (uint32_t)deltas[off + 3] << 24 | (uint32_t)deltas[off + 2] << 16 |
                     (uint32_t)deltas[off + 1] << 8 | (uint32_t)deltas[off + 0];
        off += 4;
        return v + 3u + 256u;
    }
    // encoded size model: 4 B first index + 2-bit controls + payload bytes
    size_t encoded_bytes() const {
        if (size_value == 0) return 0;
        return 4 + (2 * (size_value - 1) + 7) / 8 + bytes;
    }
};

enum SearchKind {
    SEARCH_COMPRESSED3D = 0,
    SEARCH_QUADRATIC3D = 1,
    SEARCH_HASHED2D = 2,
    SEARCH_QUADRATIC2D = 3,
    SEARCH_HASHED3D = 4
};

// ---- Sort::quick_sort (ParticleCollectionAlgorithm.hpp:28-127) ----------------
// Hoare partition around the key of the middle element, recursing on [low, j]
// and [j + 1, high]; the two halves are disjoint, so processing them in any
// order (the reference splits large ones onto threads) gives one permutation.
// perm[k] = old index now at position k.
static std::vector<size_t> quick_sort_perm(const std::vector<uint64_t>& keys_in) {
    const size_t n = keys_in.size();
    std::vector<uint64_t> key = keys_in;
    std::vector<size_t> perm(n);
    std::iota(perm.begin(), perm.end(), (size_t)0);
    if (n < 2) return perm;
    std::vector<std::pair<int64_t, int64_t>> todo{{0, (int64_t)n - 1}};
    while (!todo.empty()) {
        const int64_t low = todo.back().first, high = todo.back().second;
        todo.pop_back();
        if (low >= high) continue;
        const uint64_t pivot = key[low + (high - low) / 2];
        int64_t i = low - 1, j = high + 1;
        for (;;) {
            do {
                i++;
            } while (key[i] < pivot);
            do {
                j--;
            } while (key[j] > pivot);
            if (i >= j) break;
            std::swap(key[i], key[j]);
            std::swap(perm[i], perm[j]);
        }
        todo.push_back({low, j});
        todo.push_back({j + 1, high});
    }
    return perm;
}
enum { T_NORMAL = 0, T_BOUNDARY = 1, T_INACTIVE = 2 };

// std::abs((int)e) as the reference's x86-64 build evaluates it
// (IISPHFluidSolver.hpp:400-402, where the float error meets the integer abs):
// cvttss2si truncates toward zero in range and gives INT_MIN for NaN or a value
// beyond the int range, and abs(INT_MIN) stays INT_MIN (the negation wraps).
// C leaves both cases undefined, so they are spelled out.  The float error is
// exact in double, so one double form serves the FP32 and FP64 modes.
static inline double x86_int_abs(double e) {
    if (!(e > -2147483649.0 && e < 2147483648.0)) return -2147483648.0;
    const long long v = (long long)e;  // truncation toward zero, in int range
    if (v == -2147483648LL) return -2147483648.0;
    return (double)(v < 0 ? -v : v);
}

struct SceneBase {
    virtual ~SceneBase() {}
    virtual void translate(const float* delta) = 0;
    virtual int find_neighbors() = 0;
    virtual int simulate(float dt, fsa_step_stats* st) = 0;
    virtual int download(const fsa_particle_view* v) = 0;
    virtual int download_iisph(void* rec, uint64_t stride) = 0;
    virtual int export_neighbors(uint64_t* off, uint64_t* cols, uint64_t cap, uint64_t* nnz) = 0;
    virtual void codec_stats(uint64_t* total, uint64_t* bytes) = 0;
    virtual void codec_sizes(uint64_t* count_per_row, uint64_t* bytes_per_row) = 0;
    virtual void internal_ids(uint64_t* out) = 0;
    virtual int export_raw(uint64_t* off, uint64_t* cols, uint64_t cap, uint64_t* nnz) = 0;
    virtual int upload(const fsa_particle_view* v) = 0;
    virtual float next_timestep() const = 0;
    virtual void set_entities(const fsa_entity* e, uint32_t count) = 0;
    virtual void get_statistics(float zero_height, fsa_scene_statistics* out) const = 0;
    virtual int query_neighbors(const float* x, uint64_t count, uint64_t* off, uint64_t* rows, uint64_t cap,
                                uint64_t* nnz) = 0;
    virtual int sensor_plane(const fsa_sensor_plane* p, float* values) = 0;
    virtual int evaluate_volume(const float* x, uint64_t count, fsa_volume_sample* out) = 0;
    // IISPH phases one at a time (tests of the solver's algebra): phases 1 and 2 at dt
    // over the last search; the operator A p of the Jacobi passes for pressures p;
    // the per-particle fields of phases 1 and 2.  Arrays are in caller order.
    virtual int iisph_prepare(float dt, int solve) = 0;
    virtual int iisph_apply(float dt, const double* p, double* ap, double* acc_out) = 0;
    virtual void iisph_fields(double* out) const = 0;
};

template <class T, int D>
struct Scene : SceneBase {
    using V = Vec<T, D>;
    fsa_scene_desc desc;
    int search;
    size_t n = 0;
    // ParticleCollection components (Particle.hpp:15-50) + stable id
    std::vector<V> pos, vel, acc, ext;
    std::vector<T> mass, pressure, density, density_old;
    std::vector<uint32_t> tag;
    std::vector<uint8_t> type;
    std::vector<uint64_t> id;
    // IISPHParticleData3D (IISPHFluidSolver3D.hpp:13-19) / IISPHParticleData
    std::vector<V> vstar, anp;
    std::vector<T> source, diag, perr;
    // neighbor structures
    std::vector<NeighborStorage> storage;          // compressed
    std::vector<uint64_t> cell_key;               // compressed: ParticleInformation.cell_index
    std::vector<std::pair<uint64_t, size_t>> runs; // cell_to_particle_map
    std::vector<std::vector<uint64_t>> lists;     // quadratic / hashed lists
    // hashed 3D grid state (HashedNeighborhoodSearch3D.cpp:69-127): GridCellState.current
    // of every particle (INT_MIN until its first grid update) and the call counter of
    // improve_cache_efficiency
    std::vector<std::array<int32_t, 3>> grid_current;
    size_t calls_since_sort = 0;
    // the grid of the last hashed 3D search: each cell's non-inactive particles
    // ascending.  The reference's grid also keeps particles that turned inactive
    // after their insertion (update_grid skips them, HashedNeighborhoodSearch3D.cpp:111-114);
    // this grid holds the non-inactive particles of the last search only.
    std::map<std::array<int32_t, 3>, std::vector<uint64_t>> grid3;
    // hashed 2D grid state (HashedNeighborhoodSearch.cpp:117-151)
    std::map<std::pair<int32_t, int32_t>, std::list<uint64_t>> grid;
    std::vector<std::pair<int32_t, int32_t>> current_status;
    bool searched = false;
    CubicSpline<T, D> kernel;
    T search_radius = 0;
    // false: the reference's FP32 serial sum; true: FP64 sum (the device's
    // convergence reduction), in the same index order
    bool fp64_reduction = false;

    // neighbor iteration in the reference's order
    template <class F>
    void for_neighbors(size_t i, F&& fn) const {
        if (search == SEARCH_COMPRESSED3D) {
            const NeighborStorage& s = storage[i];
            if (s.size() == 0) return;
            // the iterator of CompressedNeighbors.cpp:381-486: the same deltas as
            // get_delta, with the payload offset carried along instead of rescanned
            uint64_t cur = s.first_neighbor;
            fn((size_t)cur);
            size_t off = 0;
            for (size_t c = 1; c < s.size(); c++) {
                cur += s.next_delta(c - 1, off);
                fn((size_t)cur);
            }
        } else {
            for (uint64_t j : lists[i]) fn((size_t)j);
        }
    }

    int load(const fsa_scene_desc& d, int search_kind, const fsa_particle_view* v) {
        desc = d;
        search = search_kind;
        n = v->count;
        pos.assign(n, vzero<T, D>());
        vel = pos;
        acc = pos;
        ext = pos;
        vstar = pos;
        anp = pos;
        mass.assign(n, 0);
        pressure = mass;
        density = mass;
        density_old = mass;
        source = mass;
        diag = mass;
        perr = mass;
        tag.assign(n, 0);
        type.assign(n, 0);
        id.resize(n);
        for (size_t i = 0; i < n; i++) {
            const float* mv = (const float*)((const char*)v->movement + i * v->movement_stride);
            for (int k = 0; k < D; k++) {
                pos[i].v[k] = (T)mv[k];
                vel[i].v[k] = (T)mv[D + k];
                acc[i].v[k] = (T)mv[2 * D + k];
            }
            const float* pd = (const float*)((const char*)v->particle_data + i * v->particle_data_stride);
            mass[i] = pd[0];
            pressure[i] = pd[1];
            density[i] = pd[2];
            const char* inf = (const char*)v->info + i * v->info_stride;
            std::memcpy(&tag[i], inf, 4);
            type[i] = (uint8_t)inf[4];
            if (v->external) {
                const float* e = (const float*)((const char*)v->external + i * v->external_stride);
                for (int k = 0; k < D; k++) ext[i].v[k] = (T)e[k];
            }
            id[i] = i;
        }
        search_radius = (T)desc.particle_size * T(2);  // Math::kernel_support_factor
        kernel.initialize((T)desc.particle_size * T(2));
        return FSA_OK;
    }

    // apply permutation (new position k takes old index perm[k]) to every component
    template <class A>
    static void permute(std::vector<A>& a, const std::vector<size_t>& perm) {
        if (a.size() != perm.size()) return;  // component not in use by this search
        std::vector<A> b(a.size());
        for (size_t k = 0; k < perm.size(); k++) b[k] = a[perm[k]];
        a.swap(b);
    }
    void permute_all(const std::vector<size_t>& perm) {
        permute(pos, perm);
        permute(vel, perm);
        permute(acc, perm);
        permute(ext, perm);
        permute(mass, perm);
        permute(pressure, perm);
        permute(density, perm);
        permute(tag, perm);
        permute(type, perm);
        permute(id, perm);
        permute(vstar, perm);
        permute(anp, perm);
        permute(source, perm);
        permute(diag, perm);
        permute(perr, perm);
        permute(cell_key, perm);
        permute(storage, perm);
        permute(grid_current, perm);
    }

    // ---- CompressedNeighborhoodSearch::find_neighbors (CompressedNeighbors.cpp:168-240)
    std::array<int32_t, 3> cell_of(const V& p) const {
        std::array<int32_t, 3> c{0, 0, 0};
        for (int k = 0; k < D; k++) c[k] = (int32_t)std::floor(p.v[k] / search_radius);
        return c;
    }
    static uint64_t key_of(const std::array<int32_t, 3>& c) {
        return morton3_64(to_unsigned(c[0]), to_unsigned(c[1]), to_unsigned(c[2]));
    }
    size_t run_lookup(uint64_t key) const {
        // exact-match binary search over the run table (CompressedNeighbors.cpp:58-81)
        size_t lo = 0, hi = runs.size();
        while (lo < hi) {
            size_t mid = lo + (hi - lo) / 2;
            if (runs[mid].first > key)
                hi = mid;
            else if (runs[mid].first < key)
                lo = mid + 1;
            else
                return runs[mid].second;
        }
        return (size_t)-1;
    }
    int search_compressed() {
        cell_key.resize(n);
        storage.resize(n);
        loop_for(n, [&](size_t i) {
            cell_key[i] = type[i] == T_INACTIVE ? (uint64_t)-1 : key_of(cell_of(pos[i]));
        });
        std::vector<size_t> perm(n);
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return cell_key[a] < cell_key[b]; });
        permute_all(perm);
        runs.clear();
        for (size_t i = 0; i < n; i++) {
            if (type[i] == T_INACTIVE) break;
            if (i == 0 || cell_key[i - 1] != cell_key[i]) runs.push_back({cell_key[i], i});
        }
        T r2 = search_radius * search_radius;  // Math::pow2(search_radius)
        std::vector<int> bad(n, 0);
        loop_for(n, [&](size_t i) {
            if (type[i] == T_INACTIVE) return;
            NeighborStorage& st = storage[i];
            auto c = cell_of(pos[i]);
            std::array<uint64_t, 27> keys;
            int q = 0;
            for (int x = -1; x <= 1; x++)
                for (int y = -1; y <= 1; y++)
                    for (int z = -1; z <= 1; z++) keys[q++] = key_of({c[0] + x, c[1] + y, c[2] + z});
            std::sort(keys.begin(), keys.end());
            st.clear();
            size_t last = (size_t)-1;
            for (int k = 0; k < 27; k++) {
                size_t cur = run_lookup(keys[k]);
                if (cur == (size_t)-1) continue;
                while (keys[k] == cell_key[cur]) {
                    V diff = pos[cur] - pos[i];
                    if (dot(diff, diff) <= r2) {
                        if (last == (size_t)-1) {
                            last = cur;
                            st.set_first(last);
                        } else {
                            if (!st.set_next(cur - last)) bad[i] = 1;
                            last = cur;
                        }
                    }
                    cur++;
                    if (cur >= n) break;
                }
            }
        });
        for (size_t i = 0; i < n; i++)
            if (bad[i]) {
                g_err = "neighbor capacity exceeded (more than 97 neighbors or 96 delta bytes)";
                return FSA_E_CAPACITY;
            }
        return FSA_OK;
    }

    // ---- HashedNeighborhoodSearch3D::find_neighbors (HashedNeighborhoodSearch3D.cpp:69-217)
    // Every 101st call the collection is re-sorted with the unstable quick sort
    // by the Morton key of each particle's cell of the last grid update
    // (improve_cache_efficiency, :219-247).  The grid then maps each cell to the
    // ascending set of its non-inactive particles, and lists walk the 27 cells
    // dx-outer, dz-inner.
    int search_hashed3d() {
        if (grid_current.size() != n) {
            const int32_t lo = std::numeric_limits<int32_t>::min();
            grid_current.assign(n, {lo, lo, lo});
        }
        calls_since_sort++;
        if (calls_since_sort > 100) {
            std::vector<uint64_t> keys(n);
            for (size_t i = 0; i < n; i++) keys[i] = key_of(grid_current[i]);
            permute_all(quick_sort_perm(keys));
            calls_since_sort = 0;
        }
        grid3.clear();
        auto& grid = grid3;
        for (size_t i = 0; i < n; i++) {
            if (type[i] == T_INACTIVE) continue;
            grid_current[i] = cell_of(pos[i]);
            grid[grid_current[i]].push_back(i);  // ascending, as std::set
        }
        lists.assign(n, {});
        const T r2 = search_radius * search_radius;
        loop_for(n, [&](size_t i) {
            if (type[i] == T_INACTIVE) return;
            const auto c = grid_current[i];
            for (int dx = -1; dx <= 1; dx++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dz = -1; dz <= 1; dz++) {
                        auto it = grid.find({c[0] + dx, c[1] + dy, c[2] + dz});
                        if (it == grid.end()) continue;
                        for (uint64_t j : it->second) {
                            V d = pos[i] - pos[j];
                            if (dot(d, d) <= r2) lists[i].push_back(j);
                        }
                    }
        });
        return FSA_OK;
    }

    // ---- QuadraticNeighborhoodSearch3D::find_neighbors (QuadraticNeighborhoodSearch3D.cpp:32-69)
    int search_quadratic3d() {
        lists.assign(n, {});
        T r2 = search_radius * search_radius;
        loop_for(n, [&](size_t i) {
            if (type[i] == T_INACTIVE) return;
            for (size_t j = 0; j < n; j++) {
                if (type[j] == T_INACTIVE) continue;
                V d = pos[i] - pos[j];
                if (dot(d, d) <= r2) lists[i].push_back(j);
            }
        });
        return FSA_OK;
    }

    // ---- HashedNeighborhoodSearch::find_neighbors (HashedNeighborhoodSearch.cpp:9-151)
    std::pair<int32_t, int32_t> cell2(const V& p) const {
        T cs = search_radius;  // grid_cell_size = search_radius (initialize, :153-157)
        return {(int32_t)(p.v[0] / cs), (int32_t)(p.v[1] / cs)};
    }
    int search_hashed2d() {
        if (current_status.size() < n) current_status.resize(n, {INT32_MAX, INT32_MAX});
        lists.resize(n);
        uint16_t count = (uint16_t)n;  // particleAmount_t truncation (:129)
        for (uint64_t i = 0; i < count; i++) {
            auto key = cell2(pos[i]);
            if (current_status[i] == key) continue;
            grid[current_status[i]].remove(i);
            current_status[i] = key;
            grid[key].push_back(i);
        }
        loop_for(n, [&](size_t i) {
            if (type[i] == T_BOUNDARY || type[i] == T_INACTIVE) return;
            auto c = cell2(pos[i]);
            int32_t rc = (int32_t)std::ceil(search_radius / search_radius);
            lists[i].clear();
            for (int32_t x = c.first - rc; x <= c.first + rc; x++)
                for (int32_t y = c.second - rc; y <= c.second + rc; y++) {
                    auto it = grid.find({x, y});
                    if (it == grid.end()) continue;
                    for (uint64_t j : it->second) {
                        V d = pos[i] - pos[j];
                        if (length(d) <= search_radius) lists[i].push_back(j);
                    }
                }
        });
        return FSA_OK;
    }
    // ---- QuadraticNeighborhoodSearchDynamicAllocated (QuadraticNeighborhoodSearchDynamicAllocated.cpp:11-38)
    int search_quadratic2d() {
        lists.assign(n, {});
        loop_for(n, [&](size_t i) {
            for (size_t j = 0; j < n; j++) {
                V d = pos[i] - pos[j];
                if (length(d) <= search_radius) lists[i].push_back(j);
            }
        });
        return FSA_OK;
    }

    void translate(const float* delta) override {
        for (size_t i = 0; i < n; i++)
            for (int k = 0; k < D; k++) pos[i].v[k] = pos[i].v[k] + (T)delta[k];
    }

    int find_neighbors() override {
        int rc;
        if (search == SEARCH_COMPRESSED3D)
            rc = search_compressed();
        else if (search == SEARCH_QUADRATIC3D)
            rc = search_quadratic3d();
        else if (search == SEARCH_HASHED3D)
            rc = search_hashed3d();
        else if (search == SEARCH_HASHED2D)
            rc = search_hashed2d();
        else
            rc = search_quadratic2d();
        searched = rc == FSA_OK;
        return rc;
    }

    // ---- get_non_cfl_validating_timestep of the scene's generator:
    // ConstantTimestepGenerator.cpp:17-41 (falls back to timestep) or
    // DynamicCflTimestepGenerator.cpp:84-109 (falls back to max_timestep)
    float cfl_timestep(float max_a, float max_v) const {
        const float fallback = desc.timestep_generator == FSA_TIMESTEP_DYNAMIC_CFL ? desc.max_timestep : desc.timestep;
        float dta = fallback, dtv = fallback;
        // unqualified sqrt on a float resolves to ::sqrt(double); the product
        // with the float lambda is formed in double and stored to float
        if (max_a > 0.0f) dta = (float)(::sqrt((double)(desc.particle_size / max_a)) * desc.lambda_a);
        if (max_v > 0.0f) dtv = desc.particle_size / max_v * desc.lambda_v;
        float nt = std::fmin(dta, dtv);
        return std::fmax(nt, 0.000001f);  // MIN_ALLOWED_TIMESTEP (TimestepGenerator.hpp:18)
    }

    // ---- the generator's next desired time step (Simulator.cpp:17-24):
    // ConstantTimestepGenerator::generate_next_timestep (ConstantTimestepGenerator.cpp:4-7) or
    // DynamicCflTimestepGenerator::generate_next_timestep (DynamicCflTimestepGenerator.cpp:46-62)
    // with calculate_maximum_velocity_and_acceleration (:6-43) over the collection.
    float generated_timestep() const {
        if (desc.timestep_generator != FSA_TIMESTEP_DYNAMIC_CFL) return desc.timestep;
        float max_v = 0.0f, max_a = 0.0f;
        for (size_t i = 0; i < n; i++) {
            if (type[i] == T_INACTIVE) continue;
            max_v = std::max(max_v, (float)std::sqrt(dot(vel[i], vel[i])));
            max_a = std::max(max_a, (float)std::sqrt(dot(acc[i], acc[i])));
        }
        float ts = desc.min_timestep;
        const float eps = std::numeric_limits<float>::epsilon();
        if (max_v > eps && max_a > eps) ts = std::max(ts, std::min(desc.max_timestep, cfl_timestep(max_a, max_v)));
        return ts;
    }
    float next_timestep() const override { return generated_timestep(); }
    void set_entities(const fsa_entity* e, uint32_t count) override { entities.assign(e, e + count); }
    void get_statistics(float zero_height, fsa_scene_statistics* out) const override { statistics(zero_height, out); }

    // ---- NeighborhoodInterface::get_neighbors(position), 3D
    // compressed: find_neighbors_and_save_in_storage(position) (CompressedNeighbors.cpp:536-595):
    //   the 27 cell keys sorted, each run walked while its key matches the
    //   particle's stored cell_index, test dot(p - x, p - x) <= pow2(search_radius);
    // hashed: NeighborsIterator::operator++ (HashedNeighborhoodSearch3D.cpp:272-349):
    //   cells dx, dy, dz from -1 to 1 (dz fastest), members in set order,
    //   test length(x - p) <= search_radius;
    // quadratic: NeighborsIterator::operator++ (QuadraticNeighborhoodSearch3D.cpp:199-219):
    //   every particle by index, inactive ones included, test length(x - p) <= search_radius.
    template <class F>
    void for_position(const V& x, F&& fn) const {
        if (search == SEARCH_COMPRESSED3D) {
            const auto c = cell_of(x);
            std::array<uint64_t, 27> keys;
            int q = 0;
            for (int a = -1; a <= 1; a++)
                for (int b = -1; b <= 1; b++)
                    for (int z = -1; z <= 1; z++) keys[q++] = key_of({c[0] + a, c[1] + b, c[2] + z});
            std::sort(keys.begin(), keys.end());
            const T r2 = search_radius * search_radius;
            for (int k = 0; k < 27; k++) {
                size_t cur = run_lookup(keys[k]);
                if (cur == (size_t)-1) continue;
                while (keys[k] == cell_key[cur]) {
                    V diff = pos[cur] - x;
                    if (dot(diff, diff) <= r2) fn(cur);
                    cur++;
                    if (cur >= n) break;
                }
            }
        } else if (search == SEARCH_HASHED3D) {
            const auto c = cell_of(x);
            for (int dx = -1; dx <= 1; dx++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dz = -1; dz <= 1; dz++) {
                        auto it = grid3.find({c[0] + dx, c[1] + dy, c[2] + dz});
                        if (it == grid3.end()) continue;
                        for (uint64_t j : it->second)
                            if (length(x - pos[j]) <= search_radius) fn((size_t)j);
                    }
        } else {
            for (size_t j = 0; j < n; j++)
                if (length(x - pos[j]) <= search_radius) fn(j);
        }
    }
    bool position_queries_supported() const {
        if (D == 3 && (search == SEARCH_COMPRESSED3D || search == SEARCH_HASHED3D || search == SEARCH_QUADRATIC3D))
            return true;
        g_err = "position queries are defined for the 3D searches";
        return false;
    }
    V point(const float* x) const {
        V p = vzero<T, D>();
        for (int k = 0; k < D; k++) p.v[k] = (T)x[k];
        return p;
    }
    int query_neighbors(const float* x, uint64_t count, uint64_t* off, uint64_t* rows, uint64_t cap,
                        uint64_t* nnz) override {
        if (!position_queries_supported()) return FSA_E_INVALID;
        uint64_t total = 0;
        if (off) off[0] = 0;
        for (uint64_t k = 0; k < count; k++) {
            for_position(point(x + 3 * k), [&](size_t j) {
                if (rows && total < cap) rows[total] = id[j];
                total++;
            });
            if (off) off[k + 1] = total;
        }
        *nnz = total;
        if (rows && total > cap) {
            g_err = "output capacity too small";
            return FSA_E_INVALID;
        }
        return FSA_OK;
    }
    // SensorPlane::calculate_for_timepoint (SensorPlane.cpp:75-137): glm::normalize
    // = v * (1 / sqrt(dot(v, v))), the sample position
    // origin + span_x * x_step * (x + sub_x) + span_y * y_step * (y + sub_y),
    // value += (mass, 0, 0) | velocity | (0, 0, pressure) times
    // GetKernelValue(p, sample), value /= Math::pow2(float(sub + 1)).
    int sensor_plane(const fsa_sensor_plane* sp, float* values) override {
        if (!position_queries_supported()) return FSA_E_INVALID;
        auto normalize = [](const float* v) {
            V r = vzero<T, D>();
            const T d = ((T)v[0] * (T)v[0] + (T)v[1] * (T)v[1]) + (T)v[2] * (T)v[2];
            const T inv = T(1) / std::sqrt(d);
            for (int k = 0; k < D; k++) r.v[k] = (T)v[k] * inv;
            return r;
        };
        const V span_x = normalize(sp->span_x), span_y = normalize(sp->span_y);
        const V origin = point(sp->origin);
        const T x_step = (T)sp->width / (T)sp->number_of_samples_x;
        const T y_step = (T)sp->height / (T)sp->number_of_samples_y;
        const size_t nx = sp->number_of_samples_x, ny = sp->number_of_samples_y, sub = sp->sub_sample_grid_size;
        for (size_t i = 0; i < nx * ny; i++) {
            const size_t y = ny - 1 - i / nx, x = i % nx;
            V value = vzero<T, D>();
            for (size_t sxc = 0; sxc <= sub; sxc++) {
                const T sub_x = (T)sxc / ((T)sub + T(1)) - T(0.5);
                for (size_t syc = 0; syc <= sub; syc++) {
                    const T sub_y = (T)syc / ((T)sub + T(1)) - T(0.5);
                    const V sample = origin + span_x * x_step * ((T)x + sub_x) + span_y * y_step * ((T)y + sub_y);
                    for_position(sample, [&](size_t j) {
                        const T w = kernel.value(pos[j], sample);
                        V term = vzero<T, D>();
                        if (sp->sensor_type == FSA_SENSOR_PLANE_DENSITY) term.v[0] = mass[j];
                        else if (sp->sensor_type == FSA_SENSOR_PLANE_VELOCITY) term = vel[j];
                        else term.v[D - 1] = pressure[j];  // z (3D only)
                        value = value + term * w;
                    });
                }
            }
            const T g = (T)(sub + 1);
            value = value / (g * g);
            for (int k = 0; k < D; k++) values[3 * i + k] = (float)value.v[k];
        }
        return FSA_OK;
    }
    // ParticleIntersectionAccelerator::evaluate_volume_at_position
    // (ParticleIntersectionAccelerator.cpp:97-131)
    int evaluate_volume(const float* xs, uint64_t count, fsa_volume_sample* out) override {
        if (!position_queries_supported()) return FSA_E_INVALID;
        for (uint64_t k = 0; k < count; k++) {
            const V x = point(xs + 3 * k);
            T rho = 0, rb = 0, rf = 0;
            V fnrm = vzero<T, D>(), bnrm = vzero<T, D>();
            for_position(x, [&](size_t j) {
                if (type[j] == T_INACTIVE) return;
                const T contribution = mass[j] * kernel.value(pos[j], x);
                const V normal_contribution = mass[j] * kernel.derivative(x - pos[j]);
                rho += contribution;
                if (type[j] == T_NORMAL) {
                    rf += contribution;
                    fnrm = fnrm + normal_contribution;
                } else if (type[j] == T_BOUNDARY) {
                    rb += contribution;
                    bnrm = bnrm + normal_contribution;
                }
            });
            fsa_volume_sample& o = out[k];
            o.density = (float)rho;
            o.boundary_only_density = (float)rb;
            o.fluid_only_density = (float)rf;
            for (int a = 0; a < 3; a++) {
                o.fluid_only_normal[a] = a < D ? (float)fnrm.v[a] : 0.0f;
                o.boundary_only_normal[a] = a < D ? (float)bnrm.v[a] : 0.0f;
            }
        }
        return FSA_OK;
    }

    V gravity_vec() const {
        V g = vzero<T, D>();
        g.v[1] = -(T)desc.gravity;
        return g;
    }

    // IISPHFluidSolver3D::ComputeViscosityAcceleration (IISPHFluidSolver3D.hpp:43-76),
    // SESPHFluidSolver3D::ComputeViscosityAcceleration (SESPHFluidSolver3D.hpp:293-324)
    // and the 2D twins; `rho` is the density array the neighbors are read from.
    V viscosity(size_t i, T nu, const std::vector<T>& rho) const {
        V tmp = vzero<T, D>();
        const V& x = pos[i];
        const V& v = vel[i];
        T ps = (T)desc.particle_size;
        for_neighbors(i, [&](size_t j) {
            if (type[j] == T_INACTIVE) return;
            T nm = mass[j], nd = rho[j];
            if (nd == T(0)) return;
            V vij = v - vel[j];
            V xij = x - pos[j];
            tmp = tmp + ((nm / nd) * (dot(vij, xij) / (dot(xij, xij) + T(0.01f) * ps * ps))) *
                            kernel.reversed(pos[j], x);
        });
        return (T(2) * nu) * tmp;
    }

    // ---- IISPHFluidSolver3D::execute_simulation_step (IISPHFluidSolver3D.hpp:120-562),
    // in its phases: phase 1 (:141-212), phase 2 (:216-339), the Jacobi passes
    // (:348-461) and the loop control, CFL and integration (:463-561)
    void iisph3d_phase1(T dt) {
        const bool sl = desc.iisph_single_layer_boundary != 0;
        const T g1 = (T)desc.iisph_single_layer_gamma_1;
        density_old = density;
        loop_for(n, [&](size_t i) {
            uint8_t t = type[i];
            if (t == T_INACTIVE) return;
            pressure[i] = T(0);
            V npa = ext[i];
            if (t != T_BOUNDARY) {
                npa = npa + gravity_vec();
                npa = npa + viscosity(i, (T)desc.iisph_viscosity, density_old);
            }
            vstar[i] = vel[i] + dt * npa;
            anp[i] = npa;
            ext[i] = vzero<T, D>();
            if (t == T_BOUNDARY) return;
            T rho = T(0);
            for_neighbors(i, [&](size_t j) {
                uint8_t tj = type[j];
                if (tj == T_INACTIVE) return;
                T m = mass[j];
                if (sl && tj == T_BOUNDARY) m = m * g1;
                rho = rho + m * kernel.value(pos[j], pos[i]);
            });
            density[i] = rho;
        });
    }
    void iisph3d_phase2(T dt) {
        const T rho0 = (T)desc.rest_density;
        const T rho0sq = rho0 * rho0;
        const bool sl = desc.iisph_single_layer_boundary != 0;
        const T gamma = (T)desc.iisph_gamma;
        loop_for(n, [&](size_t i) {
            if (type[i] != T_NORMAL) return;
            const V& xi = pos[i];
            T contrib = T(0);
            for_neighbors(i, [&](size_t j) {
                uint8_t tj = type[j];
                if (tj == T_NORMAL)
                    contrib = contrib + mass[j] * dot(vstar[i] - vstar[j], kernel.reversed(pos[j], xi));
                else if (tj == T_BOUNDARY)
                    contrib = contrib + mass[j] * dot(g_mutation == 4 ? vstar[i] - vstar[j] : vstar[i],
                                                      kernel.reversed(pos[j], xi));
            });
            source[i] = rho0 - density[i] - dt * contrib;
            V inner = vzero<T, D>();
            for_neighbors(i, [&](size_t j) {
                uint8_t tj = type[j];
                if (tj == T_NORMAL)
                    inner = inner - (mass[j] / rho0sq) * kernel.reversed(pos[j], xi);
                else if (tj == T_BOUNDARY) {
                    if (sl)
                        inner = inner - (T(2) * gamma * mass[j] / rho0sq) * kernel.reversed(pos[j], xi);
                    else
                        inner = inner - ((g_mutation == 1 ? T(1) : T(2)) * mass[j] / rho0sq) * kernel.reversed(pos[j], xi);
                }
            });
            T dg = T(0);
            for_neighbors(i, [&](size_t j) {
                uint8_t tj = type[j];
                if (tj == T_NORMAL) {
                    dg = dg + mass[j] * dot(inner, kernel.reversed(pos[j], xi));
                    dg = dg + mass[j] * dot((mass[i] / rho0sq) * (g_mutation == 2 ? kernel.reversed(pos[j], xi)
                                                                                  : kernel.reversed(xi, pos[j])),
                                            kernel.reversed(pos[j], xi));
                } else if (tj == T_BOUNDARY) {
                    dg = dg + mass[j] * dot(inner, kernel.reversed(pos[j], xi));
                }
            });
            dg = dg * (dt * dt);
            diag[i] = dg;
        });
    }
    // Jacobi pass A (:348-397): pressure accelerations from the current pressures
    void iisph3d_pass_a() {
        const T rho0 = (T)desc.rest_density;
        const T rho0sq = rho0 * rho0;
        const bool sl = desc.iisph_single_layer_boundary != 0;
        const T gamma = (T)desc.iisph_gamma;
        loop_for(n, [&](size_t i) {
            if (type[i] != T_NORMAL) return;
            const V& xi = pos[i];
            V pa = vzero<T, D>();
            for_neighbors(i, [&](size_t j) {
                uint8_t tj = type[j];
                if (tj == T_NORMAL)
                    pa = pa - (mass[j] * (pressure[i] / rho0sq + pressure[g_mutation == 6 ? i : j] / rho0sq)) *
                                  kernel.reversed(pos[j], xi);
                else if (tj == T_BOUNDARY) {
                    if (sl)
                        pa = pa - (gamma * mass[j] * T(2) * pressure[i] / rho0sq) * kernel.reversed(pos[j], xi);
                    else
                        pa = pa - (mass[j] * (g_mutation == 3 ? T(1) : T(2)) * pressure[i] / rho0sq) *
                                      kernel.reversed(pos[j], xi);
                }
            });
            acc[i] = pa;
        });
    }
    // Jacobi pass B's product (:402-440): (A p)_i = dt^2 sum_j m_j (a_i - a_j) . gradW
    T iisph3d_ap(size_t i, T dt) const {
        const V& xi = pos[i];
        T ap = T(0);
        for_neighbors(i, [&](size_t j) {
            uint8_t tj = type[j];
            if (tj == T_NORMAL)
                ap = ap + mass[j] * dot(acc[i] - acc[j], kernel.reversed(pos[j], xi));
            else if (tj == T_BOUNDARY)
                ap = ap + mass[j] * dot(acc[i], kernel.reversed(pos[j], xi));
        });
        return ap * (dt * dt);
    }

    // the Jacobi loop with its convergence test (:346-522); returns the iteration count
    size_t iisph3d_solve(T dt, T& max_v2, T& max_a2, double& avg_out) {
        const T eps = Consts<T>::eps();
        size_t iterations = 0;
        for (size_t it = 0; it < desc.iisph_max_iterations; it++) {
            iisph3d_pass_a();
            loop_for(n, [&](size_t i) {
                if (type[i] != T_NORMAL) return;
                const T ap = iisph3d_ap(i, dt);
                if (std::abs(diag[i]) > eps)
                    pressure[i] = std::fmax(T(0), pressure[i] + (g_mutation == 5 ? T(-1) : T(1)) * (T)desc.iisph_omega *
                                                                   ((source[i] - ap) / diag[i]));
                else
                    pressure[i] = T(0);
                T e = std::abs(source[i] - ap);
                if (std::abs(diag[i]) > eps) {
                    if (std::abs(pressure[i]) <= eps) e = T(0);
                    perr[i] = e;
                }
            });
            iterations = it + 1;
            // serial convergence test in index order (IISPHFluidSolver3D.hpp:467-522)
            T avg = T(0);
            double avg64 = 0.0;
            size_t counter = 0;
            max_v2 = 0;
            max_a2 = 0;
            for (size_t i = 0; i < n; i++) {
                if (type[i] != T_NORMAL) continue;
                max_a2 = std::fmax(max_a2, dot(acc[i], acc[i]));
                V pv = vstar[i] + dt * acc[i];
                max_v2 = std::fmax(max_v2, dot(pv, pv));
                if (std::abs(diag[i]) > eps) {
                    avg = avg + perr[i];
                    avg64 = avg64 + (double)perr[i];
                    counter++;
                }
            }
            if (counter > 0) {
                avg = avg / (T)counter;
                avg64 = avg64 / (double)counter;
            }
            bool done;
            if (fp64_reduction) {
                avg_out = avg64;
                done = avg64 <= (double)desc.iisph_max_density_error_allowed;
            } else {
                avg_out = (double)avg;
                done = avg <= (T)desc.iisph_max_density_error_allowed;
            }
            if (it >= (size_t)desc.iisph_min_iterations - 1) {
                if (done) break;
            }
        }
        return iterations;
    }

    int simulate_iisph3d(float desired, fsa_step_stats* st) {
        T dt = (T)desired;
        iisph3d_phase1(dt);
        iisph3d_phase2(dt);
        T max_v2 = 0, max_a2 = 0;
        double avg_out = 0.0;
        const size_t iterations = iisph3d_solve(dt, max_v2, max_a2, avg_out);
        T old_dt = dt;
        float max_v = (float)std::sqrt((double)max_v2), max_a = (float)std::sqrt((double)max_a2);
        float corrected = std::fmin(cfl_timestep(max_a, max_v), desired);
        T cur = (T)corrected;
        loop_for(n, [&](size_t i) {
            if (type[i] == T_INACTIVE) return;
            V pv = vstar[i];
            if (old_dt != cur) pv = pv - old_dt * anp[i] + cur * anp[i];
            vel[i] = pv + cur * acc[i];
            pos[i] = pos[i] + cur * vel[i];
        });
        if (st) {
            st->desired_time_step = desired;
            st->actual_time_step = corrected;
            st->iterations = (uint32_t)iterations;
            st->status = FSA_OK;
            st->average_density_error = avg_out;
            st->max_velocity = max_v;
            st->max_acceleration = max_a;
        }
        return FSA_OK;
    }

    // ---- SESPHFluidSolver3D::execute_simulation_step (SESPHFluidSolver3D.hpp:49-130)
    // and SESPHFluidSolver (2D, SESPHFluidSolver.hpp:49-105: no single layer, no CFL)
    int simulate_sesph(float desired, fsa_step_stats* st) {
        T dt = (T)desired;
        const T rho0 = (T)desc.rest_density, k = (T)desc.sesph_stiffness_k;
        const bool sl = D == 3 && desc.sesph_single_layer_boundary != 0;
        const T g1 = (T)desc.sesph_single_layer_gamma_1, g2 = (T)desc.sesph_single_layer_gamma_2;
        loop_for(n, [&](size_t i) {
            if (type[i] != T_NORMAL) return;
            T rho = T(0);
            for_neighbors(i, [&](size_t j) {
                uint8_t tj = type[j];
                if (tj == T_INACTIVE) return;
                if (tj == T_BOUNDARY && sl)
                    rho = rho + g1 * mass[j] * kernel.value(pos[j], pos[i]);
                else
                    rho = rho + mass[j] * kernel.value(pos[j], pos[i]);
            });
            density[i] = rho;
            T p = k * (rho / rho0 - T(1));
            pressure[i] = g_mutation == 8 ? p : std::max(p, T(0));
        });
        loop_for(n, [&](size_t i) {
            if (type[i] != T_NORMAL) return;
            V npa = vzero<T, D>();
            npa = npa + gravity_vec();
            npa = npa + viscosity(i, (T)desc.sesph_viscosity, density);
            const V& xi = pos[i];
            T mi = mass[i], di = density[i], pi = pressure[i];
            T pdd = di == T(0) ? T(0) : pi / (di * di);
            V pa = vzero<T, D>();
            for_neighbors(i, [&](size_t j) {
                uint8_t tj = type[j];
                if (tj == T_INACTIVE) return;
                if (tj == T_BOUNDARY) {
                    if (sl)
                        pa = pa + ((-mi * (pdd + pdd)) * kernel.reversed(pos[j], xi)) * g2;
                    else
                        pa = pa + (-mi * (pdd + pdd)) * kernel.reversed(pos[j], xi);
                } else {
                    T nd = density[j], np = pressure[j];
                    T npdd = nd == T(0) ? T(0) : np / (nd * nd);
                    pa = pa + (-mass[j] * (pdd + (g_mutation == 7 ? pdd : npdd))) * kernel.reversed(pos[j], xi);
                }
            });
            acc[i] = pa + npa;
        });
        float corrected = desired;
        float max_v = 0.0f, max_a = 0.0f;
        if (D == 3) {
            // serial scan over every particle, boundary and inactive included (:96-107)
            T mv2 = 0, ma2 = 0;
            for (size_t i = 0; i < n; i++) {
                ma2 = std::fmax(ma2, dot(acc[i], acc[i]));
                V v = vel[i] + dt * acc[i];
                mv2 = std::fmax(mv2, dot(v, v));
            }
            max_v = (float)std::sqrt((double)mv2);
            max_a = (float)std::sqrt((double)ma2);
            corrected = std::fmin(cfl_timestep(max_a, max_v), desired);
        }
        T cur = (T)corrected;
        loop_for(n, [&](size_t i) {
            if (type[i] == T_INACTIVE) return;
            vel[i] = vel[i] + cur * acc[i];
            pos[i] = pos[i] + cur * vel[i];
        });
        if (st) {
            st->desired_time_step = desired;
            st->actual_time_step = corrected;
            st->iterations = 0;
            st->status = FSA_OK;
            st->average_density_error = 0.0;
            st->max_velocity = max_v;
            st->max_acceleration = max_a;
        }
        return FSA_OK;
    }

    // ---- IISPHFluidSolver (2D) (IISPHFluidSolver.hpp:139-180, 196-485), in its
    // phases as the 3D solver above
    void iisph2d_phase1(T dt) {
        density_old = density;
        loop_for(n, [&](size_t i) {
            uint8_t t = type[i];
            if (t == T_INACTIVE) return;
            if (t != T_BOUNDARY) {  // CalculateDensity
                T rho = T(0);
                for_neighbors(i, [&](size_t j) {
                    if (type[j] == T_INACTIVE) return;
                    rho = rho + mass[j] * kernel.value(pos[j], pos[i]);
                });
                density[i] = rho;
            }
            V npa = ext[i];  // CalculateNonPressureAccelerationAndPredictedVelocity
            if (t != T_BOUNDARY) {
                npa = npa + gravity_vec();
                npa = npa + viscosity(i, (T)desc.iisph_viscosity, density_old);
            }
            vstar[i] = vel[i] + dt * npa;
            ext[i] = vzero<T, D>();
        });
    }
    void iisph2d_phase2(T dt) {
        const T rho0 = (T)desc.rest_density;
        const T gamma = (T)desc.iisph_gamma;
        loop_for(n, [&](size_t i) {
            if (type[i] != T_NORMAL) return;
            const V& xi = pos[i];
            T sum = T(0);  // ComputeSourceTerm
            for_neighbors(i, [&](size_t j) {
                if (type[j] == T_INACTIVE) return;
                sum = sum + mass[j] * dot(g_mutation == 4 && type[j] == T_BOUNDARY ? vstar[i] : vstar[i] - vstar[j],
                                          kernel.reversed(pos[j], xi));
            });
            source[i] = rho0 - density[i] - dt * sum;
            // ComputeDiagonalElement: the inner sum does not depend on the outer
            // neighbor, so it is formed once; its value is identical.
            V internal = vzero<T, D>();
            for_neighbors(i, [&](size_t k) {
                uint8_t tk = type[k];
                if (tk == T_NORMAL)
                    internal = internal + (mass[k] / rho0 / rho0) * kernel.reversed(pos[k], xi);
                else if (tk == T_BOUNDARY)
                    internal = internal + ((g_mutation == 1 ? T(1) : T(2)) * gamma * mass[k] / rho0 / rho0) *
                                              kernel.reversed(pos[k], xi);
            });
            T s = T(0);
            for_neighbors(i, [&](size_t j) {
                if (type[j] == T_INACTIVE) return;
                s = s + mass[j] * dot(-internal, kernel.reversed(pos[j], xi));
            });
            for_neighbors(i, [&](size_t j) {
                if (type[j] != T_NORMAL) return;
                s = s + mass[j] * dot((mass[i] / rho0 / rho0) * (g_mutation == 2 ? kernel.reversed(pos[j], xi)
                                                                                 : kernel.reversed(xi, pos[j])),
                                      kernel.reversed(pos[j], xi));
            });
            diag[i] = dt * dt * s;
            pressure[i] = T(0);  // InitializePressure
        });
    }
    // ComputePressure's first loop (:408-440): pressure accelerations
    void iisph2d_pass_a() {
        const T rho0 = (T)desc.rest_density;
        const T gamma = (T)desc.iisph_gamma;
        loop_for(n, [&](size_t i) {
            if (type[i] != T_NORMAL) return;
            const V& xi = pos[i];
            T pi = pressure[i];
            V sum = vzero<T, D>();
            for_neighbors(i, [&](size_t j) {
                uint8_t tj = type[j];
                if (tj == T_NORMAL)
                    sum = sum + (mass[j] * (pi / rho0 / rho0 + pressure[g_mutation == 6 ? i : j] / rho0 / rho0)) *
                                    kernel.reversed(pos[j], xi);
                else if (tj == T_BOUNDARY)
                    sum = sum + (gamma * mass[j] * (g_mutation == 3 ? T(1) : T(2)) * (pi / rho0 / rho0)) *
                                    kernel.reversed(pos[j], xi);
            });
            acc[i] = -sum;
        });
    }
    // ComputePressure's second loop, the product (:443-466): dt^2 sum_j m_j (a_i - a_j) . gradW
    T iisph2d_ap(size_t i, T dt) const {
        const V& xi = pos[i];
        const V ai = acc[i];
        T sum = T(0);
        for_neighbors(i, [&](size_t j) {
            uint8_t tj = type[j];
            if (tj == T_NORMAL)
                sum = sum + mass[j] * dot(ai - acc[j], kernel.reversed(pos[j], xi));
            else if (tj == T_BOUNDARY)
                sum = sum + mass[j] * dot(ai, kernel.reversed(pos[j], xi));
        });
        return dt * dt * sum;
    }

    // ComputePressure (:396-485): returns the iteration count, err the last average error
    size_t iisph2d_solve(T dt, T& err) {
        const T eps = Consts<T>::eps();
        size_t iteration = 0;
        err = T(0);
        const bool int_abs = (desc.flags & FSA_FLAG_EMULATE_INT_ABS) != 0;
        auto err_abs = [&](T e) -> T {
            if (int_abs) return (T)x86_int_abs((double)e);
            return std::abs(e);
        };
        while ((iteration < desc.iisph_min_iterations || err_abs(err) > (T)desc.iisph_max_density_error_allowed) &&
               iteration < desc.iisph_max_iterations) {
            err = T(0);
            size_t counter = 0;
            iisph2d_pass_a();
            std::vector<T> e_i(n, T(0));
            std::vector<uint8_t> counted(n, 0);
            loop_for(n, [&](size_t i) {
                if (type[i] != T_NORMAL) return;
                T Ap = iisph2d_ap(i, dt);
                T dgi = diag[i], si = source[i];
                if (std::abs(dgi) > eps)
                    pressure[i] = std::fmax(T(0), pressure[i] + (g_mutation == 5 ? T(-1) : T(1)) * (T)desc.iisph_omega *
                                                                   ((si - Ap) / dgi));
                else
                    pressure[i] = T(0);
                T e = std::abs(si - Ap);
                if (std::abs(dgi) > eps) {
                    if (std::abs(pressure[i]) <= eps) e = T(0);
                    e_i[i] = e;
                    counted[i] = 1;
                }
            });
            double err64 = 0.0;
            for (size_t i = 0; i < n; i++)
                if (counted[i]) {
                    err = err + e_i[i];
                    err64 = err64 + (double)e_i[i];
                    counter++;
                }
            if (counter != 0) {
                err = err / (T)counter;
                err64 = err64 / (double)counter;
            } else {
                err = T(0);
                err64 = 0.0;
            }
            if (fp64_reduction) err = (T)err64;
            iteration++;
        }
        return iteration;
    }

    int simulate_iisph2d(float desired, fsa_step_stats* st) {
        T dt = (T)desired;
        iisph2d_phase1(dt);
        iisph2d_phase2(dt);
        T err = T(0);
        const size_t iteration = iisph2d_solve(dt, err);
        loop_for(n, [&](size_t i) {
            if (type[i] == T_INACTIVE) return;
            vel[i] = vstar[i] + dt * acc[i];
            pos[i] = pos[i] + dt * vel[i];
        });
        if (st) {
            st->desired_time_step = desired;
            st->actual_time_step = desired;
            st->iterations = (uint32_t)iteration;
            st->status = FSA_OK;
            st->average_density_error = (double)err;
            st->max_velocity = 0.0f;
            st->max_acceleration = 0.0f;
        }
        return FSA_OK;
    }

    // ---- simulation entities around the solver (Simulator.cpp:26-45)
    std::vector<fsa_entity> entities;

    // BoundaryPreprocessor::execute_simulation_step and initialize
    // (BoundaryPreprocessor.cpp:7-62, 96-117): the neighbor lists of the last
    // search, in their order; float sums and Math::min = std::fmin.
    void boundary_preprocess() {
        const float ps = desc.particle_size, rho0 = desc.rest_density;
        float ideal = 0.0f;
        for (int x = 0; x <= 0; x++)
            for (int y = -2; y <= 2; y++)
                for (int z = -2; z <= 2; z++) {
                    V p = vzero<T, D>();
                    p.v[0] = (T)((float)x * ps);
                    p.v[1] = (T)((float)y * ps);
                    if (D == 3) p.v[2] = (T)((float)z * ps);
                    ideal += (float)kernel.value(p);
                }
        const float gamma_1 = ((ps * ps) * ps) * ideal;
        const float max_mass = rho0 * ((ps * ps) * ps);
        for (size_t i = 0; i < n; i++) {
            if (type[i] != T_BOUNDARY) continue;
            float vr = 0.0f;
            for_neighbors(i, [&](size_t j) {
                if (type[j] != T_BOUNDARY) return;
                vr += (float)kernel.value(pos[j], pos[i]);
            });
            const float corrected = rho0 * gamma_1 / vr;
            mass[i] = (T)std::fmin(max_mass, corrected);
        }
    }

    // VelocityAlterationByTag::execute_simulation_step (VelocityAlterationByTag.cpp:6-23)
    void velocity_alteration(const fsa_entity& e) {
        for (size_t i = 0; i < n; i++) {
            if (tag[i] != e.tag) continue;
            for (int k = 0; k < D; k++) vel[i].v[k] = (T)e.velocity[k];
        }
    }

    // ParticleRemover3D::execute_simulation_step (ParticleRemover3D.cpp:8-16) with
    // VolumeGroup::is_inside / is_member (VolumeGroup.cpp:43-58) and GroupActions::remove
    void particle_remover(const fsa_entity& e) {
        float lo[3], hi[3];
        for (int k = 0; k < 3; k++) {
            hi[k] = e.center[k] + e.distance_from_center[k];
            lo[k] = e.center[k] - e.distance_from_center[k];
        }
        const bool inside_group = !e.remove_if_outside;
        for (size_t i = 0; i < n; i++) {
            const float x = (float)pos[i].v[0], y = (float)pos[i].v[1], z = D == 3 ? (float)pos[i].v[2] : 0.0f;
            const bool inside = !(x < lo[0] || x > hi[0] || y > hi[1] || y < lo[1] || z < lo[2] || z > hi[2]);
            if (inside == inside_group) type[i] = T_INACTIVE;
        }
    }

    void run_entities(bool before) {
        for (const fsa_entity& e : entities) {
            if (e.type == FSA_ENTITY_BOUNDARY_PREPROCESSOR && before) boundary_preprocess();
            else if (e.type == FSA_ENTITY_VELOCITY_ALTERATION_BY_TAG && before) velocity_alteration(e);
            else if (e.type == FSA_ENTITY_PARTICLE_REMOVER_3D && !before) particle_remover(e);
        }
    }

    // the global sensors (ParticleStatistics.cpp:17-207) over the collection in
    // its order, as the reference computes them: float sums, std::min/std::max
    // from FLT_MAX / FLT_MIN, the averages divided in float.
    void statistics(float zero_height, fsa_scene_statistics* out) const {
        std::memset(out, 0, sizeof(*out));
        float dmax = std::numeric_limits<float>::min(), dmin = std::numeric_limits<float>::max(), dsum = 0.0f;
        float pmax = dmax, pmin = dmin, psum = 0.0f, vmax = dmax, vmin = dmin, vsum = 0.0f;
        float ke = 0.0f, pe = 0.0f;
        size_t cnt = 0;
        for (size_t i = 0; i < n; i++) {
            if (type[i] == T_NORMAL) out->normal_particles++;
            else if (type[i] == T_BOUNDARY) out->boundary_particles++;
            else out->inactive_particles++;
            if (type[i] != T_NORMAL) continue;
            const float rho = (float)density[i], p = (float)pressure[i], m = (float)mass[i];
            float v2 = 0.0f;
            for (int k = 0; k < D; k++) v2 = v2 + (float)vel[i].v[k] * (float)vel[i].v[k];
            const float v = std::sqrt(v2);
            dsum += rho;
            dmin = std::min(dmin, rho);
            dmax = std::max(dmax, rho);
            psum += p;
            pmin = std::min(pmin, p);
            pmax = std::max(pmax, p);
            vsum += v;
            vmin = std::min(vmin, v);
            vmax = std::max(vmax, v);
            ke += 0.5f * m * v * v;
            pe += desc.gravity * m * ((float)pos[i].v[1] - zero_height);
            cnt++;
        }
        if (cnt) {
            dsum = dsum / cnt;
            psum = psum / cnt;
            vsum = vsum / cnt;
        }
        out->density_max = dmax;
        out->density_min = dmin;
        out->density_avg = dsum;
        out->pressure_max = pmax;
        out->pressure_min = pmin;
        out->pressure_avg = psum;
        out->velocity_max = vmax;
        out->velocity_min = vmin;
        out->velocity_avg = vsum;
        out->kinetic_energy = ke;
        out->potential_energy = pe;
    }

    int simulate(float dt, fsa_step_stats* st) override {
        if (!searched) {
            g_err = "simulate called before find_neighbors";
            return FSA_E_STATE;
        }
        run_entities(true);
        int rc;
        if (desc.solver == FSA_SOLVER_SESPH) rc = simulate_sesph(dt, st);
        else if (D == 3) rc = simulate_iisph3d(dt, st);
        else rc = simulate_iisph2d(dt, st);
        if (rc == FSA_OK) run_entities(false);
        return rc;
    }

    bool iisph_ok() const {
        if (desc.solver != FSA_SOLVER_IISPH) {
            g_err = "not an IISPH scene";
            return false;
        }
        if (!searched) {
            g_err = "find_neighbors first";
            return false;
        }
        return true;
    }
    // phases 1 and 2 at dt; with solve, the Jacobi loop too (pressures stay with the
    // particles; no CFL correction, no integration); returns the iteration count
    int iisph_prepare(float dt, int solve) override {
        if (!iisph_ok()) return -FSA_E_STATE;
        size_t it = 0;
        if (D == 3) {
            iisph3d_phase1((T)dt);
            iisph3d_phase2((T)dt);
            T v2 = 0, a2 = 0;
            double avg = 0.0;
            if (solve) it = iisph3d_solve((T)dt, v2, a2, avg);
        } else {
            iisph2d_phase1((T)dt);
            iisph2d_phase2((T)dt);
            T err = T(0);
            if (solve) it = iisph2d_solve((T)dt, err);
        }
        return (int)it;
    }
    // pressure p of caller row id[k] -> A p at that row (0 for non-fluid rows);
    // acc_out (D per row, may be NULL) receives pass A's accelerations
    int iisph_apply(float dt, const double* p, double* ap, double* acc_out) override {
        if (!iisph_ok()) return FSA_E_STATE;
        for (size_t k = 0; k < n; k++) pressure[k] = (T)p[id[k]];
        if (D == 3)
            iisph3d_pass_a();
        else
            iisph2d_pass_a();
        for (size_t k = 0; k < n; k++) {
            ap[id[k]] = type[k] == T_NORMAL ? (double)(D == 3 ? iisph3d_ap(k, (T)dt) : iisph2d_ap(k, (T)dt)) : 0.0;
            if (acc_out)
                for (int c = 0; c < D; c++) acc_out[id[k] * D + c] = type[k] == T_NORMAL ? (double)acc[k].v[c] : 0.0;
        }
        return FSA_OK;
    }
    // per caller row: density, source term, diagonal element, pressure, v* (D), a_np (D; 0 in 2D)
    void iisph_fields(double* out) const override {
        const size_t w = 4 + 2 * D;
        for (size_t k = 0; k < n; k++) {
            double* r = out + id[k] * w;
            r[0] = (double)density[k];
            r[1] = (double)source[k];
            r[2] = (double)diag[k];
            r[3] = (double)pressure[k];
            for (int c = 0; c < D; c++) {
                r[4 + c] = (double)vstar[k].v[c];
                r[4 + D + c] = D == 3 ? (double)anp[k].v[c] : 0.0;
            }
        }
    }

    int download(const fsa_particle_view* v) override {
        if (v->count != n) {
            g_err = "download view count mismatch";
            return FSA_E_INVALID;
        }
        for (size_t k = 0; k < n; k++) {
            size_t i = id[k];
            if (v->movement) {
                float* mv = (float*)((char*)v->movement + i * v->movement_stride);
                for (int c = 0; c < D; c++) {
                    mv[c] = (float)pos[k].v[c];
                    mv[D + c] = (float)vel[k].v[c];
                    mv[2 * D + c] = (float)acc[k].v[c];
                }
            }
            if (v->particle_data) {
                float* pd = (float*)((char*)v->particle_data + i * v->particle_data_stride);
                pd[0] = (float)mass[k];
                pd[1] = (float)pressure[k];
                pd[2] = (float)density[k];
            }
            if (v->info) {
                char* inf = (char*)v->info + i * v->info_stride;
                std::memcpy(inf, &tag[k], 4);
                inf[4] = (char)type[k];
            }
            if (v->external) {
                float* e = (float*)((char*)v->external + i * v->external_stride);
                for (int c = 0; c < D; c++) e[c] = (float)ext[k].v[c];
            }
        }
        return FSA_OK;
    }
    int download_iisph(void* rec, uint64_t stride) override {
        for (size_t k = 0; k < n; k++) {
            float* r = (float*)((char*)rec + id[k] * stride);
            int o = 0;
            for (int c = 0; c < D; c++) r[o++] = (float)vstar[k].v[c];
            if (D == 3)
                for (int c = 0; c < D; c++) r[o++] = (float)anp[k].v[c];
            r[o++] = (float)source[k];
            r[o++] = (float)diag[k];
            if (D == 3) r[o++] = (float)perr[k];
        }
        return FSA_OK;
    }
    int export_neighbors(uint64_t* off, uint64_t* cols, uint64_t cap, uint64_t* nnz) override {
        // rows in caller order, columns in caller indices, ascending
        std::vector<std::vector<uint64_t>> rows(n);
        for (size_t k = 0; k < n; k++) {
            bool has = false;
            if (search == SEARCH_COMPRESSED3D)
                has = type[k] != T_INACTIVE;
            else
                has = k < lists.size();
            if (!has) continue;
            std::vector<uint64_t>& r = rows[id[k]];
            for_neighbors(k, [&](size_t j) { r.push_back(id[j]); });
            std::sort(r.begin(), r.end());
        }
        uint64_t total = 0;
        for (size_t i = 0; i < n; i++) {
            off[i] = total;
            for (uint64_t c : rows[i]) {
                if (total < cap) cols[total] = c;
                total++;
            }
        }
        off[n] = total;
        *nnz = total;
        if (total > cap) {
            g_err = "export capacity too small";
            return FSA_E_INVALID;
        }
        return FSA_OK;
    }
    // NeighborStorage size and encoded bytes of each particle's list, by caller row
    // (compressed search only; other searches report sizes and zero bytes)
    void codec_sizes(uint64_t* count_per_row, uint64_t* bytes_per_row) override {
        for (size_t k = 0; k < n; k++) {
            const size_t row = id[k];
            if (search == SEARCH_COMPRESSED3D) {
                count_per_row[row] = storage[k].size();
                bytes_per_row[row] = storage[k].encoded_bytes();
            } else {
                count_per_row[row] = k < lists.size() ? lists[k].size() : 0;
                bytes_per_row[row] = 0;
            }
        }
    }
    // the caller rewrites particle data in place: the collection keeps the order
    // its searches gave it (ParticleCollection rows are edited where they are)
    int upload(const fsa_particle_view* v) override {
        if (v->count != n) {
            g_err = "upload view count mismatch";
            return FSA_E_INVALID;
        }
        for (size_t k = 0; k < n; k++) {
            const size_t i = id[k];
            const float* mv = (const float*)((const char*)v->movement + i * v->movement_stride);
            for (int c = 0; c < D; c++) {
                pos[k].v[c] = (T)mv[c];
                vel[k].v[c] = (T)mv[D + c];
                acc[k].v[c] = (T)mv[2 * D + c];
            }
            const float* pd = (const float*)((const char*)v->particle_data + i * v->particle_data_stride);
            mass[k] = pd[0];
            pressure[k] = pd[1];
            density[k] = pd[2];
            const char* inf = (const char*)v->info + i * v->info_stride;
            std::memcpy(&tag[k], inf, 4);
            type[k] = (uint8_t)inf[4];
            if (v->external) {
                const float* e = (const float*)((const char*)v->external + i * v->external_stride);
                for (int c = 0; c < D; c++) ext[k].v[c] = (T)e[c];
            }
        }
        searched = false;
        return FSA_OK;
    }
    // lists as the search stores them: rows and columns in the collection's
    // current order, entries in list order
    int export_raw(uint64_t* off, uint64_t* cols, uint64_t cap, uint64_t* nnz) override {
        uint64_t total = 0;
        for (size_t k = 0; k < n; k++) {
            off[k] = total;
            const bool has = search == SEARCH_COMPRESSED3D ? type[k] != T_INACTIVE : k < lists.size();
            if (!has) continue;
            for_neighbors(k, [&](size_t j) {
                if (total < cap) cols[total] = j;
                total++;
            });
        }
        off[n] = total;
        *nnz = total;
        return total > cap ? FSA_E_INVALID : FSA_OK;
    }
    // caller row of each position of the collection's current order
    void internal_ids(uint64_t* out) override {
        for (size_t k = 0; k < n; k++) out[k] = id[k];
    }
    void codec_stats(uint64_t* total, uint64_t* bytes) override {
        uint64_t t = 0, b = 0;
        if (search == SEARCH_COMPRESSED3D) {
            for (size_t k = 0; k < n; k++) {
                t += storage[k].size();
                b += storage[k].encoded_bytes();
            }
        } else {
            for (auto& l : lists) t += l.size();
        }
        *total = t;
        *bytes = b;
    }
};

}  // namespace orc

using namespace orc;

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }
void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void) { return thread_count(); }

// mode: 0 = FP32 reference order, 1 = FP64 throughout, 2 = FP32 reference
// order with FP64 convergence sums (the device library's reduction)
void* orc_create(const fsa_scene_desc* d, int search_kind, int mode, const fsa_particle_view* v) {
    try {
        SceneBase* s = nullptr;
        bool r64 = mode != 0;
        if (d->dimension == 3) {
            if (mode == 1) {
                auto* p = new Scene<double, 3>();
                p->load(*d, search_kind, v);
                p->fp64_reduction = r64;
                s = p;
            } else {
                auto* p = new Scene<float, 3>();
                p->load(*d, search_kind, v);
                p->fp64_reduction = r64;
                s = p;
            }
        } else {
            if (mode == 1) {
                auto* p = new Scene<double, 2>();
                p->load(*d, search_kind, v);
                p->fp64_reduction = r64;
                s = p;
            } else {
                auto* p = new Scene<float, 2>();
                p->load(*d, search_kind, v);
                p->fp64_reduction = r64;
                s = p;
            }
        }
        return s;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void orc_destroy(void* s) { delete (SceneBase*)s; }
int orc_find_neighbors(void* s) { return ((SceneBase*)s)->find_neighbors(); }
int orc_simulate(void* s, float dt, fsa_step_stats* st) { return ((SceneBase*)s)->simulate(dt, st); }
int orc_iisph_prepare(void* s, float dt, int solve) { return ((SceneBase*)s)->iisph_prepare(dt, solve); }
int orc_iisph_apply(void* s, float dt, const double* p, double* ap, double* acc) {
    return ((SceneBase*)s)->iisph_apply(dt, p, ap, acc);
}
void orc_iisph_fields(void* s, double* out) { ((SceneBase*)s)->iisph_fields(out); }
double orc_x86_int_abs(double e) { return x86_int_abs(e); }
void orc_set_mutation(int m) { g_mutation = m; }
float orc_next_timestep(void* s) { return ((SceneBase*)s)->next_timestep(); }
void orc_set_entities(void* s, const fsa_entity* e, uint32_t count) { ((SceneBase*)s)->set_entities(e, count); }
void orc_statistics(void* s, float zero_height, fsa_scene_statistics* out) { ((SceneBase*)s)->get_statistics(zero_height, out); }
int orc_query_neighbors(void* s, const float* x, uint64_t count, uint64_t* off, uint64_t* rows, uint64_t cap,
                        uint64_t* nnz) {
    return ((SceneBase*)s)->query_neighbors(x, count, off, rows, cap, nnz);
}
int orc_sensor_plane(void* s, const fsa_sensor_plane* p, float* values) { return ((SceneBase*)s)->sensor_plane(p, values); }
int orc_evaluate_volume(void* s, const float* x, uint64_t count, fsa_volume_sample* out) {
    return ((SceneBase*)s)->evaluate_volume(x, count, out);
}
int orc_download(void* s, const fsa_particle_view* v) { return ((SceneBase*)s)->download(v); }
int orc_download_iisph(void* s, void* rec, uint64_t stride) { return ((SceneBase*)s)->download_iisph(rec, stride); }
int orc_export_neighbors(void* s, uint64_t* off, uint64_t* cols, uint64_t cap, uint64_t* nnz) {
    return ((SceneBase*)s)->export_neighbors(off, cols, cap, nnz);
}
void orc_codec_stats(void* s, uint64_t* total, uint64_t* bytes) { ((SceneBase*)s)->codec_stats(total, bytes); }

void orc_translate(void* s, const float* delta) { ((SceneBase*)s)->translate(delta); }

void orc_internal_ids(void* s, uint64_t* out) { ((SceneBase*)s)->internal_ids(out); }
int orc_upload(void* s, const fsa_particle_view* v) { return ((SceneBase*)s)->upload(v); }
int orc_export_raw(void* s, uint64_t* off, uint64_t* cols, uint64_t cap, uint64_t* nnz) {
    return ((SceneBase*)s)->export_raw(off, cols, cap, nnz);
}

void orc_codec_sizes(void* s, uint64_t* count_per_row, uint64_t* bytes_per_row) {
    ((SceneBase*)s)->codec_sizes(count_per_row, bytes_per_row);
}

// ---- pure functions, exposed for pinning tests --------------------------
void orc_kernel_values(int dim, uint64_t n, const float* xs, float support, float* out) {
    if (dim == 3) {
        CubicSpline<float, 3> k;
        k.initialize(support);
        for (uint64_t i = 0; i < n; i++) out[i] = k.value(Vec<float, 3>{{xs[3 * i], xs[3 * i + 1], xs[3 * i + 2]}});
    } else {
        CubicSpline<float, 2> k;
        k.initialize(support);
        for (uint64_t i = 0; i < n; i++) out[i] = k.value(Vec<float, 2>{{xs[2 * i], xs[2 * i + 1]}});
    }
}
void orc_kernel_derivatives(int dim, uint64_t n, const float* xs, float support, float* out) {
    if (dim == 3) {
        CubicSpline<float, 3> k;
        k.initialize(support);
        for (uint64_t i = 0; i < n; i++) {
            auto r = k.derivative(Vec<float, 3>{{xs[3 * i], xs[3 * i + 1], xs[3 * i + 2]}});
            for (int c = 0; c < 3; c++) out[3 * i + c] = r.v[c];
        }
    } else {
        CubicSpline<float, 2> k;
        k.initialize(support);
        for (uint64_t i = 0; i < n; i++) {
            auto r = k.derivative(Vec<float, 2>{{xs[2 * i], xs[2 * i + 1]}});
            for (int c = 0; c < 2; c++) out[2 * i + c] = r.v[c];
        }
    }
}
float orc_kernel_value(int dim, const float* x, float support) {
    if (dim == 3) {
        CubicSpline<float, 3> k;
        k.initialize(support);
        return k.value(Vec<float, 3>{{x[0], x[1], x[2]}});
    }
    CubicSpline<float, 2> k;
    k.initialize(support);
    return k.value(Vec<float, 2>{{x[0], x[1]}});
}
void orc_kernel_derivative(int dim, const float* x, float support, float* out) {
    if (dim == 3) {
        CubicSpline<float, 3> k;
        k.initialize(support);
        auto r = k.derivative(Vec<float, 3>{{x[0], x[1], x[2]}});
        for (int c = 0; c < 3; c++) out[c] = r.v[c];
        return;
    }
    CubicSpline<float, 2> k;
    k.initialize(support);
    auto r = k.derivative(Vec<float, 2>{{x[0], x[1]}});
    for (int c = 0; c < 2; c++) out[c] = r.v[c];
}
uint64_t orc_morton3(uint32_t x, uint32_t y, uint32_t z) { return morton3_64(x, y, z); }
uint64_t orc_cell_key(int32_t x, int32_t y, int32_t z) {
    return morton3_64(to_unsigned(x), to_unsigned(y), to_unsigned(z));
}
void orc_quick_sort_u64(const uint64_t* keys, uint64_t n, uint64_t* perm) {
    std::vector<uint64_t> k(keys, keys + n);
    const std::vector<size_t> p = quick_sort_perm(k);
    for (uint64_t i = 0; i < n; i++) perm[i] = p[i];
}

void orc_stable_sort_u64(const uint64_t* keys, uint64_t n, uint64_t* perm) {
    std::vector<uint64_t> p(n);
    std::iota(p.begin(), p.end(), 0);
    std::stable_sort(p.begin(), p.end(), [&](uint64_t a, uint64_t b) { return keys[a] < keys[b]; });
    std::memcpy(perm, p.data(), n * sizeof(uint64_t));
}
float orc_cfl_timestep(float max_a, float max_v, float particle_size, float timestep, float lambda_a, float lambda_v) {
    Scene<float, 3> s;
    std::memset(&s.desc, 0, sizeof(s.desc));
    s.desc.timestep_generator = FSA_TIMESTEP_CONSTANT;
    s.desc.particle_size = particle_size;
    s.desc.timestep = timestep;
    s.desc.lambda_a = lambda_a;
    s.desc.lambda_v = lambda_v;
    return s.cfl_timestep(max_a, max_v);
}
// NeighborStorage codec
void* orc_codec_new(void) { return new NeighborStorage(); }
void orc_codec_free(void* c) { delete (NeighborStorage*)c; }
void orc_codec_clear(void* c) { ((NeighborStorage*)c)->clear(); }
void orc_codec_set_first(void* c, uint64_t i) { ((NeighborStorage*)c)->set_first(i); }
int orc_codec_set_next(void* c, uint64_t d) { return ((NeighborStorage*)c)->set_next(d) ? 0 : 1; }
uint64_t orc_codec_size(void* c) { return ((NeighborStorage*)c)->size(); }
uint32_t orc_codec_get_delta(void* c, uint64_t k) { return ((NeighborStorage*)c)->get_delta(k); }
uint64_t orc_codec_get_first(void* c) { return ((NeighborStorage*)c)->first_neighbor; }
uint64_t orc_codec_used_bytes(void* c) { return ((NeighborStorage*)c)->bytes; }
uint64_t orc_codec_encoded_bytes(void* c) { return ((NeighborStorage*)c)->encoded_bytes(); }

}  // extern "C"
