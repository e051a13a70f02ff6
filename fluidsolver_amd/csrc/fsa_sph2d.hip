// fsa_sph2d.hip — 2D neighborhood searches and the 2D IISPH / SESPH steps.
//
// Reference (src/libFluid/fluidSolver/...):
//   HashedNeighborhoodSearch::UpdateGrid / find_neighbors
//       (neighborhoodSearch/HashedNeighborhoodSearch.cpp:9-69, 117-151)
//   QuadraticNeighborhoodSearchDynamicAllocated::find_neighbors
//       (neighborhoodSearch/QuadraticNeighborhoodSearchDynamicAllocated.cpp:11-38)
//   IISPHFluidSolver (solver/IISPHFluidSolver.hpp:139-525)
//   SESPHFluidSolver (solver/SESPHFluidSolver.hpp:49-299)
//
// 2D scenes keep the caller's particle order (neither 2D search sorts the
// collection).  The hashed grid appends a particle to its new cell whenever
// it changes cell, visiting particles in index order, so the members of a
// cell are ordered by (search count of their last cell change, index); the
// device reproduces that order with two stable radix sorts, by entry count
// and then by cell, and walks the 3x3 stencil x-outer, y-inner.  The
// quadratic search lists every particle in ascending index, which the list
// kernel restores by insertion.
#include "fsa_device.hpp"
#include "fsa_kernels.hpp"

namespace fsa {

namespace {

constexpr float kEps2 = 1.1920928955078125e-07f;  // std::numeric_limits<float>::epsilon()

// GetGridCellByPosition (HashedNeighborhoodSearch.cpp:107-114): float quotient
// truncated toward zero
__device__ __forceinline__ int2 cell2(float x, float y, float r) {
    return make_int2((int32_t)(x / r), (int32_t)(y / r));
}

__device__ __forceinline__ bool cell2_linear(const SceneDev& S, int32_t cx, int32_t cy, uint32_t& lin) {
    const int32_t ox = cx - S.cmin[0], oy = cy - S.cmin[1];
    if (ox < 0 || oy < 0 || (uint32_t)ox >= S.cdim[0] || (uint32_t)oy >= S.cdim[1]) return false;
    lin = (uint32_t)oy * S.cdim[0] + (uint32_t)ox;
    return true;
}

__device__ __forceinline__ uint32_t type_of(const SceneDev& S, uint32_t i) { return S.idt[0][i] >> kTypeShift; }

// ---- search ---------------------------------------------------------------------
// grid update: entry count of every grid member, first sort key = entry
__global__ __launch_bounds__(kBlock) void k2_grid_update(const SceneDev* __restrict__ scenes,
                                                         const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t i = br.block * kBlock + threadIdx.x;
    if (i >= S.n) return;
    const float4 p = S.pos[0][i];
    const int2 c = cell2(p.x, p.y, S.search_radius);
    uint32_t e = 0;
    if (!(S.flags & kFlagQuadratic) && i < S.grid_count) {
        const int2 prev = S.cell_prev[i];
        e = S.entry[i];
        if (prev.x != c.x || prev.y != c.y) {
            e = S.search_count;
            S.entry[i] = e;
            S.cell_prev[i] = c;
        }
    }
    S.keys[0][i] = e;
    S.vals[0][i] = i;
}

// second sort key: cell of the particle at sorted position k (non-members last)
__global__ __launch_bounds__(kBlock) void k2_cell_keys(const SceneDev* __restrict__ scenes,
                                                       const BlockRef* __restrict__ map, uint32_t buf) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t k = br.block * kBlock + threadIdx.x;
    if (k >= S.n) return;
    const uint32_t i = S.vals[buf][k];
    const bool member = (S.flags & kFlagQuadratic) || i < S.grid_count;
    uint32_t key = S.cdim[3];  // one past the last cell
    if (member) {
        const float4 p = S.pos[0][i];
        const int2 c = cell2(p.x, p.y, S.search_radius);
        uint32_t lin;
        if (cell2_linear(S, c.x, c.y, lin))
            key = lin;
        else
            atomicOr(&S.state->status, kStatusGrid);
    }
    S.keys[buf][k] = key;
}

__global__ __launch_bounds__(kBlock) void k2_cell_table(const SceneDev* __restrict__ scenes,
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

// neighbor entry of particle j: index | boundary / inactive flags
__device__ __forceinline__ uint32_t entry_of(const SceneDev& S, uint32_t j) {
    const uint32_t t = type_of(S, j);
    return j | (t == 1u ? kBoundaryBit : 0u) | (t == 2u ? kInactiveBit : 0u);
}

// one thread per fluid particle: 3x3 stencil (x outer, y inner), members of a
// cell in sorted order, accepted when length(x_i - x_j) <= r
__global__ __launch_bounds__(kBlock) void k2_lists(const SceneDev* __restrict__ scenes,
                                                   const BlockRef* __restrict__ map, uint32_t buf) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    const uint32_t k = br.block * kBlock + threadIdx.x;
    if (k >= S.nf) return;
    const uint32_t i = S.fluid_idx[k];
    const float4 xi = S.pos[0][i];
    const int2 c = cell2(xi.x, xi.y, S.search_radius);
    const uint32_t* __restrict__ vals = S.vals[buf];
    uint32_t* __restrict__ out = S.nbr + k;
    const uint32_t stride = S.list_stride;
    const bool quad = (S.flags & kFlagQuadratic) != 0;
    uint32_t cnt = 0;
    for (int dx = -1; dx <= 1; dx++) {
        for (int dy = -1; dy <= 1; dy++) {
            uint32_t lin;
            if (!cell2_linear(S, c.x + dx, c.y + dy, lin)) continue;
            const uint4 e = S.cells[lin];
            if (e.z != S.stamp) continue;
            for (uint32_t m = e.x; m < e.y; m++) {
                const uint32_t j = vals[m];
                const float4 xj = S.pos[0][j];
                const float ddx = xi.x - xj.x, ddy = xi.y - xj.y, ddz = xi.z - xj.z;
                if (!(__builtin_sqrtf(dot3(ddx, ddy, ddz, ddx, ddy, ddz)) <= S.search_radius)) continue;
                const uint32_t ent = entry_of(S, j);
                if (quad) {
                    // ascending index: insert before the first larger entry
                    uint32_t pos = min(cnt, (uint32_t)kMaxNeighbors);
                    while (pos > 0) {
                        const uint32_t prev = out[(pos - 1) * stride];
                        if ((prev & kIndexMask) < j) break;
                        if (pos < (uint32_t)kMaxNeighbors) out[pos * stride] = prev;
                        pos--;
                    }
                    if (pos < (uint32_t)kMaxNeighbors) out[pos * stride] = ent;
                } else if (cnt < (uint32_t)kMaxNeighbors) {
                    out[cnt * stride] = ent;
                }
                cnt++;
            }
        }
    }
    if (cnt > (uint32_t)kMaxNeighbors) atomicOr(&S.state->status, kStatusCapacity);
    S.nbr_cnt[k] = min(cnt, (uint32_t)kMaxNeighbors);
}

// ---- IISPHFluidSolver (2D) --------------------------------------------------------
// CubicSplineKernel (2D) through the 3-component helpers with z = 0.
struct G2 {
    float x, y, z;
};

__device__ __forceinline__ void geom2(const SceneDev& S, const float4& xi, const float4& xj, float& dx, float& dy,
                                     float& dz, float& d2, float& len, G2& g) {
    dx = xi.x - xj.x;
    dy = xi.y - xj.y;
    dz = xi.z - xj.z;
    d2 = dot3(dx, dy, dz, dx, dy, dz);
    len = __builtin_sqrtf(d2);
    kernel_grad<0>(S, dx, dy, dz, len, g.x, g.y, g.z);
}

// CalculateDensity + CalculateNonPressureAccelerationAndPredictedVelocity
// (IISPHFluidSolver.hpp:147-180, 196-266) for a fluid particle; viscosity
// reads the start-of-step densities.
__global__ __launch_bounds__(kBlock) void k2_iisph_prepare(const SceneDev* __restrict__ scenes,
                                                           const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    const uint32_t k = br.block * kBlock + threadIdx.x;
    if (k >= S.nf) return;
    const uint32_t i = S.fluid_idx[k];
    const float4* __restrict__ pos = S.pos[0];
    const float4* __restrict__ vel = S.vel[0];
    const float4 xi = pos[i], vi = vel[i];
    const float4 e = S.ext[0][i];
    const uint32_t cnt = S.nbr_cnt[k];
    float rho = 0.0f;
    for (uint32_t q = 0; q < cnt; q++) {
        const uint32_t ent = S.nbr[q * S.list_stride + k];
        if (ent & kInactiveBit) continue;
        const float4 xj = pos[ent & kIndexMask];
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        rho = rho + xj.w * kernel_w<0>(S, __builtin_sqrtf(dot3(dx, dy, dz, dx, dy, dz)));
    }
    float tx = 0.0f, ty = 0.0f, tz = 0.0f;
    for (uint32_t q = 0; q < cnt; q++) {
        const uint32_t ent = S.nbr[q * S.list_stride + k];
        if (ent & kInactiveBit) continue;
        const uint32_t j = ent & kIndexMask;
        const float4 xj = pos[j], vj = vel[j];
        const float nd = vj.w;
        if (nd == 0.0f) continue;
        float dx, dy, dz, d2, len;
        G2 g;
        geom2(S, xi, xj, dx, dy, dz, d2, len, g);
        const float vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
        const float t = (xj.w / nd) * (dot3(vx, vy, vz, dx, dy, dz) / (d2 + S.visc_c01));
        tx = tx + t * g.x;
        ty = ty + t * g.y;
        tz = tz + t * g.z;
    }
    float nx = e.x + 0.0f, ny = e.y + (-S.gravity), nz = e.z + 0.0f;
    nx = nx + S.visc2 * tx;
    ny = ny + S.visc2 * ty;
    nz = nz + S.visc2 * tz;
    const float dt = S.dt;
    S.vstar[i] = make_float4(vi.x + dt * nx, vi.y + dt * ny, vi.z + dt * nz, 0.0f);
    S.rho_new[i] = rho;
    S.ext[0][i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// boundary particles: predicted velocity from their external acceleration alone
__global__ __launch_bounds__(kBlock) void k2_prepare_boundary(const SceneDev* __restrict__ scenes,
                                                              const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    const uint32_t i = br.block * kBlock + threadIdx.x;
    if (i >= S.n || type_of(S, i) != 1u) return;
    const float4 e = S.ext[0][i];
    const float4 v = S.vel[0][i];
    const float dt = S.dt;
    S.vstar[i] = make_float4(v.x + dt * e.x, v.y + dt * e.y, v.z + dt * e.z, 0.0f);
    S.ext[0][i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ComputeSourceTerm + ComputeDiagonalElement + InitializePressure
// (IISPHFluidSolver.hpp:274-380).  The diagonal's inner sum does not depend on
// the outer neighbor and is formed once.
__global__ __launch_bounds__(kBlock) void k2_iisph_diagonal(const SceneDev* __restrict__ scenes,
                                                            const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    const uint32_t k = br.block * kBlock + threadIdx.x;
    if (k >= S.nf) return;
    const uint32_t i = S.fluid_idx[k];
    const float4* __restrict__ pos = S.pos[0];
    const float4 xi = pos[i];
    const float4 vs = S.vstar[i];
    const uint32_t cnt = S.nbr_cnt[k];
    const float rho0 = S.rest_density;
    float sum = 0.0f;
    float ix = 0.0f, iy = 0.0f, iz = 0.0f;
    for (uint32_t q = 0; q < cnt; q++) {
        const uint32_t ent = S.nbr[q * S.list_stride + k];
        if (ent & kInactiveBit) continue;
        const uint32_t j = ent & kIndexMask;
        const float4 xj = pos[j];
        const float4 vsj = S.vstar[j];
        float dx, dy, dz, d2, len;
        G2 g;
        geom2(S, xi, xj, dx, dy, dz, d2, len, g);
        sum = sum + xj.w * dot3(vs.x - vsj.x, vs.y - vsj.y, vs.z - vsj.z, g.x, g.y, g.z);
        const float f = (ent & kBoundaryBit) ? 2.0f * S.gamma * xj.w / rho0 / rho0 : xj.w / rho0 / rho0;
        ix = ix + f * g.x;
        iy = iy + f * g.y;
        iz = iz + f * g.z;
    }
    const float dt = S.dt;
    S.src[i] = rho0 - S.rho_new[i] - dt * sum;
    const float cm = xi.w / rho0 / rho0;
    float s = 0.0f;
    for (uint32_t q = 0; q < cnt; q++) {
        const uint32_t ent = S.nbr[q * S.list_stride + k];
        if (ent & kInactiveBit) continue;
        const float4 xj = pos[ent & kIndexMask];
        float dx, dy, dz, d2, len;
        G2 g;
        geom2(S, xi, xj, dx, dy, dz, d2, len, g);
        s = s + xj.w * dot3(-ix, -iy, -iz, g.x, g.y, g.z);
    }
    for (uint32_t q = 0; q < cnt; q++) {
        const uint32_t ent = S.nbr[q * S.list_stride + k];
        if (ent & (kInactiveBit | kBoundaryBit)) continue;
        const float4 xj = pos[ent & kIndexMask];
        float dx, dy, dz, d2, len;
        G2 g;
        geom2(S, xi, xj, dx, dy, dz, d2, len, g);
        // (m_i / rho0 / rho0) * gradW(x_i, x_j), gradW(x_i, x_j) = -gradW(x_j, x_i) exactly
        s = s + xj.w * dot3(cm * -g.x, cm * -g.y, cm * -g.z, g.x, g.y, g.z);
    }
    S.diag[i] = dt * dt * s;
    S.pres[0][i] = 0.0f;
}

// loop entry: while ((it < min || |err| > max_err) && it < max) with err = 0
__global__ void k2_iisph_begin_loop(const SceneDev* __restrict__ scenes, const BlockRef* __restrict__ map) {
    const SceneDev& S = scenes[map[blockIdx.x].scene];
    if (scene_failed(S)) return;
    if (threadIdx.x == 0) {
        SceneState* st = S.state;
        st->iterations = 0;
        st->avg_err = 0.0;
        st->err_count = 0;
        st->max_a2 = 0.0f;
        st->max_v2 = 0.0f;
        st->active = ((0u < S.min_it) || (0.0f > S.max_err)) && (0u < S.max_it) ? 1u : 0u;
    }
}

// pressure acceleration (IISPHFluidSolver.hpp:420-445)
__global__ __launch_bounds__(kBlock) void k2_iisph_pass_a(const SceneDev* __restrict__ scenes,
                                                          const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    if (!S.state->active) return;
    const uint32_t k = br.block * kBlock + threadIdx.x;
    if (k >= S.nf) return;
    const uint32_t i = S.fluid_idx[k];
    const float4* __restrict__ pos = S.pos[0];
    const float* __restrict__ pres = S.pres[0];
    const float4 xi = pos[i];
    const float rho0 = S.rest_density;
    const float pr = pres[i] / rho0 / rho0;
    const uint32_t cnt = S.nbr_cnt[k];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (uint32_t q = 0; q < cnt; q++) {
        const uint32_t ent = S.nbr[q * S.list_stride + k];
        if (ent & kInactiveBit) continue;
        const uint32_t j = ent & kIndexMask;
        const float4 xj = pos[j];
        float dx, dy, dz, d2, len;
        G2 g;
        geom2(S, xi, xj, dx, dy, dz, d2, len, g);
        const float f = (ent & kBoundaryBit) ? S.gamma * xj.w * 2.0f * pr : xj.w * (pr + pres[j] / rho0 / rho0);
        sx = sx + f * g.x;
        sy = sy + f * g.y;
        sz = sz + f * g.z;
    }
    S.acc[0][i] = make_float4(-sx, -sy, -sz, 0.0f);
}

// A*p, relaxed pressure update, error (IISPHFluidSolver.hpp:447-480)
__global__ __launch_bounds__(kBlock) void k2_iisph_pass_b(const SceneDev* __restrict__ scenes,
                                                          const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    if (!S.state->active) return;
    __shared__ double werr[kBlock / 64];
    __shared__ uint32_t wcnt[kBlock / 64];
    const uint32_t k = br.block * kBlock + threadIdx.x;
    double err = 0.0;
    uint32_t counted = 0;
    if (k < S.nf) {
        const uint32_t i = S.fluid_idx[k];
        const float4* __restrict__ pos = S.pos[0];
        const float4* __restrict__ acc = S.acc[0];
        const float4 xi = pos[i];
        const float4 ai = acc[i];
        const uint32_t cnt = S.nbr_cnt[k];
        float sum = 0.0f;
        for (uint32_t q = 0; q < cnt; q++) {
            const uint32_t ent = S.nbr[q * S.list_stride + k];
            if (ent & kInactiveBit) continue;
            const uint32_t j = ent & kIndexMask;
            const float4 xj = pos[j];
            float dx, dy, dz, d2, len;
            G2 g;
            geom2(S, xi, xj, dx, dy, dz, d2, len, g);
            if (ent & kBoundaryBit) {
                sum = sum + xj.w * dot3(ai.x, ai.y, ai.z, g.x, g.y, g.z);
            } else {
                const float4 aj = acc[j];
                sum = sum + xj.w * dot3(ai.x - aj.x, ai.y - aj.y, ai.z - aj.z, g.x, g.y, g.z);
            }
        }
        const float dt = S.dt;
        const float ap = dt * dt * sum;
        const float dg = S.diag[i], si = S.src[i];
        float p = S.pres[0][i];
        if (fabsf(dg) > kEps2)
            p = fmaxf(0.0f, p + S.omega * ((si - ap) / dg));
        else
            p = 0.0f;
        S.pres[0][i] = p;
        float e = fabsf(si - ap);
        if (fabsf(dg) > kEps2) {
            if (fabsf(p) <= kEps2) e = 0.0f;
            err = (double)e;
            counted = 1;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        err += __shfl_xor(err, o, 64);
        counted += __shfl_xor(counted, o, 64);
    }
    if (__lane_id() == 0) {
        werr[threadIdx.x / 64] = err;
        wcnt[threadIdx.x / 64] = counted;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Partial p{};
        for (int w = 0; w < kBlock / 64; w++) {
            p.err += werr[w];
            p.count += wcnt[w];
        }
        S.partial[br.block] = p;
    }
}

// average error of the iteration and the loop condition of the next one
// (IISPHFluidSolver.hpp:396-402, 482-485); the FP64 average is rounded to the
// solver's float err
__global__ __launch_bounds__(kBlock) void k2_iisph_converge(const SceneDev* __restrict__ scenes,
                                                            const BlockRef* __restrict__ map) {
    const SceneDev& S = scenes[map[blockIdx.x].scene];
    if (scene_failed(S)) return;
    if (!S.state->active) return;
    __shared__ double werr[kBlock / 64];
    __shared__ uint32_t wcnt[kBlock / 64];
    const uint32_t nparts = (S.nf + kBlock - 1) / kBlock;
    double err = 0.0;
    uint32_t cnt = 0;
    for (uint32_t b = threadIdx.x; b < nparts; b += kBlock) {
        err += S.partial[b].err;
        cnt += S.partial[b].count;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        err += __shfl_xor(err, o, 64);
        cnt += __shfl_xor(cnt, o, 64);
    }
    if (__lane_id() == 0) {
        werr[threadIdx.x / 64] = err;
        wcnt[threadIdx.x / 64] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0;
        uint32_t c = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            tot += werr[w];
            c += wcnt[w];
        }
        SceneState* st = S.state;
        const uint32_t it = st->iterations + 1;
        st->iterations = it;
        const float e = c != 0 ? (float)(tot / (double)c) : 0.0f;
        st->avg_err = (double)e;
        st->err_count = c;
        const float ea = (S.flags & kFlagIntAbs) ? x86_int_abs(e) : fabsf(e);
        st->active = ((it < S.min_it) || (ea > S.max_err)) && (it < S.max_it) ? 1u : 0u;
    }
}

// ---- SESPHFluidSolver (2D) ---------------------------------------------------------------
// ComputeDensity + ComputePressure (SESPHFluidSolver.hpp:64-75, 190-228)
__global__ __launch_bounds__(kBlock) void k2_sesph_density(const SceneDev* __restrict__ scenes,
                                                           const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    const uint32_t k = br.block * kBlock + threadIdx.x;
    if (k >= S.nf) return;
    const uint32_t i = S.fluid_idx[k];
    const float4* __restrict__ pos = S.pos[0];
    const float4 xi = pos[i];
    const uint32_t cnt = S.nbr_cnt[k];
    float rho = 0.0f;
    for (uint32_t q = 0; q < cnt; q++) {
        const uint32_t ent = S.nbr[q * S.list_stride + k];
        if (ent & kInactiveBit) continue;
        const float4 xj = pos[ent & kIndexMask];
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        rho = rho + xj.w * kernel_w<0>(S, __builtin_sqrtf(dot3(dx, dy, dz, dx, dy, dz)));
    }
    float p = S.stiffness * (rho / S.rest_density - 1.0f);
    p = p < 0.0f ? 0.0f : p;
    S.rho_new[i] = rho;
    S.pres[0][i] = p;
}

// the densities of the step become current before the accelerations read them
__global__ __launch_bounds__(kBlock) void k2_sesph_commit(const SceneDev* __restrict__ scenes,
                                                          const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    const uint32_t k = br.block * kBlock + threadIdx.x;
    if (k >= S.nf) return;
    const uint32_t i = S.fluid_idx[k];
    S.vel[0][i].w = S.rho_new[i];
}

// pressure + viscosity + gravity (SESPHFluidSolver.hpp:78-91, 231-299)
__global__ __launch_bounds__(kBlock) void k2_sesph_accel(const SceneDev* __restrict__ scenes,
                                                         const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    const uint32_t k = br.block * kBlock + threadIdx.x;
    if (k >= S.nf) return;
    const uint32_t i = S.fluid_idx[k];
    const float4* __restrict__ pos = S.pos[0];
    const float4* __restrict__ vel = S.vel[0];
    const float* __restrict__ pres = S.pres[0];
    const float4 xi = pos[i], vi = vel[i];
    const float rho_i = vi.w, p_i = pres[i];
    const float pdd = rho_i == 0.0f ? 0.0f : p_i / (rho_i * rho_i);
    const uint32_t cnt = S.nbr_cnt[k];
    float tx = 0.0f, ty = 0.0f, tz = 0.0f;
    for (uint32_t q = 0; q < cnt; q++) {
        const uint32_t ent = S.nbr[q * S.list_stride + k];
        if (ent & kInactiveBit) continue;
        const uint32_t j = ent & kIndexMask;
        const float4 xj = pos[j], vj = vel[j];
        const float nd = vj.w;
        if (nd == 0.0f) continue;
        float dx, dy, dz, d2, len;
        G2 g;
        geom2(S, xi, xj, dx, dy, dz, d2, len, g);
        const float vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
        const float t = (xj.w / nd) * (dot3(vx, vy, vz, dx, dy, dz) / (d2 + S.visc_c01));
        tx = tx + t * g.x;
        ty = ty + t * g.y;
        tz = tz + t * g.z;
    }
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    for (uint32_t q = 0; q < cnt; q++) {
        const uint32_t ent = S.nbr[q * S.list_stride + k];
        if (ent & kInactiveBit) continue;
        const uint32_t j = ent & kIndexMask;
        const float4 xj = pos[j];
        float dx, dy, dz, d2, len;
        G2 g;
        geom2(S, xi, xj, dx, dy, dz, d2, len, g);
        float f;
        if (ent & kBoundaryBit) {
            f = -xi.w * (pdd + pdd);
        } else {
            const float nd = vel[j].w;
            const float npdd = nd == 0.0f ? 0.0f : pres[j] / (nd * nd);
            f = -xj.w * (pdd + npdd);
        }
        ax = ax + f * g.x;
        ay = ay + f * g.y;
        az = az + f * g.z;
    }
    float nx = 0.0f + 0.0f, ny = 0.0f + (-S.gravity), nz = 0.0f + 0.0f;
    nx = nx + S.visc2 * tx;
    ny = ny + S.visc2 * ty;
    nz = nz + S.visc2 * tz;
    S.acc[0][i] = make_float4(ax + nx, ay + ny, az + nz, 0.0f);
}

// ---- integration (no CFL correction in 2D) and the cell box of the next search ------------
__global__ __launch_bounds__(kBlock) void k2_integrate(const SceneDev* __restrict__ scenes,
                                                       const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    __shared__ int32_t blo[2][kBlock / 64], bhi[2][kBlock / 64];
    const uint32_t i = br.block * kBlock + threadIdx.x;
    int32_t lo[2] = {INT32_MAX, INT32_MAX}, hi[2] = {INT32_MIN, INT32_MIN};
    float nv2 = 0.0f, na2 = 0.0f;  // |v|^2, |a|^2 of the new state (inactive: excluded)
    if (i < S.n) {
        const uint32_t type = type_of(S, i);
        float4 x = S.pos[0][i];
        if (type != 2u) {
            const float4 a = S.acc[0][i];
            float4 v = S.vel[0][i];
            const float dt = S.dt;
            float vx, vy, vz;
            if (S.flags & kFlagSesph) {
                vx = v.x + dt * a.x;
                vy = v.y + dt * a.y;
                vz = v.z + dt * a.z;
            } else {
                const float4 pv = S.vstar[i];
                vx = pv.x + dt * a.x;
                vy = pv.y + dt * a.y;
                vz = pv.z + dt * a.z;
                if (type == 0u) v.w = S.rho_new[i];
            }
            x.x = x.x + dt * vx;
            x.y = x.y + dt * vy;
            x.z = x.z + dt * vz;
            S.pos[0][i] = x;
            S.vel[0][i] = make_float4(vx, vy, vz, v.w);
            nv2 = vx * vx + vy * vy;  // glm::length of a vec2
            na2 = a.x * a.x + a.y * a.y;
        }
        const int2 c = cell2(x.x, x.y, S.search_radius);
        lo[0] = hi[0] = c.x;
        lo[1] = hi[1] = c.y;
    }
#pragma unroll
    for (int a = 0; a < 2; a++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            lo[a] = min(lo[a], __shfl_xor(lo[a], o, 64));
            hi[a] = max(hi[a], __shfl_xor(hi[a], o, 64));
        }
        if (__lane_id() == 0) {
            blo[a][threadIdx.x / 64] = lo[a];
            bhi[a][threadIdx.x / 64] = hi[a];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        BoxPartial bp{};
        for (int a = 0; a < 2; a++) {
            int32_t l = INT32_MAX, h = INT32_MIN;
            for (int w = 0; w < kBlock / 64; w++) {
                l = min(l, blo[a][w]);
                h = max(h, bhi[a][w]);
            }
            bp.lo[a] = l;
            bp.hi[a] = h;
        }
        S.box[br.block] = bp;
    }
    {  // max |v|^2, |a|^2 of the new state for the dynamic CFL time step
        __shared__ float lmax[2 * kBlock / 64];
        block_max2(nv2, na2, lmax);
        if (threadIdx.x == 0) {
            S.box[br.block].lo[3] = __float_as_int(nv2);
            S.box[br.block].hi[3] = __float_as_int(na2);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k2_integrate_finalize(const SceneDev* __restrict__ scenes,
                                                                const BlockRef* __restrict__ map) {
    const SceneDev& S = scenes[map[blockIdx.x].scene];
    if (scene_failed(S)) return;
    __shared__ int32_t blo[2][kBlock / 64], bhi[2][kBlock / 64];
    int32_t l[2] = {INT32_MAX, INT32_MAX}, h[2] = {INT32_MIN, INT32_MIN};
    for (uint32_t b = threadIdx.x; b < S.nblocks; b += kBlock) {
        const BoxPartial bp = S.box[b];
        for (int a = 0; a < 2; a++) {
            l[a] = min(l[a], bp.lo[a]);
            h[a] = max(h[a], bp.hi[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 2; a++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            l[a] = min(l[a], __shfl_xor(l[a], o, 64));
            h[a] = max(h[a], __shfl_xor(h[a], o, 64));
        }
        if (__lane_id() == 0) {
            blo[a][threadIdx.x / 64] = l[a];
            bhi[a][threadIdx.x / 64] = h[a];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        SceneState* st = S.state;
        for (int a = 0; a < 2; a++) {
            int32_t lo2 = INT32_MAX, hi2 = INT32_MIN;
            for (int w = 0; w < kBlock / 64; w++) {
                lo2 = min(lo2, blo[a][w]);
                hi2 = max(hi2, bhi[a][w]);
            }
            st->bbox_min[a] = lo2;
            st->bbox_max[a] = hi2;
        }
        st->bbox_min[2] = st->bbox_max[2] = 0;
        st->dt_actual = S.dt;
        st->max_v = 0.0f;
        st->max_a = 0.0f;
        if (S.flags & kFlagSesph) {
            st->iterations = 0;
            st->avg_err = 0.0;
        }
    }
    {  // max |v|^2, |a|^2 of the new state for the dynamic CFL time step
        __shared__ float lmax[2 * kBlock / 64];
        float v2 = 0.0f, a2 = 0.0f;
        for (uint32_t b = threadIdx.x; b < S.nblocks; b += kBlock) {
            const BoxPartial bp = S.box[b];
            v2 = fmaxf(v2, __int_as_float(bp.lo[3]));
            a2 = fmaxf(a2, __int_as_float(bp.hi[3]));
        }
        block_max2(v2, a2, lmax);
        if (threadIdx.x == 0) {
            S.state->next_v2 = v2;
            S.state->next_a2 = a2;
        }
    }
}

}  // namespace

// ---- launchers ------------------------------------------------------------------------
void launch2_search(const LaunchSet& L, uint32_t entry_passes, uint32_t cell_passes, hipStream_t st) {
    k2_grid_update<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
    uint32_t buf = 0;
    if (entry_passes) {
        launch_sort_range(L, 0u, 0u, entry_passes, st);
        buf = entry_passes & 1u;
    }
    k2_cell_keys<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all, buf);
    launch_sort_range(L, buf, 0u, cell_passes, st);
    buf ^= cell_passes & 1u;
    k2_cell_table<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all, buf);
    if (L.nb_fluid) k2_lists<<<L.nb_fluid, kBlock, 0, st>>>(L.scenes, L.map_fluid, buf);
}

void launch2_iisph_prepare(const LaunchSet& L, hipStream_t st) {
    if (L.nb_fluid) k2_iisph_prepare<<<L.nb_fluid, kBlock, 0, st>>>(L.scenes, L.map_fluid);
    if (L.nb_all) k2_prepare_boundary<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
}
void launch2_iisph_diagonal(const LaunchSet& L, hipStream_t st) {
    if (L.nb_fluid) k2_iisph_diagonal<<<L.nb_fluid, kBlock, 0, st>>>(L.scenes, L.map_fluid);
}
void launch2_iisph_begin_loop(const LaunchSet& L, hipStream_t st) {
    k2_iisph_begin_loop<<<L.nb_scene, 64, 0, st>>>(L.scenes, L.map_scene);
}
void launch2_iisph_pass_a(const LaunchSet& L, hipStream_t st) {
    if (L.nb_fluid) k2_iisph_pass_a<<<L.nb_fluid, kBlock, 0, st>>>(L.scenes, L.map_fluid);
}
void launch2_iisph_pass_b(const LaunchSet& L, hipStream_t st) {
    if (L.nb_fluid) k2_iisph_pass_b<<<L.nb_fluid, kBlock, 0, st>>>(L.scenes, L.map_fluid);
    k2_iisph_converge<<<L.nb_scene, kBlock, 0, st>>>(L.scenes, L.map_scene);
}
void launch2_sesph(const LaunchSet& L, hipStream_t st) {
    if (L.nb_fluid) {
        k2_sesph_density<<<L.nb_fluid, kBlock, 0, st>>>(L.scenes, L.map_fluid);
        k2_sesph_commit<<<L.nb_fluid, kBlock, 0, st>>>(L.scenes, L.map_fluid);
        k2_sesph_accel<<<L.nb_fluid, kBlock, 0, st>>>(L.scenes, L.map_fluid);
    }
}
void launch2_integrate(const LaunchSet& L, hipStream_t st) {
    k2_integrate<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
    k2_integrate_finalize<<<L.nb_scene, kBlock, 0, st>>>(L.scenes, L.map_scene);
}

}  // namespace fsa
