// fsa_entities.hip — simulation entities around the solver and the global
// sensors, on gfx950.
//
// Entities (Simulator.cpp:26-45 runs them after the search, before and after
// the solver, in list order):
//   k_entities_before  VelocityAlterationByTag (VelocityAlterationByTag.cpp:6-23) and
//                      BoundaryPreprocessor (BoundaryPreprocessor.cpp:7-48), one thread
//                      per particle; the boundary sums walk the neighbors in the order
//                      the scene's search lists them
//   k_remover_check    ParticleRemover3D (ParticleRemover3D.cpp:8-16, VolumeGroup.cpp:43-58):
//                      flags that some active particle is a member of a remover's group;
//                      the host then marks the members inactive and re-lays the scene out
// Sensors (ParticleStatistics.cpp:17-207): k_statistics / k_statistics_finalize.
#include "fsa_device.hpp"
#include "fsa_kernels.hpp"

namespace fsa {

// Visits the particles of the 27 cells around (cx, cy, cz) in the order the
// scene's search iterates candidates (its NeighborhoodInterface order):
//   compressed-3d  ascending sorted index: the 27 cell runs taken by start
//                  (CompressedNeighbors.cpp:538-592 sorts the 27 cell keys);
//   hashed-3d      cells dx-outer, dz-inner, each cell's members ascending
//                  (HashedNeighborhoodSearch3D.cpp:272-349, std::set per cell);
//   quadratic-3d   ascending index (QuadraticNeighborhoodSearch3D.cpp:200-220):
//                  a merge of the cells' ascending member lists.
template <class Visit>
__device__ void walk_in_search_order(const SceneDev& S, int32_t cx, int32_t cy, int32_t cz, Visit&& visit) {
    constexpr uint32_t kNone = 0xffffffffu;
    if (!(S.flags & kFlagUnsorted)) {
        uint32_t rs[27], re[27];
#pragma unroll
        for (int q = 0; q < 27; q++) {
            rs[q] = kNone;
            re[q] = 0;
            uint32_t lin;
            if (cell_linear(S, cx + q / 9 - 1, cy + (q / 3) % 3 - 1, cz + q % 3 - 1, lin)) {
                const uint4 e = S.cells[lin];
                if (e.z == S.stamp) {
                    rs[q] = e.x;
                    re[q] = e.y;
                }
            }
        }
        // runs are disjoint: take them by ascending start
        uint32_t last = 0;
        bool first = true;
        for (int it = 0; it < 27; it++) {
            uint32_t best = kNone, bend = 0;
#pragma unroll
            for (int q = 0; q < 27; q++) {
                if (rs[q] != kNone && (first || rs[q] > last) && rs[q] < best) {
                    best = rs[q];
                    bend = re[q];
                }
            }
            if (best == kNone) break;
            for (uint32_t j = best; j < bend; j++) visit(j);
            last = best;
            first = false;
        }
        return;
    }
    const uint32_t* __restrict__ vals = S.vals[S.vals_buf];
    if (!(S.flags & kFlagQuadratic)) {
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    uint32_t lin;
                    if (!cell_linear(S, cx + dx, cy + dy, cz + dz, lin)) continue;
                    const uint4 e = S.cells[lin];
                    if (e.z != S.stamp) continue;
                    for (uint32_t m = e.x; m < e.y; m++) visit(vals[m]);
                }
        return;
    }
    uint32_t hb[27], he[27];
#pragma unroll
    for (int q = 0; q < 27; q++) {
        hb[q] = he[q] = 0;
        uint32_t lin;
        if (cell_linear(S, cx + q / 9 - 1, cy + (q / 3) % 3 - 1, cz + q % 3 - 1, lin)) {
            const uint4 e = S.cells[lin];
            if (e.z == S.stamp) {
                hb[q] = e.x;
                he[q] = e.y;
            }
        }
    }
    while (true) {
        uint32_t best = kNone;
        int bq = -1;
#pragma unroll
        for (int q = 0; q < 27; q++) {
            if (hb[q] < he[q]) {
                const uint32_t j = vals[hb[q]];
                if (j < best) {
                    best = j;
                    bq = q;
                }
            }
        }
        if (bq < 0) break;
        visit(best);
#pragma unroll
        for (int q = 0; q < 27; q++)
            if (q == bq) hb[q]++;
    }
}

// Sum of W over the boundary neighbors of particle i in its neighbor list's
// order, the order BoundaryPreprocessor walks; the list test is
// dot(x_i - x_j, x_i - x_j) <= r^2 and W is of |x_i - x_j| (IEEE square root).
// The walk starts from the particle's cell of this search.
__device__ float boundary_volume_reciprocal(const SceneDev& S, uint32_t i, uint32_t c) {
    const float4* __restrict__ pos = S.pos[c];
    const uint32_t* __restrict__ idt = S.idt[c];
    const float4 xi = pos[i];
    const float r2 = S.r2;
    float vr = 0.0f;
    walk_in_search_order(S, cell_of(xi.x, S.search_radius), cell_of(xi.y, S.search_radius),
                         cell_of(xi.z, S.search_radius), [&](uint32_t j) {
                             if ((idt[j] >> kTypeShift) != 1u) return;
                             const float4 xj = pos[j];
                             const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                             const float d2 = dot3(dx, dy, dz, dx, dy, dz);
                             if (!(d2 <= r2)) return;
                             vr = vr + kernel_w<0>(S, __builtin_sqrtf(d2));
                         });
    return vr;
}

__global__ __launch_bounds__(kBlock) void k_entities_before(const SceneDev* __restrict__ scenes,
                                                            const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    if (!(S.entity_bits & (kEntityBitBoundary | kEntityBitVelocity))) return;
    const uint32_t i = br.block * kBlock + threadIdx.x;
    if (i >= S.n) return;
    const uint32_t c = S.state->cur;
    const uint32_t w = S.idt[c][i];
    const uint32_t type = w >> kTypeShift;
    for (uint32_t k = 0; k < S.n_entities; k++) {
        const EntityDev& e = S.entities[k];
        if (e.type == kEntityVelocityAlteration) {
            if (S.tag_of_row[w & kIdMask] == e.tag) {
                float4 v = S.vel[c][i];
                v.x = e.velocity[0];
                v.y = e.velocity[1];
                v.z = e.velocity[2];
                S.vel[c][i] = v;
            }
        } else if (e.type == kEntityBoundaryPreprocessor && type == 1u) {
            // corrected mass = rest_density * gamma_1 / volume_reciprocal, capped (Math::min = std::fmin)
            const float vr = boundary_volume_reciprocal(S, i, c);
            const float corrected = S.rest_density * S.bp_gamma1 / vr;
            S.pos[c][i].w = fminf(S.bp_max_mass, corrected);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_remover_check(const SceneDev* __restrict__ scenes,
                                                          const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    if (!(S.entity_bits & kEntityBitRemover)) return;
    const uint32_t i = br.block * kBlock + threadIdx.x;
    bool hit = false;
    if (i < S.n) {
        const float4 p = S.pos[S.state->cur][i];
        for (uint32_t k = 0; k < S.n_entities; k++) {
            const EntityDev& e = S.entities[k];
            if (e.type != kEntityParticleRemover) continue;
            float lo[3], hi[3];
            for (int a = 0; a < 3; a++) {
                hi[a] = e.center[a] + e.distance_from_center[a];
                lo[a] = e.center[a] - e.distance_from_center[a];
            }
            const bool inside = !(p.x < lo[0] || p.x > hi[0] || p.y > hi[1] || p.y < lo[1] || p.z < lo[2] || p.z > hi[2]);
            if (inside == (e.remove_if_outside == 0u)) hit = true;
        }
    }
    if (__any(hit) && __lane_id() == 0) atomicOr(&S.state->status, kStatusRemoval);
}

void launch_entities_before(const LaunchSet& L, hipStream_t st) {
    if (L.nb_all) k_entities_before<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
}

void launch_remover_check(const LaunchSet& L, hipStream_t st) {
    if (L.nb_all) k_remover_check<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
}

// ---- position queries (NeighborhoodInterface::get_neighbors(position)) --------------------
// The particles near a query position in the search's iteration order, against
// the cell structure of the last search and the current positions (the state a
// sensor sees after a step):
//   compressed: (p - x) . (p - x) <= r^2 over the 27 cells' runs by ascending
//               Morton key (CompressedNeighbors.cpp:536-592);
//   hashed:     length(x - p) <= search_radius, cells dx-outer, each cell's members
//               ascending (HashedNeighborhoodSearch3D.cpp:272-349);
//   quadratic:  length(x - p) <= search_radius over every particle by ascending
//               index (QuadraticNeighborhoodSearch3D.cpp:199-219).
// Inactive particles are not resident, so they are never reported.
// visit(j, dx, dy, dz, d2) receives the combined index and d = x - p.
template <class Visit>
__device__ void for_query_neighbors(const SceneDev& S, const float4* __restrict__ pos, float x, float y, float z,
                                    Visit&& visit) {
    const bool sorted = !(S.flags & kFlagUnsorted);
    auto test = [&](uint32_t j) {
        const float4 p = pos[j];
        const float dx = x - p.x, dy = y - p.y, dz = z - p.z;
        const float d2 = dot3(dx, dy, dz, dx, dy, dz);  // bitwise equal for p - x
        const bool near = sorted ? d2 <= S.r2 : __builtin_sqrtf(d2) <= S.search_radius;
        if (near) visit(j, dx, dy, dz, d2);
    };
    if (!sorted && (S.flags & kFlagQuadratic)) {
        for (uint32_t j = 0; j < S.n; j++) test(j);
        return;
    }
    // the host passes finite positions; cells beyond the grid's one-cell rim
    // hold no neighbors (and the +-1 offsets stay clear of int overflow)
    const float cxf = floorf(x / S.search_radius), cyf = floorf(y / S.search_radius),
                czf = floorf(z / S.search_radius);
    const bool inside = cxf >= (float)(S.cmin[0] - 1) && cxf <= (float)S.cmin[0] + (float)S.cdim[0] &&
                        cyf >= (float)(S.cmin[1] - 1) && cyf <= (float)S.cmin[1] + (float)S.cdim[1] &&
                        czf >= (float)(S.cmin[2] - 1) && czf <= (float)S.cmin[2] + (float)S.cdim[2];
    if (inside) walk_in_search_order(S, (int32_t)cxf, (int32_t)cyf, (int32_t)czf, test);
}

// Neighbor rows of each query: count pass, then a fill pass at the scanned offsets.
template <bool Fill>
__global__ __launch_bounds__(kBlock) void k_query(const SceneDev* __restrict__ scene, const float4* __restrict__ q,
                                                  uint32_t nq, uint32_t* __restrict__ counts,
                                                  const uint64_t* __restrict__ offsets, uint64_t* __restrict__ out) {
    const SceneDev& S = *scene;
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nq) return;
    const float4 x = q[k];
    const uint32_t c = S.state->cur;
    const uint32_t* __restrict__ idt = S.idt[c];
    uint32_t cnt = 0;
    const uint64_t o = Fill ? offsets[k] : 0;
    for_query_neighbors(S, S.pos[c], x.x, x.y, x.z, [&](uint32_t j, float, float, float, float) {
        if (Fill) out[o + cnt] = idt[j] & kIdMask;
        cnt++;
    });
    if (!Fill) counts[k] = cnt;
}

void launch_query_count(const SceneDev* d_scene, const float4* q, uint32_t nq, uint32_t* counts, hipStream_t st) {
    if (nq) k_query<false><<<(nq + kBlock - 1) / kBlock, kBlock, 0, st>>>(d_scene, q, nq, counts, nullptr, nullptr);
}

void launch_query_fill(const SceneDev* d_scene, const float4* q, uint32_t nq, const uint64_t* offsets, uint64_t* out,
                       hipStream_t st) {
    if (nq) k_query<true><<<(nq + kBlock - 1) / kBlock, kBlock, 0, st>>>(d_scene, q, nq, nullptr, offsets, out);
}

// SensorPlane::calculate_for_timepoint (SensorPlane.cpp:75-137), one thread per
// sample: for each sub-sample position
//   origin + span_x * x_step * (x + sub_x) + span_y * y_step * (y + sub_y)
// (spans normalized and scaled by the host: ax = normalize(span_x) * x_step)
// the neighbors in search order add mass, velocity or pressure times
// GetKernelValue(p, sample) = W(|sample - p|) in float; the sum is divided by
// (sub_sample_grid_size + 1)^2.  Row y of sample i counts from the top.
__global__ __launch_bounds__(kBlock) void k_sensor_plane(const SceneDev* __restrict__ scene, PlaneParams P,
                                                         float* __restrict__ out) {
    const SceneDev& S = *scene;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.nx * P.ny) return;
    const uint32_t c = S.state->cur;
    const float4* __restrict__ pos = S.pos[c];
    const float4* __restrict__ vel = S.vel[c];
    const float* __restrict__ pres = S.pres[c];
    const uint32_t y = P.ny - 1 - i / P.nx, x = i % P.nx;
    float vx = 0.0f, vy = 0.0f, vz = 0.0f;
    for (uint32_t sxc = 0; sxc <= P.sub; sxc++) {
        const float sub_x = (float)sxc / P.sub_den - 0.5f;
        const float fx = (float)x + sub_x;
        for (uint32_t syc = 0; syc <= P.sub; syc++) {
            const float sub_y = (float)syc / P.sub_den - 0.5f;
            const float fy = (float)y + sub_y;
            const float px = (P.origin[0] + P.ax[0] * fx) + P.ay[0] * fy;
            const float py = (P.origin[1] + P.ax[1] * fx) + P.ay[1] * fy;
            const float pz = (P.origin[2] + P.ax[2] * fx) + P.ay[2] * fy;
            for_query_neighbors(S, pos, px, py, pz, [&](uint32_t j, float, float, float, float d2) {
                const float w = kernel_w<0>(S, __builtin_sqrtf(d2));
                if (P.type == 0) {
                    vx = vx + pos[j].w * w;  // (mass, 0, 0) * W: the zero lanes add +0
                } else if (P.type == 1) {
                    const float4 v = vel[j];
                    vx = vx + v.x * w;
                    vy = vy + v.y * w;
                    vz = vz + v.z * w;
                } else {
                    vz = vz + pres[j] * w;
                }
            });
        }
    }
    out[3ull * i + 0] = vx / P.div;
    out[3ull * i + 1] = vy / P.div;
    out[3ull * i + 2] = vz / P.div;
}

void launch_sensor_plane(const SceneDev* d_scene, const PlaneParams& P, float* out, hipStream_t st) {
    const uint32_t n = P.nx * P.ny;
    if (n) k_sensor_plane<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(d_scene, P, out);
}

// ParticleIntersectionAccelerator::evaluate_volume_at_position
// (ParticleIntersectionAccelerator.cpp:97-131): over the non-inactive neighbors
// in search order, density += m W(|x - p|), split by type, and the normals
// m * GetKernelDerivativeValue(p, x).  The derivative is minus the reversed
// gradient of d = x - p, negation being exact.  Output per query: density,
// boundary-only density, fluid-only density, fluid-only normal, boundary-only normal.
__global__ __launch_bounds__(kBlock) void k_volume(const SceneDev* __restrict__ scene, const float4* __restrict__ q,
                                                   uint32_t nq, float* __restrict__ out) {
    const SceneDev& S = *scene;
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nq) return;
    const float4 x = q[k];
    const uint32_t c = S.state->cur;
    const float4* __restrict__ pos = S.pos[c];
    const uint32_t* __restrict__ idt = S.idt[c];
    float rho = 0.0f, rb = 0.0f, rf = 0.0f;
    float fnx = 0.0f, fny = 0.0f, fnz = 0.0f, bnx = 0.0f, bny = 0.0f, bnz = 0.0f;
    for_query_neighbors(S, pos, x.x, x.y, x.z, [&](uint32_t j, float dx, float dy, float dz, float d2) {
        const float m = pos[j].w;
        const float len = __builtin_sqrtf(d2);
        const float contribution = m * kernel_w<0>(S, len);
        float gx, gy, gz;
        kernel_grad<0>(S, dx, dy, dz, len, gx, gy, gz);
        const float nx = m * -gx, ny = m * -gy, nz = m * -gz;
        rho = rho + contribution;
        const uint32_t type = idt[j] >> kTypeShift;
        if (type == 0u) {
            rf = rf + contribution;
            fnx = fnx + nx;
            fny = fny + ny;
            fnz = fnz + nz;
        } else if (type == 1u) {
            rb = rb + contribution;
            bnx = bnx + nx;
            bny = bny + ny;
            bnz = bnz + nz;
        }
    });
    float* o = out + 9ull * k;
    o[0] = rho;
    o[1] = rb;
    o[2] = rf;
    o[3] = fnx;
    o[4] = fny;
    o[5] = fnz;
    o[6] = bnx;
    o[7] = bny;
    o[8] = bnz;
}

void launch_volume(const SceneDev* d_scene, const float4* q, uint32_t nq, float* out, hipStream_t st) {
    if (nq) k_volume<<<(nq + kBlock - 1) / kBlock, kBlock, 0, st>>>(d_scene, q, nq, out);
}

// ---- global sensors ---------------------------------------------------------------------
// Per block: counts by type; over normal particles the density, pressure and
// |v| extrema (maxima from FLT_MIN, minima from FLT_MAX, as the reference's
// MaxMinAvgSensorData starts) and FP64 sums of those and of the energies.
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

constexpr float kFltMin = 1.17549435082228750797e-38f;  // std::numeric_limits<float>::min()
constexpr float kFltMax = 3.40282346638528859812e+38f;  // std::numeric_limits<float>::max()

struct StatAcc {
    double sum[5];  // density, pressure, |v|, kinetic, potential
    float mx[3], mn[3];
    uint32_t cnt[3];  // normal, boundary, inactive
    uint32_t pad;
};
static_assert(sizeof(StatAcc) == kStatAccBytes, "kStatAccBytes sizes the statistics scratch");

__device__ void stat_block_reduce(StatAcc& a) {
    __shared__ StatAcc part[kBlock / 64];
    for (int k = 0; k < 5; k++) a.sum[k] = wave_sum_f64(a.sum[k]);
    for (int k = 0; k < 3; k++) {
        a.mx[k] = wave_max(a.mx[k]);
        a.mn[k] = wave_min(a.mn[k]);
        a.cnt[k] = wave_sum_u32(a.cnt[k]);
    }
    __syncthreads();
    if (__lane_id() == 0) part[threadIdx.x / 64] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; w++) {
            for (int k = 0; k < 5; k++) a.sum[k] += part[w].sum[k];
            for (int k = 0; k < 3; k++) {
                a.mx[k] = fmaxf(a.mx[k], part[w].mx[k]);
                a.mn[k] = fminf(a.mn[k], part[w].mn[k]);
                a.cnt[k] += part[w].cnt[k];
            }
        }
    }
}

__device__ void stat_init(StatAcc& a) {
    for (int k = 0; k < 5; k++) a.sum[k] = 0.0;
    for (int k = 0; k < 3; k++) {
        a.mx[k] = kFltMin;
        a.mn[k] = kFltMax;
        a.cnt[k] = 0;
    }
    a.pad = 0;
}

__global__ __launch_bounds__(kBlock) void k_statistics(const SceneDev* __restrict__ scene, float zero_height,
                                                       StatAcc* __restrict__ parts) {
    const SceneDev& S = *scene;
    StatAcc a;
    stat_init(a);
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < S.n) {
        const uint32_t c = S.state->cur;
        const uint32_t type = S.idt[c][i] >> kTypeShift;
        a.cnt[0] = type == 0u ? 1u : 0u;
        a.cnt[1] = type == 1u ? 1u : 0u;
        a.cnt[2] = type >= 2u ? 1u : 0u;
        if (type == 0u) {
            const float4 x = S.pos[c][i];
            const float4 v4 = S.vel[c][i];
            const float rho = v4.w, p = S.pres[c][i], m = x.w;
            const float v = __builtin_sqrtf(dot3(v4.x, v4.y, v4.z, v4.x, v4.y, v4.z));  // glm::length
            a.sum[0] = rho;
            a.sum[1] = p;
            a.sum[2] = v;
            a.sum[3] = 0.5f * m * v * v;                         // GlobalEnergySensor, float terms
            a.sum[4] = S.gravity * m * (x.y - zero_height);
            a.mx[0] = fmaxf(a.mx[0], rho);
            a.mn[0] = fminf(a.mn[0], rho);
            a.mx[1] = fmaxf(a.mx[1], p);
            a.mn[1] = fminf(a.mn[1], p);
            a.mx[2] = fmaxf(a.mx[2], v);
            a.mn[2] = fminf(a.mn[2], v);
        }
    }
    stat_block_reduce(a);
    if (threadIdx.x == 0) parts[blockIdx.x] = a;
}

__global__ __launch_bounds__(kBlock) void k_statistics_finalize(const StatAcc* __restrict__ parts, uint32_t nparts,
                                                                StatAcc* __restrict__ out) {
    StatAcc a;
    stat_init(a);
    for (uint32_t b = threadIdx.x; b < nparts; b += kBlock) {
        const StatAcc& q = parts[b];
        for (int k = 0; k < 5; k++) a.sum[k] += q.sum[k];
        for (int k = 0; k < 3; k++) {
            a.mx[k] = fmaxf(a.mx[k], q.mx[k]);
            a.mn[k] = fminf(a.mn[k], q.mn[k]);
            a.cnt[k] += q.cnt[k];
        }
    }
    stat_block_reduce(a);
    if (threadIdx.x == 0) *out = a;
}

uint64_t statistics_scratch_bytes(uint32_t nblocks) { return sizeof(StatAcc) * ((uint64_t)nblocks + 1); }

uint64_t statistics_result_bytes() { return sizeof(StatAcc); }

void launch_statistics(const SceneDev* d_scene, uint32_t nblocks, float zero_height, void* scratch, hipStream_t st) {
    StatAcc* parts = (StatAcc*)scratch;
    const uint32_t nb = nblocks ? nblocks : 1;
    k_statistics<<<nb, kBlock, 0, st>>>(d_scene, zero_height, parts);
    k_statistics_finalize<<<1, kBlock, 0, st>>>(parts, nb, parts + nb);
}

void read_statistics(const void* host_copy, uint64_t counts[3], float mx[3], float mn[3], double sums[5]) {
    const StatAcc& a = *(const StatAcc*)host_copy;
    for (int k = 0; k < 3; k++) {
        counts[k] = a.cnt[k];
        mx[k] = a.mx[k];
        mn[k] = a.mn[k];
    }
    for (int k = 0; k < 5; k++) sums[k] = a.sum[k];
}

}  // namespace fsa
