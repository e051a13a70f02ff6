// fsa_sph.hip — IISPH-3D and SESPH-3D particle updates on gfx950.
//
// One thread per fluid particle (fluid rank k -> combined index i), walking
// its neighbor list in ascending combined index, which is the order libFluid
// iterates a compressed-search neighbor list.  Accumulators are kept per
// reference loop and updated in the reference's operand order, so each
// per-particle result equals the reference-order CPU restatement bit for bit.
//
// Reference loops (IISPHFluidSolver3D.hpp):
//   k_iisph_prepare   phase 1, :141-212 (viscosity :43-76)
//   k_iisph_diagonal  phase 2, :216-339
//   k_iisph_pass_a    Jacobi pressure acceleration, :348-397
//   k_iisph_pass_b    Jacobi pressure update + error, :402-461, and the
//                     convergence / CFL reductions of :467-522 (FP64 sum)
//   k_integrate       CFL correction :525-535 + integration :539-561
// (SESPHFluidSolver3D.hpp): k_sesph_density :64-75, 196-228;
//   k_sesph_accel :78-91, 231-324; k_cfl_all :93-107; k_integrate :109-129.
#include "fsa_device.hpp"
#include "fsa_kernels.hpp"

namespace fsa {

constexpr float kEps = 1.1920928955078125e-07f;  // std::numeric_limits<float>::epsilon()

struct Grad {
    float x, y, z;
};

// d = x_i - x_j, squared length and length, kernel gradient of the pair
template <uint32_t Spec>
__device__ __forceinline__ void pair_geometry(const SceneDev& S, const float4& xi, const float4& xj, float& dx,
                                              float& dy, float& dz, float& d2, float& len, Grad& g) {
    dx = xi.x - xj.x;
    dy = xi.y - xj.y;
    dz = xi.z - xj.z;
    d2 = dot3(dx, dy, dz, dx, dy, dz);
    len = pair_length<Spec>(d2);
    kernel_grad<Spec>(S, dx, dy, dz, len, g.x, g.y, g.z);
}

__device__ __forceinline__ uint32_t scene_fluid_blocks(const SceneDev& S) { return (S.nf + kBlock - 1) / kBlock; }

// ---- ConstantTimestepGenerator::get_non_cfl_validating_timestep -----------------
// (ConstantTimestepGenerator.cpp:17-41).  The reference's unqualified sqrt on
// a float is ::sqrt(double); the product with lambda_a is formed in double.
__device__ __forceinline__ float cfl_timestep(const SceneDev& S, float max_a, float max_v) {
    float dta = S.gen_dt, dtv = S.gen_dt;
    if (max_a > 0.0f) dta = (float)(sqrt((double)(S.particle_size / max_a)) * (double)S.lambda_a);
    if (max_v > 0.0f) dtv = S.particle_size / max_v * S.lambda_v;
    const float nt = fminf(dta, dtv);
    return fmaxf(nt, 0.000001f);
}

// ---- block reductions with a fixed tree (deterministic) ---------------------------
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_sum_u(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

struct RedLds {
    double err[kBlock / 64];
    uint32_t cnt[kBlock / 64];
    float a2[kBlock / 64];
    float v2[kBlock / 64];
};

__device__ __forceinline__ Partial block_reduce(double err, uint32_t cnt, float a2, float v2, RedLds& lds) {
    err = wave_sum_d(err);
    cnt = wave_sum_u(cnt);
    a2 = wave_max_f(a2);
    v2 = wave_max_f(v2);
    const uint32_t w = threadIdx.x / 64;
    if (__lane_id() == 0) {
        lds.err[w] = err;
        lds.cnt[w] = cnt;
        lds.a2[w] = a2;
        lds.v2[w] = v2;
    }
    __syncthreads();
    Partial p;
    p.err = 0.0;
    p.count = 0;
    p.max_a2 = 0.0f;
    p.max_v2 = 0.0f;
    p.aux = 0;
    for (int k = 0; k < kBlock / 64; k++) {
        p.err += lds.err[k];
        p.count += lds.cnt[k];
        p.max_a2 = fmaxf(p.max_a2, lds.a2[k]);
        p.max_v2 = fmaxf(p.max_v2, lds.v2[k]);
    }
    __syncthreads();
    return p;
}

// Reduce partials[0..nparts) in a fixed order inside one block.
__device__ __forceinline__ Partial reduce_partials(const Partial* parts, uint32_t nparts, RedLds& lds) {
    double err = 0.0;
    uint32_t cnt = 0;
    float a2 = 0.0f, v2 = 0.0f;
    // each thread adds partials t, t + kBlock, t + 2 kBlock, ... in that order;
    // 16 are loaded ahead of their adds (one round of memory latency for up to
    // 16 * kBlock partials, the 1M-particle scenes' 3907 among them)
    constexpr uint32_t kAhead = 16;
    for (uint32_t b0 = threadIdx.x; b0 < nparts; b0 += kAhead * kBlock) {
        Partial q[kAhead];
#pragma unroll
        for (uint32_t u = 0; u < kAhead; u++)
            if (b0 + u * kBlock < nparts) q[u] = parts[b0 + u * kBlock];
#pragma unroll
        for (uint32_t u = 0; u < kAhead; u++) {
            if (b0 + u * kBlock < nparts) {
                err += q[u].err;
                cnt += q[u].count;
                a2 = fmaxf(a2, q[u].max_a2);
                v2 = fmaxf(v2, q[u].max_v2);
            }
        }
    }
    return block_reduce(err, cnt, a2, v2, lds);
}

// ---- IISPH phase 1: viscosity, predicted velocity, density -------------------------
constexpr int kUnroll = 4;

struct PosVel {
    float4 x, v;
};
struct PosAcc {
    float4 x, a;
};
struct PosP {
    float4 x;
    float p;
};

// The Jacobi loop's state at the start of a step's loop (3D IISPH).  first_done:
// the first iteration's convergence check is not launched (every scene of the
// launch has min and max iterations of at least 2, so it could only continue
// the loop, and pass B's check overwrites what it records): the loop starts
// with that iteration counted.
__device__ __forceinline__ void begin_loop_state(const SceneDev& S, uint32_t first_done) {
    SceneState* st = S.state;
    st->active = S.max_it > 0 ? 1u : 0u;
    st->pending = 1u;
    st->iterations = first_done;
    st->max_a2 = 0.0f;
    st->max_v2 = 0.0f;
    st->avg_err = 0.0;
    st->err_count = 0;
}

// loop_start: 0, or 1 + first_done, and the first thread of the scene's block 0
// sets the loop state (no kernel between phase 1 and phase 2 reads it; a scene
// without fluid has no block here and gets it from k_iisph_begin_loop)
template <uint32_t Spec, int U>
__device__ __forceinline__ void iisph_prepare_body(const SceneDev* __restrict__ scenes,
                                                          const BlockRef* __restrict__ map, uint32_t loop_start) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    if (loop_start && br.block == 0 && threadIdx.x == 0) begin_loop_state(S, loop_start - 1);
    const uint32_t k = br.block * kBlock + threadIdx.x;
    if (k >= S.nf) return;
    const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)S.state->cur);  // uniform: the arrays' bases stay scalar
    const uint32_t i = S.fluid_idx[k];
    const auto pos = gload_ptr(S.pos[c]);
    const auto vel = gload_ptr(S.vel[c]);
    const float4 xi = pos[i];
    const float4 vi = vel[i];
    float ex = 0.0f, ey = 0.0f, ez = 0.0f;
    if (S.flags & kFlagExternal) {
        const float4 e = S.ext[c][i];
        ex = e.x;
        ey = e.y;
        ez = e.z;
    }
    const bool sl = (S.flags & kFlagSingleLayer) != 0;
    float tx = 0.0f, ty = 0.0f, tz = 0.0f, rho = 0.0f;
    float ix = 0.0f, iy = 0.0f, iz = 0.0f;  // d_ii of phase 2, from the same pair gradients
    walk_list<Spec, U>(
        S, k, S.nbr_cnt[k],
        [&](uint32_t j) { return PosVel{pos.at(j), vel.at(j)}; },
        [&](uint32_t ent, const PosVel& d) {
            const float4 xj = d.x, vj = d.v;
            float dx, dy, dz, d2, len;
            Grad g;
            pair_geometry<Spec>(S, xi, xj, dx, dy, dz, d2, len, g);
            // d_ii (IISPHFluidSolver3D.hpp:231-262): -sum f_j gradW_ij in list order
            float f;
            if (!(ent & kBoundaryBit))
                f = div_rho0sq<Spec>(S, xj.w);
            else
                f = sl ? div_rho0sq<Spec>(S, 2.0f * S.gamma * xj.w) : div_rho0sq<Spec>(S, 2.0f * xj.w);
            ix = ix - f * g.x;
            iy = iy - f * g.y;
            iz = iz - f * g.z;
            // ComputeViscosityAcceleration: densities of the start-of-step snapshot
            const float nd = vj.w;
            if (nd != 0.0f) {
                const float vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
                const float t = (xj.w / nd) * (dot3(vx, vy, vz, dx, dy, dz) / (d2 + S.visc_c01));
                tx = tx + t * g.x;
                ty = ty + t * g.y;
                tz = tz + t * g.z;
            }
            float m = xj.w;
            if (sl && (ent & kBoundaryBit)) m = m * S.gamma1;
            rho = rho + m * kernel_w<Spec>(S, len);
        });
    float nx = ex + 0.0f, ny = ey + (-S.gravity), nz = ez + 0.0f;
    nx = nx + S.visc2 * tx;
    ny = ny + S.visc2 * ty;
    nz = nz + S.visc2 * tz;
    const float dt = S.dt;
    S.vstar[i] = make_float4(vi.x + dt * nx, vi.y + dt * ny, vi.z + dt * nz, 0.0f);
    S.anp[i] = make_float4(nx, ny, nz, 0.0f);
    S.rho_new[i] = rho;
    S.pres[c][i] = 0.0f;
    S.dii[k] = make_float4(ix, iy, iz, 0.0f);
    if (S.flags & kFlagExternal) S.ext[c][i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// the walk above as a kernel, and with an occupancy hint (at least W waves per
// SIMD): variants the autotuner times
template <uint32_t Spec, int U>
__global__ __launch_bounds__(kBlock) void k_iisph_prepare(const SceneDev* __restrict__ scenes,
                                                  const BlockRef* __restrict__ map, uint32_t loop_start) {
    iisph_prepare_body<Spec, U>(scenes, map, loop_start);
}
template <uint32_t Spec, int U, int W>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W))) void k_iisph_prepare_w(
    const SceneDev* __restrict__ scenes, const BlockRef* __restrict__ map, uint32_t loop_start) {
    iisph_prepare_body<Spec, U>(scenes, map, loop_start);
}

// phase 1 for boundary particles: no forces of their own, density untouched
// (IISPHFluidSolver3D.hpp:152-176, 180-182)
__global__ __launch_bounds__(kBlock) void k_iisph_prepare_boundary(const SceneDev* __restrict__ scenes,
                                                                   const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    const uint32_t i = br.block * kBlock + threadIdx.x;
    if (i >= S.n) return;
    const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)S.state->cur);  // uniform: the arrays' bases stay scalar
    if ((S.idt[c][i] >> kTypeShift) != 1u) return;
    float4 e = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (S.flags & kFlagExternal) e = S.ext[c][i];
    const float4 v = S.vel[c][i];
    const float dt = S.dt;
    S.vstar[i] = make_float4(v.x + dt * e.x, v.y + dt * e.y, v.z + dt * e.z, 0.0f);
    S.anp[i] = make_float4(e.x, e.y, e.z, 0.0f);
    S.pres[c][i] = 0.0f;
    if (S.flags & kFlagExternal) S.ext[c][i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Pass B's end for fluid particle i (IISPHFluidSolver3D.hpp:440-462): ap *= dt^2,
// the relaxed Jacobi pressure update, the predicted density error, and this
// particle's share of the convergence and CFL partials.
__device__ __forceinline__ void pressure_update(const SceneDev& S, uint32_t c, uint32_t i, const float4& ai, float ap,
                                                double& err_sum, uint32_t& counted, float& a2, float& v2) {
    const float dt = S.dt;
    ap = ap * (dt * dt);
    const float dg = S.diag[i];
    const float s = S.src[i];
    float p = S.pres[c][i];
    if (fabsf(dg) > kEps)
        p = fmaxf(0.0f, p + S.omega * ((s - ap) / dg));
    else
        p = 0.0f;
    float e = fabsf(s - ap);
    if (fabsf(dg) > kEps) {
        if (fabsf(p) <= kEps) e = 0.0f;
        S.err[c][i] = e;
        err_sum = (double)e;
        counted = 1;
    }
    S.pres[c][i] = p;
    a2 = dot3(ai.x, ai.y, ai.z, ai.x, ai.y, ai.z);
    const float4 vs = S.vstar[i];
    const float px = vs.x + dt * ai.x, py = vs.y + dt * ai.y, pz = vs.z + dt * ai.z;
    v2 = dot3(px, py, pz, px, py, pz);
}

// ---- IISPH phase 2: source term and diagonal element -----------------------------
// IISPHFluidSolver3D.hpp:216-339.  The reference walks a particle's neighbors
// twice: for the source term and d_ii, then for a_ii, which needs the whole
// d_ii.  Phase 1 already forms every pair gradient, so it sums d_ii there (the
// same terms in the same order, bit for bit), and phase 2 walks once for the
// source term and a_ii together.
template <uint32_t Spec, int U>
__device__ __forceinline__ void iisph_diagonal_body(const SceneDev* __restrict__ scenes,
                                                           const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    // the first Jacobi iteration rides along (the loop state is set before this kernel)
    const bool first = S.state->active != 0;
    __shared__ RedLds lds;
    double err_sum = 0.0;
    uint32_t counted = 0;
    float a2 = 0.0f, v2 = 0.0f;
    const uint32_t k = br.block * kBlock + threadIdx.x;
    if (k < S.nf) {
        const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)S.state->cur);  // uniform: the arrays' bases stay scalar
        const uint32_t i = S.fluid_idx[k];
        const auto pos = gload_ptr(S.pos[c]);
        const auto vstar = gload_ptr(S.vstar);
        const float4 xi = pos[i];
        const float4 vs = vstar[i];
        const uint32_t cnt = S.nbr_cnt[k];
        // d_ii was summed in phase 1 over the same pairs, in list order
        const float4 di = S.dii[k];
        const float ix = di.x, iy = di.y, iz = di.z;
        const float cm = div_rho0sq<Spec>(S, xi.w);
        float contrib = 0.0f, dg = 0.0f;
        walk_list<Spec, U>(
            S, k, cnt, [&](uint32_t j) { return PosVel{pos.at(j), vstar.at(j)}; },
            [&](uint32_t ent, const PosVel& d) {
                const float4 xj = d.x;
                float dx, dy, dz, d2, len;
                Grad g;
                pair_geometry<Spec>(S, xi, xj, dx, dy, dz, d2, len, g);
                if (!(ent & kBoundaryBit)) {
                    const float4 vsj = d.v;
                    contrib = contrib + xj.w * dot3(vs.x - vsj.x, vs.y - vsj.y, vs.z - vsj.z, g.x, g.y, g.z);
                } else {
                    contrib = contrib + xj.w * dot3(vs.x, vs.y, vs.z, g.x, g.y, g.z);
                }
                dg = dg + xj.w * dot3(ix, iy, iz, g.x, g.y, g.z);
                if (!(ent & kBoundaryBit)) {
                    // (m_i / rho0^2) * gradW(x_i, x_j) with gradW(x_i, x_j) = -gradW(x_j, x_i) exactly
                    dg = dg + xj.w * dot3(-(cm * g.x), -(cm * g.y), -(cm * g.z), g.x, g.y, g.z);
                }
            });
        const float dt = S.dt;
        S.src[i] = S.rest_density - S.rho_new[i] - dt * contrib;
        S.diag[i] = dg * (dt * dt);
        if (first) {
            // Phase 1 set every pressure to zero (IISPHFluidSolver3D.hpp:150-151): the first
            // iteration's pass A subtracts only signed zeros from +0 (accelerations exactly
            // (+0, +0, +0)) and its pass B adds only m_j * dot(+0 - +0, gradW) to +0 (ap exactly
            // +0), so its pressure update follows from this particle's own values.
            const float4 ai = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            S.acc[c][i] = ai;
            pressure_update(S, c, i, ai, 0.0f, err_sum, counted, a2, v2);
        }
    }
    if (first) {
        const Partial part = block_reduce(err_sum, counted, a2, v2, lds);
        if (threadIdx.x == 0) S.partial[br.block] = part;
    }
}

// the walk above as a kernel, and with an occupancy hint (at least W waves per
// SIMD): variants the autotuner times
template <uint32_t Spec, int U>
__global__ __launch_bounds__(kBlock) void k_iisph_diagonal(const SceneDev* __restrict__ scenes,
                                                  const BlockRef* __restrict__ map) {
    iisph_diagonal_body<Spec, U>(scenes, map);
}
template <uint32_t Spec, int U, int W>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W))) void k_iisph_diagonal_w(
    const SceneDev* __restrict__ scenes, const BlockRef* __restrict__ map) {
    iisph_diagonal_body<Spec, U>(scenes, map);
}

// ---- Jacobi loop control -------------------------------------------------------------
// the loop state of every scene (launched when some scene of the set has no
// fluid: phase 1 sets it for the others)
__global__ void k_iisph_begin_loop(const SceneDev* __restrict__ scenes, const BlockRef* __restrict__ map,
                                   uint32_t first_done) {
    const SceneDev& S = scenes[map[blockIdx.x].scene];
    if (scene_failed(S)) return;
    if (threadIdx.x == 0) begin_loop_state(S, first_done);
}

// ---- Jacobi pass A: pressure acceleration ------------------------------------------
template <uint32_t Spec, int U>
__device__ __forceinline__ void iisph_pass_a_body(const SceneDev* __restrict__ scenes,
                                                         const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    if (!S.state->active) return;
    const uint32_t k = br.block * kBlock + threadIdx.x;
    if (k >= S.nf) return;
    const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)S.state->cur);  // uniform: the arrays' bases stay scalar
    const uint32_t i = S.fluid_idx[k];
    const auto pos = gload_ptr(S.pos[c]);
    const auto pres = gload_ptr(S.pres[c]);
    const float4 xi = pos[i];
    const float pi = pres[i];
    const float pir = div_rho0sq<Spec>(S, pi);
    const bool sl = (S.flags & kFlagSingleLayer) != 0;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    walk_list<Spec, U>(
        S, k, S.nbr_cnt[k], [&](uint32_t j) { return PosP{pos.at(j), pres.at(j)}; },
        [&](uint32_t ent, const PosP& d) {
            const float4 xj = d.x;
            float dx, dy, dz, d2, len;
            Grad g;
            pair_geometry<Spec>(S, xi, xj, dx, dy, dz, d2, len, g);
            float f;
            if (!(ent & kBoundaryBit)) {
                f = xj.w * (pir + div_rho0sq<Spec>(S, d.p));
            } else {
                f = sl ? div_rho0sq<Spec>(S, S.gamma * xj.w * 2.0f * pi) : div_rho0sq<Spec>(S, xj.w * 2.0f * pi);
            }
            ax = ax - f * g.x;
            ay = ay - f * g.y;
            az = az - f * g.z;
        });
    S.acc[c][i] = make_float4(ax, ay, az, 0.0f);
}

// the walk above as a kernel, and with an occupancy hint (at least W waves per
// SIMD): variants the autotuner times
template <uint32_t Spec, int U>
__global__ __launch_bounds__(kBlock) void k_iisph_pass_a(const SceneDev* __restrict__ scenes,
                                                  const BlockRef* __restrict__ map) {
    iisph_pass_a_body<Spec, U>(scenes, map);
}
template <uint32_t Spec, int U, int W>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W))) void k_iisph_pass_a_w(
    const SceneDev* __restrict__ scenes, const BlockRef* __restrict__ map) {
    iisph_pass_a_body<Spec, U>(scenes, map);
}

// ---- Jacobi pass B: A*p, relaxed pressure update, error, reductions -------------------
template <uint32_t Spec, int U>
__device__ __forceinline__ void iisph_pass_b_body(const SceneDev* __restrict__ scenes,
                                                         const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    if (!S.state->active) return;
    __shared__ RedLds lds;
    const uint32_t k = br.block * kBlock + threadIdx.x;
    const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)S.state->cur);  // uniform: the arrays' bases stay scalar
    double err_sum = 0.0;
    uint32_t counted = 0;
    float a2 = 0.0f, v2 = 0.0f;
    if (k < S.nf) {
        const uint32_t i = S.fluid_idx[k];
        const auto pos = gload_ptr(S.pos[c]);
        const auto acc = gload_ptr(S.acc[c]);
        const float4 xi = pos[i];
        const float4 ai = acc[i];
        float ap = 0.0f;
        walk_list<Spec, U>(
            S, k, S.nbr_cnt[k], [&](uint32_t j) { return PosAcc{pos.at(j), acc.at(j)}; },
            [&](uint32_t ent, const PosAcc& d) {
                const float4 xj = d.x;
                float dx, dy, dz, d2, len;
                Grad g;
                pair_geometry<Spec>(S, xi, xj, dx, dy, dz, d2, len, g);
                if (!(ent & kBoundaryBit)) {
                    const float4 aj = d.a;
                    ap = ap + xj.w * dot3(ai.x - aj.x, ai.y - aj.y, ai.z - aj.z, g.x, g.y, g.z);
                } else {
                    ap = ap + xj.w * dot3(ai.x, ai.y, ai.z, g.x, g.y, g.z);
                }
            });
        pressure_update(S, c, i, ai, ap, err_sum, counted, a2, v2);
    }
    const Partial part = block_reduce(err_sum, counted, a2, v2, lds);
    if (threadIdx.x == 0) S.partial[br.block] = part;
}

// the walk above as a kernel, and with an occupancy hint (at least W waves per
// SIMD): variants the autotuner times
template <uint32_t Spec, int U>
__global__ __launch_bounds__(kBlock) void k_iisph_pass_b(const SceneDev* __restrict__ scenes,
                                                  const BlockRef* __restrict__ map) {
    iisph_pass_b_body<Spec, U>(scenes, map);
}
template <uint32_t Spec, int U, int W>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W))) void k_iisph_pass_b_w(
    const SceneDev* __restrict__ scenes, const BlockRef* __restrict__ map) {
    iisph_pass_b_body<Spec, U>(scenes, map);
}

// convergence decision of one Jacobi iteration (IISPHFluidSolver3D.hpp:463-522),
// one block per scene over the pass-B partials in block order
__global__ __launch_bounds__(kBlock) void k_iisph_converge(const SceneDev* __restrict__ scenes,
                                                           const BlockRef* __restrict__ map) {
    const SceneDev& S = scenes[map[blockIdx.x].scene];
    if (scene_failed(S)) return;
    if (!S.state->active) return;
    __shared__ RedLds lds;
    const Partial tot = reduce_partials(S.partial, scene_fluid_blocks(S), lds);
    if (threadIdx.x == 0) {
        SceneState* st = S.state;
        const uint64_t it = st->iterations;
        st->iterations = (uint32_t)(it + 1);
        const double avg = tot.count > 0 ? tot.err / (double)tot.count : 0.0;
        st->avg_err = avg;
        st->err_count = tot.count;
        st->max_a2 = tot.max_a2;
        st->max_v2 = tot.max_v2;
        const bool reached_min = it >= (uint64_t)S.min_it - 1ull;  // size_t arithmetic as in the reference
        if ((reached_min && avg <= (double)S.max_err) || it + 1 >= (uint64_t)S.max_it) st->active = 0;
    }
}

// ---- SESPH --------------------------------------------------------------------------
template <uint32_t Spec>
__global__ __launch_bounds__(kBlock) void k_sesph_density(const SceneDev* __restrict__ scenes,
                                                          const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    const uint32_t k = br.block * kBlock + threadIdx.x;
    if (k >= S.nf) return;
    const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)S.state->cur);  // uniform: the arrays' bases stay scalar
    const uint32_t i = S.fluid_idx[k];
    const auto pos = gload_ptr(S.pos[c]);
    const float4 xi = pos[i];
    const bool sl = (S.flags & kFlagSingleLayer) != 0;
    float rho = 0.0f;
    walk_list<Spec, kUnroll>(
        S, k, S.nbr_cnt[k], [&](uint32_t j) { return pos.at(j); },
        [&](uint32_t ent, const float4& xj) {
            const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            const float len = pair_length<Spec>(dot3(dx, dy, dz, dx, dy, dz));
            if (sl && (ent & kBoundaryBit))
                rho = rho + S.gamma1 * xj.w * kernel_w<Spec>(S, len);
            else
                rho = rho + xj.w * kernel_w<Spec>(S, len);
        });
    float p = S.stiffness * (rho / S.rest_density - 1.0f);
    p = p < 0.0f ? 0.0f : p;  // std::max(pressure, 0.0f)
    float4 v = S.vel[c][i];
    v.w = rho;
    S.vel[c][i] = v;
    S.pres[c][i] = p;
}

struct PosVelP {
    float4 x, v;
    float p;
};

template <uint32_t Spec>
__global__ __launch_bounds__(kBlock) void k_sesph_accel(const SceneDev* __restrict__ scenes,
                                                        const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    const uint32_t k = br.block * kBlock + threadIdx.x;
    if (k >= S.nf) return;
    const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)S.state->cur);  // uniform: the arrays' bases stay scalar
    const uint32_t i = S.fluid_idx[k];
    const auto pos = gload_ptr(S.pos[c]);
    const auto vel = gload_ptr(S.vel[c]);
    const auto pres = gload_ptr(S.pres[c]);
    const float4 xi = pos[i];
    const float4 vi = vel[i];
    const float rho_i = vi.w, p_i = pres[i];
    const float pdd = rho_i == 0.0f ? 0.0f : p_i / (rho_i * rho_i);
    const bool sl = (S.flags & kFlagSingleLayer) != 0;
    float tx = 0.0f, ty = 0.0f, tz = 0.0f;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    walk_list<Spec, kUnroll>(
        S, k, S.nbr_cnt[k],
        [&](uint32_t j) { return PosVelP{pos.at(j), vel.at(j), pres.at(j)}; },
        [&](uint32_t ent, const PosVelP& d) {
            const float4 xj = d.x, vj = d.v;
            float dx, dy, dz, d2, len;
            Grad g;
            pair_geometry<Spec>(S, xi, xj, dx, dy, dz, d2, len, g);
            const float nd = vj.w;
            if (nd != 0.0f) {
                const float vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
                const float t = (xj.w / nd) * (dot3(vx, vy, vz, dx, dy, dz) / (d2 + S.visc_c01));
                tx = tx + t * g.x;
                ty = ty + t * g.y;
                tz = tz + t * g.z;
            }
            if (ent & kBoundaryBit) {
                const float f = -xi.w * (pdd + pdd);
                if (sl) {
                    ax = ax + (f * g.x) * S.gamma2;
                    ay = ay + (f * g.y) * S.gamma2;
                    az = az + (f * g.z) * S.gamma2;
                } else {
                    ax = ax + f * g.x;
                    ay = ay + f * g.y;
                    az = az + f * g.z;
                }
            } else {
                const float npdd = nd == 0.0f ? 0.0f : d.p / (nd * nd);
                const float f = -xj.w * (pdd + npdd);
                ax = ax + f * g.x;
                ay = ay + f * g.y;
                az = az + f * g.z;
            }
        });
    float nx = 0.0f + 0.0f, ny = 0.0f + (-S.gravity), nz = 0.0f + 0.0f;
    nx = nx + S.visc2 * tx;
    ny = ny + S.visc2 * ty;
    nz = nz + S.visc2 * tz;
    S.acc[c][i] = make_float4(ax + nx, ay + ny, az + nz, 0.0f);
}

// SESPH CFL maxima over every active particle, boundary included
__global__ __launch_bounds__(kBlock) void k_cfl_all(const SceneDev* __restrict__ scenes,
                                                    const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    __shared__ RedLds lds;
    const uint32_t i = br.block * kBlock + threadIdx.x;
    const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)S.state->cur);  // uniform: the arrays' bases stay scalar
    float a2 = 0.0f, v2 = 0.0f;
    if (i < S.n) {
        const float4 a = S.acc[c][i];
        const float4 v = S.vel[c][i];
        a2 = dot3(a.x, a.y, a.z, a.x, a.y, a.z);
        const float dt = S.dt;
        const float px = v.x + dt * a.x, py = v.y + dt * a.y, pz = v.z + dt * a.z;
        v2 = dot3(px, py, pz, px, py, pz);
    }
    const Partial part = block_reduce(0.0, 0u, a2, v2, lds);
    if (threadIdx.x == 0) S.partial[br.block] = part;
}

__global__ __launch_bounds__(kBlock) void k_cfl_finalize(const SceneDev* __restrict__ scenes,
                                                         const BlockRef* __restrict__ map) {
    const SceneDev& S = scenes[map[blockIdx.x].scene];
    if (scene_failed(S)) return;
    __shared__ RedLds lds;
    const Partial tot = reduce_partials(S.partial, S.nblocks, lds);
    if (threadIdx.x == 0) {
        S.state->max_a2 = fmaxf(tot.max_a2, S.inactive_a2);
        S.state->max_v2 = fmaxf(tot.max_v2, S.inactive_v2);
        S.state->iterations = 0;
        S.state->avg_err = 0.0;
        S.state->active = 0u;
        S.state->pending = 1u;
    }
}

// ---- CFL correction + Euler-Cromer integration ------------------------------------------
// A scene integrates once per step, after its Jacobi loop has ended: a launch that
// finds the loop still running (it runs on in later passes) or the step already
// integrated leaves the scene alone, so integration can be enqueued right behind
// the predicted passes and again after further ones.
__device__ __forceinline__ bool integration_due(const SceneDev& S) {
    return S.state->active == 0u && S.state->pending != 0u;
}

__global__ __launch_bounds__(kBlock) void k_integrate(const SceneDev* __restrict__ scenes,
                                                      const BlockRef* __restrict__ map) {
    const BlockRef br = map[blockIdx.x];
    const SceneDev& S = scenes[br.scene];
    if (scene_failed(S)) return;
    if (!integration_due(S)) return;
    __shared__ int32_t blo[3][kBlock / 64], bhi[3][kBlock / 64];
    const uint32_t i = br.block * kBlock + threadIdx.x;
    const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)S.state->cur);  // uniform: the arrays' bases stay scalar
    const float max_v = (float)sqrt((double)S.state->max_v2);
    const float max_a = (float)sqrt((double)S.state->max_a2);
    const float corrected = fminf(cfl_timestep(S, max_a, max_v), S.dt);
    const bool sesph = (S.flags & kFlagSesph) != 0;
    int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    float nv2 = 0.0f, na2 = 0.0f;  // |v|^2, |a|^2 of the new state (static boundary: 0)
    if (i < S.n) {
        const uint32_t type = S.idt[c][i] >> kTypeShift;
        float4 x = S.pos[c][i];
        const bool moves = type == 0u || (S.flags & kFlagDynamicBoundary);
        if (moves) {
            const float4 a = S.acc[c][i];
            float4 v = S.vel[c][i];
            float vx, vy, vz;
            if (sesph) {
                vx = v.x + corrected * a.x;
                vy = v.y + corrected * a.y;
                vz = v.z + corrected * a.z;
            } else {
                float4 pv = S.vstar[i];
                if (S.dt != corrected) {
                    const float4 an = S.anp[i];
                    pv.x = pv.x - S.dt * an.x + corrected * an.x;
                    pv.y = pv.y - S.dt * an.y + corrected * an.y;
                    pv.z = pv.z - S.dt * an.z + corrected * an.z;
                }
                vx = pv.x + corrected * a.x;
                vy = pv.y + corrected * a.y;
                vz = pv.z + corrected * a.z;
                if (type == 0u) v.w = S.rho_new[i];
            }
            x.x = x.x + corrected * vx;
            x.y = x.y + corrected * vy;
            x.z = x.z + corrected * vz;
            S.pos[c][i] = x;
            const float ax_ = fabsf(x.x), ay_ = fabsf(x.y), az_ = fabsf(x.z);
            if ((ax_ > 0.0f && ax_ < kTinyCoord) || (ay_ > 0.0f && ay_ < kTinyCoord) || (az_ > 0.0f && az_ < kTinyCoord))
                atomicOr(&S.state->status, kStatusTinyCoord);
            S.vel[c][i] = make_float4(vx, vy, vz, v.w);
            nv2 = dot3(vx, vy, vz, vx, vy, vz);
            na2 = dot3(a.x, a.y, a.z, a.x, a.y, a.z);
        }
        lo[0] = hi[0] = cell_of(x.x, S.search_radius);
        lo[1] = hi[1] = cell_of(x.y, S.search_radius);
        lo[2] = hi[2] = cell_of(x.z, S.search_radius);
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
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
        BoxPartial bp;
        for (int a = 0; a < 3; a++) {
            int32_t l = INT32_MAX, h = INT32_MIN;
            for (int w = 0; w < kBlock / 64; w++) {
                l = min(l, blo[a][w]);
                h = max(h, bhi[a][w]);
            }
            bp.lo[a] = l;
            bp.hi[a] = h;
        }
        bp.lo[3] = 0;
        bp.hi[3] = 0;
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

// bounding box of the cells of the next search, the step's dt statistics, and
// the maxima of |v|^2 and |a|^2 of the new state over the integrate blocks'
// partials: DynamicCflTimestepGenerator::calculate_maximum_velocity_and_acceleration
// (DynamicCflTimestepGenerator.cpp:6-43) for the next desired time step (sqrt
// is correctly rounded and monotone, so sqrt of the maximum is the maximum length).  One wide block per
// scene; minima and maxima do not depend on the order.
constexpr uint32_t kFinalizeThreads = 1024;

__global__ __launch_bounds__(kFinalizeThreads) void k_integrate_finalize(const SceneDev* __restrict__ scenes,
                                                                         const BlockRef* __restrict__ map) {
    const SceneDev& S = scenes[map[blockIdx.x].scene];
    if (scene_failed(S)) return;
    if (!integration_due(S)) return;
    constexpr uint32_t kWaves = kFinalizeThreads / 64;
    __shared__ int32_t red[8][kWaves];
    int32_t l0 = INT32_MAX, l1 = INT32_MAX, l2 = INT32_MAX, h0 = INT32_MIN, h1 = INT32_MIN, h2 = INT32_MIN;
    float v2 = 0.0f, a2 = 0.0f;
    // minima and maxima: 8 partials per thread are loaded ahead of their use
    constexpr uint32_t kAhead = 8;
    for (uint32_t b0 = threadIdx.x; b0 < S.nblocks; b0 += kAhead * kFinalizeThreads) {
        BoxPartial q[kAhead];
#pragma unroll
        for (uint32_t u = 0; u < kAhead; u++)
            if (b0 + u * kFinalizeThreads < S.nblocks) q[u] = S.box[b0 + u * kFinalizeThreads];
#pragma unroll
        for (uint32_t u = 0; u < kAhead; u++) {
            if (b0 + u * kFinalizeThreads >= S.nblocks) break;
            const BoxPartial& bp = q[u];
            l0 = min(l0, bp.lo[0]);
            l1 = min(l1, bp.lo[1]);
            l2 = min(l2, bp.lo[2]);
            h0 = max(h0, bp.hi[0]);
            h1 = max(h1, bp.hi[1]);
            h2 = max(h2, bp.hi[2]);
            v2 = fmaxf(v2, __int_as_float(bp.lo[3]));
            a2 = fmaxf(a2, __int_as_float(bp.hi[3]));
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        l0 = min(l0, __shfl_xor(l0, o, 64));
        l1 = min(l1, __shfl_xor(l1, o, 64));
        l2 = min(l2, __shfl_xor(l2, o, 64));
        h0 = max(h0, __shfl_xor(h0, o, 64));
        h1 = max(h1, __shfl_xor(h1, o, 64));
        h2 = max(h2, __shfl_xor(h2, o, 64));
        v2 = fmaxf(v2, __shfl_xor(v2, o, 64));
        a2 = fmaxf(a2, __shfl_xor(a2, o, 64));
    }
    const uint32_t w = threadIdx.x / 64;
    if (__lane_id() == 0) {
        red[0][w] = l0;
        red[1][w] = l1;
        red[2][w] = l2;
        red[3][w] = h0;
        red[4][w] = h1;
        red[5][w] = h2;
        red[6][w] = __float_as_int(v2);
        red[7][w] = __float_as_int(a2);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        SceneState* st = S.state;
        for (int a = 0; a < 3; a++) {
            int32_t lo = INT32_MAX, hi = INT32_MIN;
            for (uint32_t q = 0; q < kWaves; q++) {
                lo = min(lo, red[a][q]);
                hi = max(hi, red[3 + a][q]);
            }
            st->bbox_min[a] = lo;
            st->bbox_max[a] = hi;
        }
        float nv2 = 0.0f, na2 = 0.0f;
        for (uint32_t q = 0; q < kWaves; q++) {
            nv2 = fmaxf(nv2, __int_as_float(red[6][q]));
            na2 = fmaxf(na2, __int_as_float(red[7][q]));
        }
        st->next_v2 = nv2;
        st->next_a2 = na2;
        const float max_v = (float)sqrt((double)st->max_v2);
        const float max_a = (float)sqrt((double)st->max_a2);
        st->dt_actual = fminf(cfl_timestep(S, max_a, max_v), S.dt);
        st->max_v = max_v;
        st->max_a = max_a;
        st->pending = 0u;
    }
}

// ---- launchers ------------------------------------------------------------------------
#define FSA_SPEC_DISPATCH(kernel, grid, map)                                                        \
    do {                                                                                            \
        switch (L.spec) {                                                                           \
            case 15: kernel<15><<<grid, kBlock, 0, st>>>(L.scenes, map); break;                     \
            case 13: kernel<13><<<grid, kBlock, 0, st>>>(L.scenes, map); break;                     \
            case 11: kernel<11><<<grid, kBlock, 0, st>>>(L.scenes, map); break;                     \
            case 9: kernel<9><<<grid, kBlock, 0, st>>>(L.scenes, map); break;                       \
            case 8: kernel<8><<<grid, kBlock, 0, st>>>(L.scenes, map); break;                       \
            case 7: kernel<7><<<grid, kBlock, 0, st>>>(L.scenes, map); break;                       \
            case 5: kernel<5><<<grid, kBlock, 0, st>>>(L.scenes, map); break;                       \
            case 3: kernel<3><<<grid, kBlock, 0, st>>>(L.scenes, map); break;                       \
            case 1: kernel<1><<<grid, kBlock, 0, st>>>(L.scenes, map); break;                       \
            default: kernel<0><<<grid, kBlock, 0, st>>>(L.scenes, map); break;                      \
        }                                                                                           \
    } while (0)

// spec x neighbor-walk unroll (the unroll chosen by the autotuner per kernel family)
#define FSA_SPEC_U(kernel, U, grid, ...)                                                            \
    do {                                                                                            \
        switch (L.spec) {                                                                           \
            case 15: kernel<15, U><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;                  \
            case 13: kernel<13, U><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;                  \
            case 11: kernel<11, U><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;                  \
            case 9: kernel<9, U><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;                    \
            case 8: kernel<8, U><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;                    \
            case 7: kernel<7, U><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;                    \
            case 5: kernel<5, U><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;                    \
            case 3: kernel<3, U><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;                    \
            case 1: kernel<1, U><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;                    \
            default: kernel<0, U><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;                   \
        }                                                                                           \
    } while (0)
// spec x unroll x occupancy hint (the variant chosen by the autotuner per kernel family:
// the unroll in the low nibble, the waves-per-SIMD hint in the high nibble)
#define FSA_SPEC_UW(kernel, U, W, grid, ...)                                                        \
    do {                                                                                            \
        switch (L.spec) {                                                                           \
            case 15: kernel##_w<15, U, W><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;           \
            case 13: kernel##_w<13, U, W><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;           \
            case 11: kernel##_w<11, U, W><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;           \
            case 9: kernel##_w<9, U, W><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;             \
            case 8: kernel##_w<8, U, W><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;             \
            case 7: kernel##_w<7, U, W><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;             \
            case 5: kernel##_w<5, U, W><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;             \
            case 3: kernel##_w<3, U, W><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;             \
            case 1: kernel##_w<1, U, W><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;             \
            default: kernel##_w<0, U, W><<<grid, kBlock, 0, st>>>(L.scenes, __VA_ARGS__); break;            \
        }                                                                                           \
    } while (0)
#define FSA_UNROLL_DISPATCH(kernel, family, grid, ...)                                              \
    do {                                                                                            \
        switch (L.unroll[family]) {                                                                 \
            case kVariantU2: FSA_SPEC_U(kernel, 2, grid, __VA_ARGS__); break;                               \
            case kVariantU8: FSA_SPEC_U(kernel, 8, grid, __VA_ARGS__); break;                               \
            case kVariantU4W8: FSA_SPEC_UW(kernel, 4, 8, grid, __VA_ARGS__); break;                         \
            case kVariantU8W6: FSA_SPEC_UW(kernel, 8, 6, grid, __VA_ARGS__); break;                         \
            default: FSA_SPEC_U(kernel, 4, grid, __VA_ARGS__); break;                                       \
        }                                                                                           \
    } while (0)

// phase 1 also sets the Jacobi loop state of every scene with fluid
uint32_t loop_start(const LaunchSet& L) { return L.skip_first_check ? 2u : 1u; }
void launch_iisph_prepare(const LaunchSet& L, bool boundary_pass, hipStream_t st) {
    if (L.nb_fluid) FSA_UNROLL_DISPATCH(k_iisph_prepare, kTunePrepare, L.nb_fluid, L.map_fluid, loop_start(L));
    if (boundary_pass && L.nb_all) k_iisph_prepare_boundary<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
}
void launch_iisph_diagonal(const LaunchSet& L, hipStream_t st) {
    if (L.nb_fluid) FSA_UNROLL_DISPATCH(k_iisph_diagonal, kTuneDiagonal, L.nb_fluid, L.map_fluid);
}
void launch_iisph_begin_loop(const LaunchSet& L, hipStream_t st) {
    if (L.fluidless_scene)
        k_iisph_begin_loop<<<L.nb_scene, 64, 0, st>>>(L.scenes, L.map_scene, L.skip_first_check ? 1u : 0u);
}
void launch_iisph_pass_a(const LaunchSet& L, hipStream_t st) {
    if (L.nb_fluid) FSA_UNROLL_DISPATCH(k_iisph_pass_a, kTunePassA, L.nb_fluid, L.map_fluid);
}
// the first Jacobi iteration's convergence decision (its update ran in phase 2)
void launch_iisph_first_converge(const LaunchSet& L, hipStream_t st) {
    if (!L.skip_first_check) k_iisph_converge<<<L.nb_scene, kBlock, 0, st>>>(L.scenes, L.map_scene);
}
void launch_iisph_pass_b(const LaunchSet& L, hipStream_t st) {
    if (L.nb_fluid) FSA_UNROLL_DISPATCH(k_iisph_pass_b, kTunePassB, L.nb_fluid, L.map_fluid);
    k_iisph_converge<<<L.nb_scene, kBlock, 0, st>>>(L.scenes, L.map_scene);
}
void launch_tuned_kernel(const LaunchSet& L, int family, hipStream_t st) {
    if (!L.nb_fluid) return;
    switch (family) {
        case kTunePrepare:
            FSA_UNROLL_DISPATCH(k_iisph_prepare, kTunePrepare, L.nb_fluid, L.map_fluid, loop_start(L));
            break;
        case kTuneDiagonal: FSA_UNROLL_DISPATCH(k_iisph_diagonal, kTuneDiagonal, L.nb_fluid, L.map_fluid); break;
        case kTunePassA: FSA_UNROLL_DISPATCH(k_iisph_pass_a, kTunePassA, L.nb_fluid, L.map_fluid); break;
        default: FSA_UNROLL_DISPATCH(k_iisph_pass_b, kTunePassB, L.nb_fluid, L.map_fluid); break;
    }
}
void launch_sesph_density(const LaunchSet& L, hipStream_t st) {
    if (L.nb_fluid) FSA_SPEC_DISPATCH(k_sesph_density, L.nb_fluid, L.map_fluid);
}
void launch_sesph_accel(const LaunchSet& L, hipStream_t st) {
    if (L.nb_fluid) FSA_SPEC_DISPATCH(k_sesph_accel, L.nb_fluid, L.map_fluid);
    k_cfl_all<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
    k_cfl_finalize<<<L.nb_scene, kBlock, 0, st>>>(L.scenes, L.map_scene);
}
void launch_integrate(const LaunchSet& L, hipStream_t st) {
    k_integrate<<<L.nb_all, kBlock, 0, st>>>(L.scenes, L.map_all);
    k_integrate_finalize<<<L.nb_scene, kFinalizeThreads, 0, st>>>(L.scenes, L.map_scene);
}

}  // namespace fsa
