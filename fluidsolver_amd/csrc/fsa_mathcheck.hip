// fsa_mathcheck.hip — device checks of the pair-math helpers against the IEEE
// operations (a test library: only tests/test_gpu_math.py loads it).
//   fsa_mc_sqrt:   every non-negative float below +inf; mismatches in [2^-96, inf)
//                  and results outside [0, 2^-40] below 2^-96
//   fsa_mc_recip:  recip_pair against the IEEE 1 / b, every mantissa at the given exponents
//   fsa_mc_div:    every divisor mantissa in [1, 2) at the given exponents, against
//                  numerators drawn from a hash (zeros, tiny, extreme and random);
//                  div_pair bit for bit, div_pair_zs by value (zeros of either sign)
//   fsa_mc_neighbor_order: the code order of a cell's 27 neighbors (neighbor_order_word)
//   fsa_mc_int_abs: x86_int_abs of given values
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fsa_device.hpp"

using namespace fsa;

namespace {

__global__ void k_mc_sqrt(uint32_t begin, uint32_t end, unsigned long long* mism, unsigned long long* bad_small) {
    const uint32_t stride = gridDim.x * blockDim.x;
    unsigned long long m = 0, b = 0;
    for (uint32_t u = begin + blockIdx.x * blockDim.x + threadIdx.x; u < end && u >= begin; u += stride) {
        const float x = __uint_as_float(u);
        const float mine = sqrt_rn_pair(x);
        if (u >= 0x0f800000u) {  // x >= 2^-96
            if (__float_as_uint(mine) != __float_as_uint(__builtin_sqrtf(x))) m++;
        } else {
            if (!(mine >= 0.0f && mine <= 9.094947017729282e-13f)) b++;  // 2^-40
        }
        if (u > 0xffffffffu - stride) break;
    }
    if (m) atomicAdd(mism, m);
    if (b) atomicAdd(bad_small, b);
}

__device__ __forceinline__ uint32_t mix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return (uint32_t)(x ^ (x >> 31));
}

// numerator k of divisor b: zeros, values near the range limits, random magnitudes up to b
__device__ __forceinline__ float numerator(float b, uint32_t k, uint64_t seed) {
    const uint32_t h = mix(seed * 1315423911ull + k);
    const float sign = (h & 1u) ? -1.0f : 1.0f;
    switch (k & 7u) {
        case 0: return sign * 0.0f;
        case 1: return sign * b;
        case 2: return sign * 1.7347234759768071e-18f * (1.0f + (float)(h >> 9) * 1.1920929e-07f);  // ~2^-59
        default: {
            // uniform exponent in [2^-59, b], random mantissa
            const int eb = (int)((__float_as_uint(b) >> 23) & 0xff) - 127;
            const int lo = -59;
            const int ex = lo + (int)((h >> 8) % (uint32_t)(eb - lo + 1));
            float a = __uint_as_float(((uint32_t)(ex + 127) << 23) | (mix(h) & 0x7fffffu));
            if (a > b) a = b;
            return sign * a;
        }
    }
}

__global__ void k_mc_div(int exp_lo, int exp_hi, uint32_t per_b, uint64_t seed, unsigned long long* mism,
                         unsigned long long* tested) {
    const uint32_t nmant = 1u << 23;
    const uint64_t total = (uint64_t)nmant * (uint64_t)(exp_hi - exp_lo + 1);
    unsigned long long m = 0, t = 0;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t mant = (uint32_t)(g % nmant);
        const int ex = exp_lo + (int)(g / nmant);
        const float b = __uint_as_float(((uint32_t)(ex + 127) << 23) | mant);
        const float y = recip_pair(b);
        for (uint32_t k = 0; k < per_b; k++) {
            const float a = numerator(b, k, seed ^ g);
            const float want = a / b;
            const float got = div_pair(a, b, y);
            // div_pair_zs: the same value, a zero of either sign
            const float got_zs = div_pair_zs(a, b, y);
            const bool zs_ok = want == 0.0f ? got_zs == 0.0f : __float_as_uint(got_zs) == __float_as_uint(want);
            if (__float_as_uint(got) != __float_as_uint(want) || !zs_ok) m++;
            t++;
        }
    }
    if (m) atomicAdd(mism, m);
    atomicAdd(tested, t);
}

// mismatches of recip_pair against the IEEE 1 / b, every mantissa of each exponent
__global__ void k_mc_recip(int exp_lo, int exp_hi, unsigned long long* mism) {
    const uint32_t nmant = 1u << 23;
    const uint64_t total = (uint64_t)nmant * (uint64_t)(exp_hi - exp_lo + 1);
    unsigned long long m = 0;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t mant = (uint32_t)(g % nmant);
        const int ex = exp_lo + (int)(g / nmant);
        const float b = __uint_as_float(((uint32_t)(ex + 127) << 23) | mant);
        if (__float_as_uint(recip_pair(b)) != __float_as_uint(1.0f / b)) m++;
    }
    if (m) atomicAdd(mism, m);
}

// x86_int_abs of n values
__global__ void k_mc_int_abs(const float* in, float* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = x86_int_abs(in[i]);
}

// the neighbor order of n cells: out[64 * i + k] = the neighbor offset index
// (dx+1)*9 + (dy+1)*3 + (dz+1) of code k, or 255 where no neighbor has code k;
// self[i] = the cell's own code
__global__ void k_mc_neighbor_order(const int32_t* cells, uint32_t L, uint32_t n, uint8_t* out, uint8_t* self) {
    const uint32_t i = blockIdx.x;
    if (i >= n) return;
    const uint32_t k = threadIdx.x;
    const NeighborOrder o = unpack_order(neighbor_order_word(cells[3 * i], cells[3 * i + 1], cells[3 * i + 2], L));
    int32_t dx, dy, dz;
    const bool has = neighbor_offset(o, k, 0, dx) & neighbor_offset(o, k, 1, dy) & neighbor_offset(o, k, 2, dz);
    out[64 * i + k] = has ? (uint8_t)((dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)) : (uint8_t)255;
    if (k == 0) self[i] = (uint8_t)o.self;
}

}  // namespace

// device neighbor order of the cells (n x 3 coordinates) at key level L into
// out (n x 64) and self (n) (n <= 1 << 20)
extern "C" __attribute__((visibility("default"))) int fsa_mc_neighbor_order(const int32_t* cells, uint32_t L,
                                                                           uint32_t n, uint8_t* out, uint8_t* self) {
    if (n == 0 || n > (1u << 20) || L < 1 || L > 9) return 3;
    uint8_t* d = nullptr;
    const size_t cb = 12 * (size_t)n, ob = 64 * (size_t)n;
    if (hipMalloc(&d, cb + ob + n) != hipSuccess) return 1;
    (void)hipMemcpy(d, cells, cb, hipMemcpyHostToDevice);
    k_mc_neighbor_order<<<n, 64>>>((const int32_t*)d, L, n, d + cb, d + cb + ob);
    hipError_t e = hipMemcpy(out, d + cb, ob, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(self, d + cb + ob, n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e == hipSuccess ? 0 : 2;
}

// device x86_int_abs of host values in[0..n) into out (n <= 4096)
extern "C" __attribute__((visibility("default"))) int fsa_mc_int_abs(const float* in, float* out, uint32_t n) {
    if (n == 0 || n > 4096) return 3;
    float* d = nullptr;
    if (hipMalloc(&d, 8 * (size_t)n) != hipSuccess) return 1;
    (void)hipMemcpy(d, in, 4 * (size_t)n, hipMemcpyHostToDevice);
    k_mc_int_abs<<<(n + 255) / 256, 256>>>(d, d + n, n);
    const hipError_t e = hipMemcpy(out, d + n, 4 * (size_t)n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e == hipSuccess ? 0 : 2;
}

extern "C" __attribute__((visibility("default"))) int fsa_mc_recip(int exp_lo, int exp_hi, uint64_t* mism) {
    if (exp_lo < -23 || exp_hi > 21 || exp_lo > exp_hi) return 3;
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, 8) != hipSuccess) return 1;
    (void)hipMemset(d, 0, 8);
    k_mc_recip<<<8192, 256>>>(exp_lo, exp_hi, d);
    unsigned long long h = 0;
    const hipError_t e = hipMemcpy(&h, d, 8, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    *mism = h;
    return e == hipSuccess ? 0 : 2;
}

extern "C" __attribute__((visibility("default"))) int fsa_mc_sqrt(uint64_t* mism, uint64_t* bad_small) {
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, 16) != hipSuccess) return 1;
    (void)hipMemset(d, 0, 16);
    // all non-negative finite floats: [0, 0x7f800000)
    k_mc_sqrt<<<4096, 256>>>(0u, 0x7f800000u, d, d + 1);
    unsigned long long h[2] = {0, 0};
    const hipError_t e = hipMemcpy(h, d, 16, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    *mism = h[0];
    *bad_small = h[1];
    return e == hipSuccess ? 0 : 2;
}

extern "C" __attribute__((visibility("default"))) int fsa_mc_div(int exp_lo, int exp_hi, uint32_t per_b, uint64_t seed,
                                                                 uint64_t* mism, uint64_t* tested) {
    if (exp_lo < -23 || exp_hi > 21 || exp_lo > exp_hi) return 3;
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, 16) != hipSuccess) return 1;
    (void)hipMemset(d, 0, 16);
    k_mc_div<<<8192, 256>>>(exp_lo, exp_hi, per_b, seed, d, d + 1);
    unsigned long long h[2] = {0, 0};
    const hipError_t e = hipMemcpy(h, d, 16, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    *mism = h[0];
    *tested = h[1];
    return e == hipSuccess ? 0 : 2;
}
