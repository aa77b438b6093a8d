// rounding_probe.hip -- checks, on the GPU, the arithmetic identities the float step
// kernels rely on, with the kernels' own functions (shll_device.h):
//   rcp_fast(b)    == 1.0f / b   for every normal b whose reciprocal is normal (exhaustive)
//   div_fast(a, b) == a / b      on random pairs whose quotient is normal
//   sqrt_fast(x)   == sqrtf(x)   for every x in [2^-100, 2^101) (exhaustive)
// and the faithful FP64 sequences of the F64F kernels:
//   rcp_nr2(b)     == 1.0 / b    on random doubles with exponents -1000..1000
//   rsq_nr2(x)     within a stated number of ulp of 1.0 / sqrt(x) on the same kind of arguments
// The right-hand sides are the IEEE operators (-fhip-fp32-correctly-rounded-divide-sqrt).
// Test infrastructure: loaded only by tests/test_gpu_rounding.py.
#include <stdint.h>
#include <stdlib.h>

#include "shll_device.h"

namespace {

constexpr uint32_t NMANT = 1u << 23;

__device__ __forceinline__ uint32_t mix(uint64_t x) {  // SplitMix64 finaliser
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return (uint32_t)(x ^ (x >> 31));
}

// one thread per mantissa; out[t] = mismatches over the exponents
__global__ void k_rcp(uint32_t* out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NMANT) return;
    uint32_t c = 0;
    for (int e = -126; e <= 125; ++e) {
        const float b = __uint_as_float(((uint32_t)(e + 127) << 23) | t);
        c += __float_as_uint(shll::rcp_fast(b)) != __float_as_uint(1.0f / b);
    }
    out[t] = c;
}

__global__ void k_sqrt(uint32_t* out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NMANT) return;
    uint32_t c = 0;
    for (int e = -100; e <= 100; ++e) {
        const float x = __uint_as_float(((uint32_t)(e + 127) << 23) | t);
        c += __float_as_uint(shll::sqrt_fast(x)) != __float_as_uint(sqrtf(x));
    }
    out[t] = c;
}

// reps random pairs per thread: mantissas uniform, exponents in [-60, 60], either sign
__global__ void k_div(uint32_t* out, int reps, uint64_t seed) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NMANT) return;
    uint32_t c = 0;
    for (int i = 0; i < reps; ++i) {
        const uint64_t base = ((uint64_t)t * (uint64_t)reps + (uint64_t)i) * 4ull + seed;
        const uint32_t ma = mix(base) & 0x7fffffu, mb = mix(base + 1) & 0x7fffffu, h = mix(base + 2);
        const int ea = (int)(h % 121u) - 60, eb = (int)((h >> 8) % 121u) - 60;
        const float a = __uint_as_float(((h >> 20) & 1u) << 31 | (uint32_t)(ea + 127) << 23 | ma);
        const float b = __uint_as_float(((h >> 21) & 1u) << 31 | (uint32_t)(eb + 127) << 23 | mb);
        c += __float_as_uint(shll::div_fast(a, b)) != __float_as_uint(a / b);
    }
    out[t] = c;
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// reps random positive doubles per thread: mantissas uniform, exponents in [-1000, 1000]
__device__ __forceinline__ double random_double(uint32_t t, int reps, int i, uint64_t seed) {
    const uint64_t h = mix64(((uint64_t)t * (uint64_t)reps + (uint64_t)i) * 2ull + seed);
    const uint64_t e = (uint64_t)((int)(mix64(h) % 2001ull) - 1000 + 1023);
    return __longlong_as_double((long long)(e << 52 | (h & 0xfffffffffffffull)));
}

__global__ void k_rcp64(uint32_t* out, int reps, uint64_t seed) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NMANT) return;
    uint32_t c = 0;
    for (int i = 0; i < reps; ++i) {
        const double b = random_double(t, reps, i, seed);
        c += __double_as_longlong(shll::rcp_nr2(b)) != __double_as_longlong(1.0 / b);
    }
    out[t] = c;
}

// out[t] = the largest distance in ulp (positive doubles: the difference of the bit patterns)
__global__ void k_rsq64(uint32_t* out, int reps, uint64_t seed) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NMANT) return;
    long long m = 0;
    for (int i = 0; i < reps; ++i) {
        const double x = random_double(t, reps, i, seed);
        long long d = __double_as_longlong(shll::rsq_nr2(x)) - __double_as_longlong(1.0 / sqrt(x));
        d = d < 0 ? -d : d;
        m = d > m ? d : m;
    }
    out[t] = m > 0xffffffffll ? 0xffffffffu : (uint32_t)m;
}

// launch, copy the per-thread counts back and sum them (or take their maximum); -1 on a HIP error
template <class L>
long long count(L launch, bool maximum = false) {
    uint32_t* d = nullptr;
    if (hipMalloc((void**)&d, NMANT * sizeof(uint32_t)) != hipSuccess) return -1;
    uint32_t* h = (uint32_t*)malloc(NMANT * sizeof(uint32_t));
    long long total = -1;
    if (h) {
        launch(d);
        if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
            hipMemcpy(h, d, NMANT * sizeof(uint32_t), hipMemcpyDeviceToHost) == hipSuccess) {
            total = 0;
            for (uint32_t i = 0; i < NMANT; ++i) total = maximum ? (h[i] > total ? h[i] : total) : total + h[i];
        }
        free(h);
    }
    (void)hipFree(d);
    return total;
}

}  // namespace

extern "C" long long probe_rcp_mismatches(void) {
    return count([](uint32_t* d) { hipLaunchKernelGGL(k_rcp, dim3(NMANT / 256), dim3(256), 0, 0, d); });
}
extern "C" long long probe_sqrt_mismatches(void) {
    return count([](uint32_t* d) { hipLaunchKernelGGL(k_sqrt, dim3(NMANT / 256), dim3(256), 0, 0, d); });
}
// pairs tested: 2^23 * reps
extern "C" long long probe_div_mismatches(int reps, unsigned long long seed) {
    return count([=](uint32_t* d) {
        hipLaunchKernelGGL(k_div, dim3(NMANT / 256), dim3(256), 0, 0, d, reps, (uint64_t)seed);
    });
}
// doubles tested: 2^23 * reps each
extern "C" long long probe_rcp64_mismatches(int reps, unsigned long long seed) {
    return count([=](uint32_t* d) {
        hipLaunchKernelGGL(k_rcp64, dim3(NMANT / 256), dim3(256), 0, 0, d, reps, (uint64_t)seed);
    });
}
extern "C" long long probe_rsq64_max_ulp(int reps, unsigned long long seed) {
    return count(
        [=](uint32_t* d) {
            hipLaunchKernelGGL(k_rsq64, dim3(NMANT / 256), dim3(256), 0, 0, d, reps, (uint64_t)seed);
        },
        true);
}
