// valu_ceiling.hip -- the VALU issue ceiling of the GPU, for the step kernels' roofline.
//
// The 1D and 2D step kernels with several steps per launch read and write HBM once per k
// steps, so their bound is instruction issue, not HBM bandwidth.  This program measures
// how many wavefront VALU instructions per second the device issues at most, for FP32
// and FP64 fused multiply-adds: every thread runs NCH independent fma chains (no memory
// traffic inside the loop), a launch fills every SIMD with 8 wavefronts, and HIP events
// time it.  It prints the rate per SIMD per nanosecond and per clock cycle at the
// device's maximum clock.
//
// Usage: valu_ceiling [iterations]   (default 20000)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                            \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                    \
            std::exit(1);                                                                   \
        }                                                                                   \
    } while (0)

constexpr int NCH = 8;  // independent chains per thread

__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <class T>
__global__ __launch_bounds__(256) void k_fma_chains(T* out, int iters, T b, T c) {
    T a[NCH];
#pragma unroll
    for (int q = 0; q < NCH; ++q) a[q] = (T)(threadIdx.x + q);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int q = 0; q < NCH; ++q) a[q] = fma_t(a[q], b, c);
    }
    T s = 0;
#pragma unroll
    for (int q = 0; q < NCH; ++q) s += a[q];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;  // one value per thread of the grid
}

template <class T>
static double rate(const char* name, int iters, int nblocks, double simds, double clock_ghz) {
    T* out = nullptr;
    const size_t n = (size_t)nblocks * 256;
    CHECK(hipMalloc(&out, n * sizeof(T)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    // b slightly below 1 and c small keep the chains finite and normal
    hipLaunchKernelGGL(k_fma_chains<T>, dim3(nblocks), dim3(256), 0, 0, out, iters / 10, (T)0.999, (T)1e-3);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    float best = 1e30f;
    for (int r = 0; r < 5; ++r) {
        CHECK(hipEventRecord(e0, 0));
        hipLaunchKernelGGL(k_fma_chains<T>, dim3(nblocks), dim3(256), 0, 0, out, iters, (T)0.999, (T)1e-3);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms = 0;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) best = ms;
    }
    const double waves = (double)nblocks * 4.0;
    const double insts = waves * (double)iters * NCH;  // wavefront fma instructions of the loop
    const double per_ns = insts / (best * 1e6) / simds;
    std::printf("%s: %d blocks x 256 threads, %d iterations x %d chains: %.3f ms, %.4g wave-instructions/s, "
                "%.4f per SIMD per ns, %.4f per SIMD per cycle at %.3f GHz\n",
                name, nblocks, iters, NCH, best, insts / (best * 1e-3), per_ns, per_ns / clock_ghz, clock_ghz);
    CHECK(hipFree(out));
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    return per_ns;
}

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 20000;
    hipDeviceProp_t p;
    CHECK(hipGetDeviceProperties(&p, 0));
    const double simds = 4.0 * p.multiProcessorCount;
    const double clock_ghz = p.clockRate * 1e-6;  // clockRate is in kHz
    std::printf("device %s, %d CUs (%g SIMDs), max clock %.3f GHz\n", p.gcnArchName, p.multiProcessorCount, simds,
                clock_ghz);
    const int nblocks = p.multiProcessorCount * 8;  // 8 wavefronts per SIMD
    rate<float>("fp32 fma", iters, nblocks, simds, clock_ghz);
    rate<double>("fp64 fma", iters, nblocks, simds, clock_ghz);
    return 0;
}
