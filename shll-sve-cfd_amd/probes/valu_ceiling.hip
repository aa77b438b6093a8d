// valu_ceiling.hip -- VALU issue ceilings of the GPU, per instruction class, for the step
// kernels' compute roofline.
//
// The 1D and 2D step kernels with several steps per launch read and write HBM once per k
// steps, so their bound is instruction issue, not HBM bandwidth, and what one wavefront
// instruction costs depends on its class.  For each class the step kernels use, every
// thread runs NCH independent chains of that instruction (no memory traffic inside the
// loop), a launch puts 8 wavefronts on every SIMD, and HIP events time it.  The program
// prints, per class, wavefront instructions per SIMD per nanosecond and per cycle at the
// device's maximum clock; the roofline weights these by a kernel's own instruction mix.
// Each class runs twice: with the constant operands in scalar registers (as the compiler
// places kernel arguments) and in vector registers (as most operands of the step kernels).
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

enum Op { ADD32, MUL32, FMA32, ADD64, MUL64, FMA64, MAX64, DPP32, RCP64, MIX32, MIX64, NOPS };
static const char* kNames[NOPS] = {"v_add_f32", "v_mul_f32", "v_fma_f32", "v_add_f64", "v_mul_f64",
                                   "v_fma_f64", "v_max_f64", "v_mov_b32_dpp", "v_rcp_f64",
                                   "f32 add/mul/fma mix", "f64 add/mul/fma mix"};

template <int OP>
struct Ty { using T = float; };

__device__ __forceinline__ float fma_any(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_any(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <> struct Ty<ADD64> { using T = double; };
template <> struct Ty<MUL64> { using T = double; };
template <> struct Ty<FMA64> { using T = double; };
template <> struct Ty<MAX64> { using T = double; };
template <> struct Ty<RCP64> { using T = double; };
template <> struct Ty<DPP32> { using T = int; };
template <> struct Ty<MIX64> { using T = double; };

template <int OP, class T>
__device__ __forceinline__ T apply(T a, T b, T c, int q) {
    // the mixes: chain q takes add, mul or fma by q % 3
    if constexpr (OP == MIX32 || OP == MIX64) return q % 3 == 0 ? a + c : (q % 3 == 1 ? a * b : fma_any(a, b, c));
    if constexpr (OP == ADD32 || OP == ADD64) return a + c;
    else if constexpr (OP == MUL32 || OP == MUL64) return a * b;
    else if constexpr (OP == FMA32) return __builtin_fmaf(a, b, c);
    else if constexpr (OP == FMA64) return __builtin_fma(a, b, c);
    else if constexpr (OP == MAX64) return __builtin_fmax(a, c);
    else if constexpr (OP == RCP64) return __builtin_amdgcn_rcp(a);
    else return __builtin_amdgcn_update_dpp(a, a, 0x111, 0xf, 0xf, false);  // row_shr:1
}

template <int OP, bool VREG>
__global__ __launch_bounds__(256) void k_chains(typename Ty<OP>::T* out, int iters, typename Ty<OP>::T b,
                                                typename Ty<OP>::T c) {
    using T = typename Ty<OP>::T;
    if constexpr (VREG) asm volatile("" : "+v"(b), "+v"(c));  // operands in vector registers
    T a[NCH];
#pragma unroll
    for (int q = 0; q < NCH; ++q) a[q] = (T)(threadIdx.x + q + 1);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int q = 0; q < NCH; ++q) a[q] = apply<OP, T>(a[q], b, c, q);
    }
    T s = 0;
#pragma unroll
    for (int q = 0; q < NCH; ++q) s += a[q];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;  // one value per thread of the grid
}

template <int OP, bool VREG>
static void rate(int iters, int nblocks, double simds, double clock_ghz) {
    using T = typename Ty<OP>::T;
    T* out = nullptr;
    const size_t n = (size_t)nblocks * 256;
    CHECK(hipMalloc(&out, n * sizeof(T)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    // b slightly below 1 and c small keep the float chains finite and normal
    const T b = (T)0.999, c = (T)1e-3;
    hipLaunchKernelGGL((k_chains<OP, VREG>), dim3(nblocks), dim3(256), 0, 0, out, iters / 10, b, c);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    float best = 1e30f;
    for (int r = 0; r < 5; ++r) {
        CHECK(hipEventRecord(e0, 0));
        hipLaunchKernelGGL((k_chains<OP, VREG>), dim3(nblocks), dim3(256), 0, 0, out, iters, b, c);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms = 0;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) best = ms;
    }
    const double waves = (double)nblocks * 4.0;
    const double insts = waves * (double)iters * NCH;  // wavefront instructions of the loop
    const double per_ns = insts / (best * 1e6) / simds;
    std::printf("%-14s %s  %.3f ms  %.4f per SIMD per ns  %.4f per SIMD per cycle at %.3f GHz\n", kNames[OP],
                VREG ? "vgpr" : "sgpr", best, per_ns, per_ns / clock_ghz, clock_ghz);
    CHECK(hipFree(out));
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
}

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 20000;
    hipDeviceProp_t p;
    CHECK(hipGetDeviceProperties(&p, 0));
    const double simds = 4.0 * p.multiProcessorCount;
    const double clock_ghz = p.clockRate * 1e-6;  // clockRate is in kHz
    std::printf("device %s, %d CUs (%g SIMDs), max clock %.3f GHz; %d iterations x %d chains, 8 waves/SIMD\n",
                p.gcnArchName, p.multiProcessorCount, simds, clock_ghz, iters, NCH);
    const int nblocks = p.multiProcessorCount * 8;  // 8 wavefronts per SIMD
    rate<ADD32, false>(iters, nblocks, simds, clock_ghz);
    rate<ADD32, true>(iters, nblocks, simds, clock_ghz);
    rate<MUL32, false>(iters, nblocks, simds, clock_ghz);
    rate<MUL32, true>(iters, nblocks, simds, clock_ghz);
    rate<FMA32, false>(iters, nblocks, simds, clock_ghz);
    rate<FMA32, true>(iters, nblocks, simds, clock_ghz);
    rate<ADD64, false>(iters, nblocks, simds, clock_ghz);
    rate<ADD64, true>(iters, nblocks, simds, clock_ghz);
    rate<MUL64, false>(iters, nblocks, simds, clock_ghz);
    rate<MUL64, true>(iters, nblocks, simds, clock_ghz);
    rate<FMA64, false>(iters, nblocks, simds, clock_ghz);
    rate<FMA64, true>(iters, nblocks, simds, clock_ghz);
    rate<MAX64, true>(iters, nblocks, simds, clock_ghz);
    rate<DPP32, true>(iters, nblocks, simds, clock_ghz);
    rate<RCP64, true>(iters / 4, nblocks, simds, clock_ghz);
    rate<MIX32, true>(iters, nblocks, simds, clock_ghz);
    rate<MIX64, true>(iters, nblocks, simds, clock_ghz);
    return 0;
}
