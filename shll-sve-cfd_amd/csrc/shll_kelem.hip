// shll_kelem.hip -- per-cell kernels around the step loop.
// F64F stores its state as F64 does, so it shares F64's k_init, k_output and k_sums.
//   k_init    : primitives -> state      (Compute_U_from_P, base_shll.c:74-85,
//                                          base_shll_2d.c:124-137)
//   k_output  : state -> primitives, conserved (Compute_P_from_U + the arrays
//                                          Save_Results prints, base_shll.c:161-192)
//   k_sums    : FP64 conservation sums and a non-finite / non-positive census
//   k_compare : FP64 L1 / Linf / max-relative differences of two primitive sets
//   k_stage   : sound speed and split fluxes of every cell from the state, as the step
//               kernels compute them (Compute_P_from_U + Compute_F_from_P,
//               base_shll.c:130-177, base_shll_2d.c:239-319), for stage-level checks
// Reductions: per-lane sums in a fixed cell order, a wave64 xor-butterfly
// (__shfl_xor), then the waves of a block in index order; blocks write partials
// that the host adds in index order, so results are deterministic.
#include <mutex>

#include "shll_flux.h"

namespace shll {

// primitives (rho, u, (v), T) -> conserved, per precision
template <int PREC, int DIM>
__global__ __launch_bounds__(256) void k_init(const Inst* insts, const double* __restrict__ prim, void* out0,
                                              void* out1, int64_t plane) {
    const Inst in = insts[blockIdx.y];
    const int64_t ncell = (int64_t)in.nx * in.ny;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = in.off + c;
        const double r = prim[o], vx = prim[plane + o];
        const double vy = DIM == 2 ? prim[2 * plane + o] : 0.0;
        const double T = prim[(DIM == 2 ? 3 : 2) * plane + o];
        double U[4];
        if constexpr (PREC == PREC_REF32) {
            const float rf = (float)r, vxf = (float)vx, vyf = (float)vy, Tf = (float)T, cv = in.k.f_cv;
            float* u = (float*)out0;
            u[o] = rf;
            u[plane + o] = rf * vxf;
            if constexpr (DIM == 1) {
                double kin = in.ke == 0 ? (0.5 * (double)vxf) * (double)vxf : 0.5 * (double)(vxf * vxf);
                u[2 * plane + o] = (float)((double)rf * ((double)(Tf * cv) + kin));
            } else {
                u[2 * plane + o] = rf * vyf;
                u[3 * plane + o] = (float)((double)rf * ((double)(Tf * cv) + 0.5 * (double)(vxf * vxf + vyf * vyf)));
            }
            continue;
        }
        const double cv = in.k.d_cv;
        U[0] = r;
        U[1] = r * vx;
        if constexpr (DIM == 1) {
            U[2] = in.ke == 0 ? r * (T * cv + (0.5 * vx) * vx) : r * (T * cv + 0.5 * (vx * vx));
        } else {
            U[2] = r * vy;
            U[3] = r * (T * cv + 0.5 * (vx * vx + vy * vy));
        }
        if constexpr (PREC == PREC_F64) {
            double* u = (double*)out0;
#pragma unroll
            for (int v = 0; v < DIM + 2; ++v) u[v * plane + o] = U[v];
        } else {
            float* h = (float*)out0;
            float* l = (float*)out1;
#pragma unroll
            for (int v = 0; v < DIM + 2; ++v) {
                float hv = (float)U[v];
                h[v * plane + o] = hv;
                l[v * plane + o] = (float)(U[v] - (double)hv);
            }
        }
    }
}

// conserved values of one cell as doubles
template <int PREC, int DIM>
__device__ __forceinline__ void load_cons(const void* s0, const void* s1, int64_t plane, int64_t o, double (&U)[4]) {
#pragma unroll
    for (int v = 0; v < DIM + 2; ++v) {
        if constexpr (PREC == PREC_F64)
            U[v] = ((const double*)s0)[v * plane + o];
        else if constexpr (PREC == PREC_REF32)
            U[v] = (double)((const float*)s0)[v * plane + o];
        else
            U[v] = (double)((const float*)s0)[v * plane + o] + (double)((const float*)s1)[v * plane + o];
    }
}

template <int PREC, int DIM>
__global__ __launch_bounds__(256) void k_output(const Inst* insts, const void* s0, const void* s1, int64_t plane,
                                                double* __restrict__ prim, double* __restrict__ cons) {
    const Inst in = insts[blockIdx.y];
    const int64_t ncell = (int64_t)in.nx * in.ny;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = in.off + c;
        double U[4] = {0, 0, 0, 0}, p[4] = {0, 0, 0, 0};
        load_cons<PREC, DIM>(s0, s1, plane, o, U);
        if constexpr (PREC == PREC_REF32) {
            const float* u = (const float*)s0;
            float vx, vy = 0.f, T, a;
            if constexpr (DIM == 1) {
                RefArith<float>::prim1(u[o], u[plane + o], u[2 * plane + o], in.ke, in.k.f_cv, in.k.f_gR, vx, T, a);
                p[0] = u[o]; p[1] = vx; p[2] = T;
            } else {
                RefArith<float>::prim2(u[o], u[plane + o], u[2 * plane + o], u[3 * plane + o], in.k.f_cv, in.k.f_gR,
                                       vx, vy, T, a);
                p[0] = u[o]; p[1] = vx; p[2] = vy; p[3] = T;
            }
        } else {
            const double cv = in.k.d_cv, r = U[0], vx = U[1] / r;
            p[0] = r;
            p[1] = vx;
            if constexpr (DIM == 1) {
                p[2] = in.ke == 0 ? ((U[2] / r) - (0.5 * vx) * vx) / cv : ((U[2] / r) - 0.5 * (vx * vx)) / cv;
            } else {
                const double vy = U[2] / r;
                p[2] = vy;
                p[3] = ((U[3] / r) - 0.5 * (vx * vx + vy * vy)) / cv;
            }
        }
#pragma unroll
        for (int v = 0; v < DIM + 2; ++v) {
            if (prim) prim[v * plane + o] = p[v];
            if (cons) cons[v * plane + o] = U[v];
        }
    }
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}
__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x = fmax(x, __shfl_xor(x, d, 64));
    return x;
}

// Block reduction of NS sums then NM maxima; thread 0 writes them to out.
template <int NS, int NM>
__device__ __forceinline__ void block_reduce(double (&s)[NS], double (&mx)[NM], double* out) {
    __shared__ double red[4][NS + NM];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < NS; ++q) s[q] = wave_sum(s[q]);
#pragma unroll
    for (int q = 0; q < NM; ++q) mx[q] = wave_max(mx[q]);
    if (l == 0) {
#pragma unroll
        for (int q = 0; q < NS; ++q) red[w][q] = s[q];
#pragma unroll
        for (int q = 0; q < NM; ++q) red[w][NS + q] = mx[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        for (int q = 0; q < NS; ++q) {
            double t = 0;
            for (int k = 0; k < nw; ++k) t += red[k][q];
            out[q] = t;
        }
        for (int q = 0; q < NM; ++q) {
            double t = red[0][NS + q];
            for (int k = 1; k < nw; ++k) t = fmax(t, red[k][NS + q]);
            out[NS + q] = t;
        }
    }
}

// partial[(inst * gridDim.x + block) * 6 + q]: sums of the DIM+2 conserved variables
// (unused slots 0) and, at q = 4, the count of cells whose density or energy is
// non-finite or not positive (q = 5 is an unused maximum slot).
template <int PREC, int DIM>
__global__ __launch_bounds__(256) void k_sums(const Inst* insts, const void* s0, const void* s1, int64_t plane,
                                              double* partial) {
    const Inst in = insts[blockIdx.y];
    const int64_t ncell = (int64_t)in.nx * in.ny;
    double s[5] = {0, 0, 0, 0, 0};
    double none[1] = {0};
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += (int64_t)gridDim.x * blockDim.x) {
        double U[4] = {0, 0, 0, 0};
        load_cons<PREC, DIM>(s0, s1, plane, in.off + c, U);
#pragma unroll
        for (int v = 0; v < 4; ++v) s[v] += U[v];
        const double E = U[DIM + 1];
        const bool good = U[0] > 0.0 && E > 0.0 && isfinite(U[0]) && isfinite(U[1]) && isfinite(E) &&
                          (DIM == 1 || isfinite(U[2]));
        s[4] += good ? 0.0 : 1.0;
    }
    block_reduce<5, 1>(s, none, partial + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 6);
}

// Differences of two primitive sets (nv planes, ncell cells from offset off, plane stride):
// sums |drho|, |du|, |dp| (p = rho*T, R = 1), maxima |drho|/rho_b, |dp|/p_b, |du|, |u_b|,
// |drho|, |dp|, and a count of non-finite entries.
__global__ __launch_bounds__(256) void k_compare(const double* __restrict__ a, int64_t plane_a,
                                                 const double* __restrict__ b, int64_t plane_b, int64_t ncell,
                                                 int dim, double* partial) {
    double s[4] = {0, 0, 0, 0};
    double mx[6] = {0, 0, 0, 0, 0, 0};
    const int tv = dim == 2 ? 3 : 2;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += (int64_t)gridDim.x * blockDim.x) {
        const double ra = a[c], rb = b[c], ua = a[plane_a + c], ub = b[plane_b + c];
        const double pa = ra * a[tv * plane_a + c], pb = rb * b[tv * plane_b + c];
        const double dr = fabs(ra - rb), du = fabs(ua - ub), dp = fabs(pa - pb);
        double dv = 0, vb = 0;
        if (dim == 2) {
            dv = fabs(a[2 * plane_a + c] - b[2 * plane_b + c]);
            vb = fabs(b[2 * plane_b + c]);
        }
        const bool fin = isfinite(ra) && isfinite(ua) && isfinite(pa) && isfinite(dv);
        s[0] += dr;
        s[1] += du;
        s[2] += dp;
        s[3] += fin ? 0.0 : 1.0;
        mx[0] = fmax(mx[0], dr / fabs(rb));
        mx[1] = fmax(mx[1], dp / fabs(pb));
        mx[2] = fmax(mx[2], fmax(du, dv));
        mx[3] = fmax(mx[3], fmax(fabs(ub), vb));
        mx[4] = fmax(mx[4], dr);
        mx[5] = fmax(mx[5], dp);
    }
    block_reduce<4, 6>(s, mx, partial + (int64_t)blockIdx.x * 10);
}

// Device counter {j0, anchor, jmax} of the constant-parameter batch graphs: set before a
// run of graph chunks (add = 0: ctl = {v0, v1, v2}), its launch index advanced by the last
// node of every chunk (add = 1: ctl[0] += v0).
__global__ __launch_bounds__(64) void k_ctl(int64_t* ctl, int64_t v0, int64_t v1, int64_t v2, int add) {
    if (threadIdx.x == 0) {
        if (add) {
            ctl[0] = ctl[0] + v0;
        } else {
            ctl[0] = v0;
            ctl[1] = v1;
            ctl[2] = v2;
        }
    }
}

hipError_t launch_ctl(hipStream_t st, int64_t* ctl, int64_t v0, int64_t v1, int64_t v2, int add) {
    hipLaunchKernelGGL(k_ctl, dim3(1), dim3(64), 0, st, ctl, v0, v1, v2, add);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
template <int PREC>
static hipError_t init_p(int dim, dim3 g, hipStream_t st, const Inst* insts, const double* prim, void* o0,
                         void* o1, int64_t plane) {
    if (dim == 1) hipLaunchKernelGGL((k_init<PREC, 1>), g, dim3(256), 0, st, insts, prim, o0, o1, plane);
    else hipLaunchKernelGGL((k_init<PREC, 2>), g, dim3(256), 0, st, insts, prim, o0, o1, plane);
    return hipGetLastError();
}

hipError_t launch_init(int prec, int dim, int nblocks, int n_inst, hipStream_t st, const Inst* insts,
                       const double* prim, void* o0, void* o1, int64_t plane) {
    dim3 g(nblocks, n_inst);
    // F64F keeps its state as F64 does (one double plane per variable, same initial values)
    if (prec == PREC_F64 || prec == PREC_F64F) return init_p<PREC_F64>(dim, g, st, insts, prim, o0, o1, plane);
    if (prec == PREC_REF32) return init_p<PREC_REF32>(dim, g, st, insts, prim, o0, o1, plane);
    return init_p<PREC_C32>(dim, g, st, insts, prim, o0, o1, plane);
}

// sound speed a and split fluxes (x: fp, fm; y: hp, hm) of every cell of one instance;
// out holds 1 + 4*nvar planes of `plane` doubles: a, fp, fm, hp, hm
template <int PREC, int DIM>
__global__ __launch_bounds__(256) void k_stage(const Inst* insts, const void* s0, int64_t plane,
                                               double* __restrict__ out) {
    using RT = typename Real<PREC>::T;
    constexpr int NV = DIM + 2;
    const Inst in = insts[blockIdx.y];
    const int64_t ncell = (int64_t)in.nx * in.ny;
    const RT* u = (const RT*)s0;  // the state; the hi word in the compensated forms
    double* a = out;
    double* fp = out + plane;
    double* fm = fp + NV * plane;
    double* hp = fm + NV * plane;
    double* hm = hp + NV * plane;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = in.off + c;
        RT aa;
        if constexpr (DIM == 1) {
            RT p[3], q[3];
            flux1<PREC, RT>(u[o], u[plane + o], u[2 * plane + o], in, p, q, aa);
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                fp[v * plane + o] = p[v];
                fm[v * plane + o] = q[v];
            }
        } else {
            const RT U[4] = {u[o], u[plane + o], u[2 * plane + o], u[3 * plane + o]};
            RT xp[4], xm[4], yp[4], ym[4];
            fluxes2<PREC, RT>(U, in, xp, xm, yp, ym, aa);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                fp[v * plane + o] = xp[v];
                fm[v * plane + o] = xm[v];
                hp[v * plane + o] = yp[v];
                hm[v * plane + o] = ym[v];
            }
        }
        a[o] = aa;
    }
}

template <int PREC>
static hipError_t stage_p(int dim, dim3 g, hipStream_t st, const Inst* insts, const void* s0, int64_t plane,
                          double* out) {
    if (dim == 1) {
        hipLaunchKernelGGL((k_stage<PREC, 1>), g, dim3(256), 0, st, insts, s0, plane, out);
    } else {
        if constexpr (PREC == PREC_F64F) return hipErrorInvalidValue;  // a 1D form
        else hipLaunchKernelGGL((k_stage<PREC, 2>), g, dim3(256), 0, st, insts, s0, plane, out);
    }
    return hipGetLastError();
}

hipError_t launch_stage(int prec, int dim, int nblocks, hipStream_t st, const Inst* inst, const void* s0,
                        int64_t plane, double* out) {
    dim3 g(nblocks, 1);
    switch (prec) {
    case PREC_F64: return stage_p<PREC_F64>(dim, g, st, inst, s0, plane, out);
    case PREC_REF32: return stage_p<PREC_REF32>(dim, g, st, inst, s0, plane, out);
    case PREC_C32: return stage_p<PREC_C32>(dim, g, st, inst, s0, plane, out);
    case PREC_C32K: return stage_p<PREC_C32K>(dim, g, st, inst, s0, plane, out);
    case PREC_F64F: return stage_p<PREC_F64F>(dim, g, st, inst, s0, plane, out);
    default: return hipErrorInvalidValue;
    }
}

template <int PREC>
static hipError_t output_p(int dim, dim3 g, hipStream_t st, const Inst* insts, const void* s0, const void* s1,
                           int64_t plane, double* prim, double* cons) {
    if (dim == 1) hipLaunchKernelGGL((k_output<PREC, 1>), g, dim3(256), 0, st, insts, s0, s1, plane, prim, cons);
    else hipLaunchKernelGGL((k_output<PREC, 2>), g, dim3(256), 0, st, insts, s0, s1, plane, prim, cons);
    return hipGetLastError();
}

hipError_t launch_output(int prec, int dim, int nblocks, int n_inst, hipStream_t st, const Inst* insts,
                         const void* s0, const void* s1, int64_t plane, double* prim, double* cons) {
    dim3 g(nblocks, n_inst);
    if (prec == PREC_F64 || prec == PREC_F64F) return output_p<PREC_F64>(dim, g, st, insts, s0, s1, plane, prim, cons);
    if (prec == PREC_REF32) return output_p<PREC_REF32>(dim, g, st, insts, s0, s1, plane, prim, cons);
    return output_p<PREC_C32>(dim, g, st, insts, s0, s1, plane, prim, cons);
}

template <int PREC>
static hipError_t sums_p(int dim, dim3 g, hipStream_t st, const Inst* insts, const void* s0, const void* s1,
                         int64_t plane, double* partial) {
    if (dim == 1) hipLaunchKernelGGL((k_sums<PREC, 1>), g, dim3(256), 0, st, insts, s0, s1, plane, partial);
    else hipLaunchKernelGGL((k_sums<PREC, 2>), g, dim3(256), 0, st, insts, s0, s1, plane, partial);
    return hipGetLastError();
}

hipError_t launch_sums(int prec, int dim, int nblocks, int n_inst, hipStream_t st, const Inst* insts,
                       const void* s0, const void* s1, int64_t plane, double* partial) {
    dim3 g(nblocks, n_inst);
    if (prec == PREC_F64 || prec == PREC_F64F) return sums_p<PREC_F64>(dim, g, st, insts, s0, s1, plane, partial);
    if (prec == PREC_REF32) return sums_p<PREC_REF32>(dim, g, st, insts, s0, s1, plane, partial);
    return sums_p<PREC_C32>(dim, g, st, insts, s0, s1, plane, partial);
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point (shll_conserved_sums_noalloc): the block
// partials of k_sums for one instance, at most 256 blocks of 6 doubles (the sums kernels
// never use more blocks: their grid strides over the cells).  Every GPU holds its own copy
// of the array at its own address; host code must use the address hipGetSymbolAddress
// returns on that device, never the array's name.  k_sums writes every slot it uses, so
// nothing needs clearing between calls; calls that overlap on one GPU share it and race.
// ---------------------------------------------------------------------------
constexpr int kNoallocScratchDoubles = 256 * 6;
static __device__ __align__(256) double scratch[kNoallocScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on `device` (which must be the current device), looked up
// once per device; capacity in doubles
hipError_t noalloc_scratch(int device, double** out, size_t* capacity) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = (double*)p;
    }
    *out = g_scratch_ptr[device];
    *capacity = kNoallocScratchDoubles;
    return hipSuccess;
}

hipError_t launch_compare(int nblocks, hipStream_t st, const double* a, int64_t plane_a, const double* b,
                          int64_t plane_b, int64_t ncell, int dim, double* partial) {
    hipLaunchKernelGGL(k_compare, dim3(nblocks), dim3(256), 0, st, a, plane_a, b, plane_b, ncell, dim, partial);
    return hipGetLastError();
}

}  // namespace shll
