// Fused per-element solve X = inv(D) [F Ru] of the static condensation
// (assembly/elementschur.cpp:94-105), one workgroup per element, the element's
// square operands resident in LDS and every column tile of the result kept in
// one wave's registers from load to store.
//
// FP64 mode: inv(D) (FP64, from the inverse kernel) in LDS; X = inv(D) [F Ru]
// on the FP64 MFMA.
// Mixed mode: D (FP64) and inv32 (the FP32 Gauss-Jordan inverse) in LDS;
//   X0 = inv32 [F Ru]                 FP32 MFMA
//   per sweep: R = [F Ru] - D X       FP64 MFMA (the FP64 residual)
//              X += inv32 R           FP32 MFMA, update in FP64
// with the per-element maxima of the last correction and of X recorded for the
// convergence test (refine_count_kernel).  A pass can also start from the X in
// memory (further sweeps).
//
// Register layouts (16x16x4 MFMAs, lane l, lr = l % 16, lk = l / 16): a B
// operand of k step s holds B(4s + lk, lr); the FP64 result holds rows lk + 4v
// of its tile, the FP32 result rows 4lk + v.  So the FP64 result of tile r is
// the B operand of k steps 4r .. 4r+3 as it is, and the FP32 result is a B
// operand whose k order within the tile is permuted (k = 16r + 4lk + v); the A
// operand is read from LDS in the same permuted order, which leaves the
// contraction unchanged.  No data leaves the wave between the products.
#include "ctx.h"

#include <algorithm>

namespace hfem {

typedef double xs_d4 __attribute__((ext_vector_type(4)));
typedef float xs_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ xs_d4 xs_mfma64(double a, double b, xs_d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ xs_f4 xs_mfma32(float a, float b, xs_f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ unsigned long long xs_wave_max(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    return v;
}

// The A operands of one k step for the NT row tiles: row k of the staged
// operand at columns 16r + lr (p points at column lr of row k).
template <typename T, int NT>
__device__ __forceinline__ void xs_lds_row(T (&a)[NT], const T *p)
{
#pragma unroll
    for (int r = 0; r < NT; r++) a[r] = p[16 * r];
}

// LDS row strides (elements): k rows of the staged square operands, padded so
// that the lane groups of an operand read start on different banks.
template <int NT> struct XsLd {
    static constexpr int NN = 16 * NT;
    static constexpr int D64 = NN + 4;    // D: lane groups 4 rows apart (mixed residual)
    static constexpr int I64 = NN + 16 - (NN % 32 == 16 ? 16 : 0);   // inv(D) FP64: groups 1 row apart
    static constexpr int I32 = NN + ((16 - NN) % 64 + 64) % 64;     // inv32: groups 1 row apart
};

size_t xsolve_lds(int mode, int na)
{
    const int NT = (na + 15) / 16, NN = 16 * NT;
    if (mode == 0) return (size_t)na * (NN + 16 - (NN % 32 == 16 ? 16 : 0)) * 8;
    const int i32 = NN + ((16 - NN) % 64 + 64) % 64;
    return (size_t)na * (NN + 4) * 8 + (size_t)na * i32 * 4;
}

template <int NT, int MODE>
__global__ void __launch_bounds__(256, 1)
xsolve_kernel(int na, int N, int ldx, int e0, int n, const double *__restrict__ Dm, const void *__restrict__ Inv,
              const double *__restrict__ Fx, double *__restrict__ X, int nsweep, int from_x,
              unsigned long long *__restrict__ dmax, unsigned long long *__restrict__ cmax)
{
    typedef XsLd<NT> L;
    constexpr int NN = L::NN, KS = 4 * NT;
    extern __shared__ double xs_lds[];
    const int tid = threadIdx.x, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int TC = (N + 15) >> 4;
    double *DL = xs_lds;                                    // MODE 1: D(k, i) at DL[k*D64 + i]
    float *IL = (float *)(xs_lds + (size_t)na * L::D64);    // MODE 1: inv32(k, i) at IL[k*I32 + i]
    double *IL64 = xs_lds;                                  // MODE 0: inv(D)(k, i) at IL64[k*I64 + i]
    // rows i in [na, NN) of every staged k row are zero for the whole launch
    for (int idx = tid; idx < na * (NN - na); idx += 256) {
        const int k = idx / (NN - na), i = na + idx % (NN - na);
        if (MODE == 0) IL64[k * L::I64 + i] = 0.0;
        else {
            DL[k * L::D64 + i] = 0.0;
            IL[k * L::I32 + i] = 0.f;
        }
    }
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const long e = (long)e0 + t;
        // the square operands into LDS, column-major (k, i) -> row k of the stage,
        // eight loads in flight per thread before their stores (rows i >= na stay
        // at the zero written before the element loop)
        {
            const int nn = na * na;
            const float rna = 1.0f / (float)na;   // idx / na in FP32: exact for idx < 2^20
            if (MODE == 0) {
                const double *Ie = (const double *)Inv + e * (long)nn;
                for (int b0 = tid; b0 < nn; b0 += 256 * 8) {
                    double v[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int idx = b0 + 256 * u;
                        v[u] = Ie[idx < nn ? idx : 0];
                    }
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int idx = b0 + 256 * u, k = (int)(((float)idx + 0.5f) * rna);
                        if (idx < nn) IL64[k * L::I64 + idx - k * na] = v[u];
                    }
                }
            } else {
                const double *De = Dm + e * (long)nn;
                const float *Ie = (const float *)Inv + e * (long)nn;
                for (int b0 = tid; b0 < nn; b0 += 256 * 8) {
                    double v[8];
                    float w[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int idx = b0 + 256 * u, q = idx < nn ? idx : 0;
                        v[u] = De[q];
                        w[u] = Ie[q];
                    }
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int idx = b0 + 256 * u, k = (int)(((float)idx + 0.5f) * rna);
                        if (idx < nn) {
                            DL[k * L::D64 + idx - k * na] = v[u];
                            IL[k * L::I32 + idx - k * na] = w[u];
                        }
                    }
                }
            }
        }
        __syncthreads();
        const double *Fe = Fx + e * na * (long)ldx;
        double *Xe = X + e * na * (long)ldx;
        unsigned long long dm = 0, cm = 0;
        for (int ct = wave; ct < TC; ct += 4) {
            const int j = 16 * ct + lr;
            const bool jok = j < N;
            // [F Ru] column tile as B operand: Fv[s] = F(4s + lk, j), zero beyond
            double Fv[KS];
#pragma unroll
            for (int s = 0; s < KS; s++) {
                const int k = 4 * s + lk;
                const bool ok = jok && k < na;
                const double v = Fe[ok ? (long)k * ldx + j : 0];
                Fv[s] = ok ? v : 0.0;
            }
            if (MODE == 0) {
                // X = inv(D) [F Ru]: the NT row tiles as independent accumulators
                xs_d4 acc[NT];
#pragma unroll
                for (int r = 0; r < NT; r++) acc[r] = xs_d4{0.0, 0.0, 0.0, 0.0};
                double a0[NT], a1[NT];
                xs_lds_row(a0, IL64 + min(lk, na - 1) * L::I64 + lr);
#pragma unroll
                for (int s = 0; s < KS; s++) {
                    // k = 4s + lk; k >= na: B is zero (the clamped row is finite)
                    if (s + 1 < KS) xs_lds_row(a1, IL64 + min(4 * (s + 1) + lk, na - 1) * L::I64 + lr);
#pragma unroll
                    for (int r = 0; r < NT; r++) acc[r] = xs_mfma64(a0[r], Fv[s], acc[r]);
#pragma unroll
                    for (int r = 0; r < NT; r++) a0[r] = a1[r];
                    __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll
                for (int r = 0; r < NT; r++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int row = 16 * r + lk + 4 * v;
                        if (row < na && jok) Xe[(long)row * ldx + j] = acc[r][v];
                    }
                continue;
            }
            // mixed: Xd[r][v] = X(16r + 4lk + v, j) (the FP32 result layout)
            double Xd[NT][4];
            if (from_x) {
#pragma unroll
                for (int r = 0; r < NT; r++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int row = 16 * r + 4 * lk + v;
                        const bool ok = jok && row < na;
                        const double x = Xe[ok ? (long)row * ldx + j : 0];
                        Xd[r][v] = ok ? x : 0.0;
                    }
            } else {
                // X0 = inv32 [F Ru]
                xs_f4 acc[NT];
#pragma unroll
                for (int r = 0; r < NT; r++) acc[r] = xs_f4{0.f, 0.f, 0.f, 0.f};
                float a0[NT], a1[NT];
                xs_lds_row(a0, IL + min(lk, na - 1) * L::I32 + lr);
#pragma unroll
                for (int s = 0; s < KS; s++) {
                    if (s + 1 < KS) xs_lds_row(a1, IL + min(4 * (s + 1) + lk, na - 1) * L::I32 + lr);
                    const float b = (float)Fv[s];
#pragma unroll
                    for (int r = 0; r < NT; r++) acc[r] = xs_mfma32(a0[r], b, acc[r]);
#pragma unroll
                    for (int r = 0; r < NT; r++) a0[r] = a1[r];
                    __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll
                for (int r = 0; r < NT; r++)
#pragma unroll
                    for (int v = 0; v < 4; v++) Xd[r][v] = (double)acc[r][v];
            }
            double dmx = 0.0;
            for (int sw = 0; sw < nsweep; sw++) {
                // R = [F Ru] - D X (FP64): R tile r' starts at F(16r' + lk + 4v) = Fv[4r' + v]
                xs_d4 R[NT];
#pragma unroll
                for (int rp = 0; rp < NT; rp++)
#pragma unroll
                    for (int v = 0; v < 4; v++) R[rp][v] = Fv[4 * rp + v];
                double a0[NT], a1[NT];
                xs_lds_row(a0, DL + min(4 * lk, na - 1) * L::D64 + lr);
#pragma unroll
                for (int q = 0; q < 4 * NT; q++) {
                    // k step q = (r, v): k = 16r + 4lk + v; k >= na: b is zero
                    const int r = q >> 2, v = q & 3;
                    if (q + 1 < 4 * NT) {
                        const int r1 = (q + 1) >> 2, v1 = (q + 1) & 3;
                        xs_lds_row(a1, DL + min(16 * r1 + 4 * lk + v1, na - 1) * L::D64 + lr);
                    }
                    const double b = -Xd[r][v];
#pragma unroll
                    for (int rp = 0; rp < NT; rp++) R[rp] = xs_mfma64(a0[rp], b, R[rp]);
#pragma unroll
                    for (int rp = 0; rp < NT; rp++) a0[rp] = a1[rp];
                    __builtin_amdgcn_sched_barrier(0);
                }
                // X += inv32 R: k step (r', v) takes R rows 16r' + lk + 4v
                xs_f4 acc[NT];
#pragma unroll
                for (int r = 0; r < NT; r++) acc[r] = xs_f4{0.f, 0.f, 0.f, 0.f};
                float c0[NT], c1[NT];
                xs_lds_row(c0, IL + min(lk, na - 1) * L::I32 + lr);
#pragma unroll
                for (int q = 0; q < 4 * NT; q++) {
                    // k step q = (r', v): k = 16r' + lk + 4v
                    const int rp = q >> 2, v = q & 3;
                    if (q + 1 < 4 * NT) {
                        const int r1 = (q + 1) >> 2, v1 = (q + 1) & 3;
                        xs_lds_row(c1, IL + min(16 * r1 + lk + 4 * v1, na - 1) * L::I32 + lr);
                    }
                    const float b = (float)R[rp][v];
#pragma unroll
                    for (int r = 0; r < NT; r++) acc[r] = xs_mfma32(c0[r], b, acc[r]);
#pragma unroll
                    for (int r = 0; r < NT; r++) c0[r] = c1[r];
                    __builtin_amdgcn_sched_barrier(0);
                }
                dmx = 0.0;
#pragma unroll
                for (int r = 0; r < NT; r++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const double dx = (double)acc[r][v];
                        Xd[r][v] += dx;
                        dmx = fmax(dmx, fabs(dx));
                    }
            }
            double cmx = 0.0;
#pragma unroll
            for (int r = 0; r < NT; r++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int row = 16 * r + 4 * lk + v;
                    if (row < na && jok) {
                        Xe[(long)row * ldx + j] = Xd[r][v];
                        cmx = fmax(cmx, fabs(Xd[r][v]));
                    }
                }
            const unsigned long long bd = (unsigned long long)__double_as_longlong(dmx);
            const unsigned long long bc = (unsigned long long)__double_as_longlong(cmx);
            dm = bd > dm ? bd : dm;
            cm = bc > cm ? bc : cm;
        }
        if (MODE == 1 && dmax) {
            dm = xs_wave_max(dm);
            cm = xs_wave_max(cm);
            if (lane == 0) {
                atomicMax(dmax + t, dm);
                atomicMax(cmax + t, cm);
            }
        }
        __syncthreads();
    }
}

// Whether the fused solve runs this shape (the square operands fit the LDS).
bool xsolve_fits(int mode, int na, size_t maxlds)
{
    const int NT = (na + 15) / 16;
    return NT >= 1 && NT <= 7 && xsolve_lds(mode, na) <= maxlds;
}

// mode 0: X = inv(D)[F Ru] with the FP64 inverse; mode 1: mixed (inv32 FP32),
// nsweep refinement sweeps, from X0 (from_x 0) or from the X in memory.
hipError_t launch_xsolve(int mode, int na, int N, int ldx, int e0, int n, const double *D, const void *inv,
                         const double *Fx, double *X, int nsweep, int from_x, unsigned long long *dmax,
                         unsigned long long *cmax, hipStream_t s)
{
    const int NT = (na + 15) / 16;
    if (NT < 1 || NT > 7 || n <= 0) return n <= 0 ? hipSuccess : hipErrorInvalidValue;
    const size_t lds = xsolve_lds(mode, na);
    const int grid = std::min(n, 65536);
#define HFEM_XS(NTV)                                                                                              \
    case NTV: {                                                                                                   \
        auto k = mode == 0 ? xsolve_kernel<NTV, 0> : xsolve_kernel<NTV, 1>;                                     \
        hipError_t err = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        if (err != hipSuccess) return err;                                                                        \
        hipLaunchKernelGGL(k, dim3(grid), dim3(256), lds, s, na, N, ldx, e0, n, D, inv, Fx, X, nsweep, from_x,   \
                           dmax, cmax);                                                                           \
        break;                                                                                                    \
    }
    switch (NT) {
        HFEM_XS(1) HFEM_XS(2) HFEM_XS(3) HFEM_XS(4) HFEM_XS(5) HFEM_XS(6) HFEM_XS(7)
    }
#undef HFEM_XS
    return hipGetLastError();
}

}  // namespace hfem
