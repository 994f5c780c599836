// Static condensation of element blocks (assembly/elementschur.cpp:6-329) on gfx950.
//
// FP64 path: inv(D) per element by blocked Gauss-Jordan on the FP64 MFMA
// (dinv_mfma_kernel); then X = inv(D) [F | Ru] and [AH | FH] = [H | Rh] - K X as
// strided-batched DGEMMs over all elements (host side, hfem_condense).
// Mixed path: inv(D) by the same blocked Gauss-Jordan on the FP32 MFMA (stored
// in FP32), X0 = inv32 [F | Ru] on the FP32 MFMA, then refinement sweeps
// X += inv32 ([F | Ru] - D X) with the residual in FP64, then the same FP64
// Schur product.
// Internal layouts: D column-major (na x na); Fx row-major (na x ldx) holding
// F then Ru in column nb; X the same layout; K column-major (nb x na); Hx
// column-major (nb x (nb+1)) holding H then Rh in column nb.
#include "dinv_body.h"

#include <algorithm>

namespace hfem {

template <typename T, typename TO, int NT, int W, int MW>
__global__ void __launch_bounds__(64 * W, MW)
dinv_mfma_kernel(int na, int e0, int n, const double *__restrict__ D, TO *__restrict__ Dinv, int *__restrict__ info)
{
    extern __shared__ double dv_lds[];
    StrideSource src;
    dinv_body<T, TO, NT, W, 1>(na, e0, n, D, Dinv, info, (char *)dv_lds, src);
}

// Launch shapes of the MFMA inverse per tile count NT (variant 0 is the
// default; the others are the candidates the condensation tuner times):
//   variant 0: 4 waves per element (8 waves for NT > 8), two workgroups per CU
//   variant 1: 2 waves per element
//   variant 2: 1 wave per element (the whole matrix in one wave's registers)
template <typename T, typename TO>
static hipError_t launch_dinv_t(int na, int e0, int n, const double *D, TO *Dinv, int *info, int variant,
                                hipStream_t s)
{
    const int NT = (na + 15) / 16;
    if (NT < 1 || NT > 11 || variant < 0 || variant > 2) return hipErrorInvalidValue;
    if (n <= 0) return hipSuccess;
    const int grid = std::min(n, 65536);
#define HFEM_DINV(NTV, W, MW)                                                                                 \
    hipLaunchKernelGGL((dinv_mfma_kernel<T, TO, NTV, W, MW>), dim3(grid), dim3(64 * W),                         \
                       (dinv_group_lds<T, NTV>()), s, na, e0, n, D, Dinv, info)
    switch (NT) {
#define CASE(NTV, W0)                                                                                         \
    case NTV:                                                                                                 \
        if (variant == 0) HFEM_DINV(NTV, W0, 2);                                                              \
        else if (variant == 1) HFEM_DINV(NTV, 2, 2);                                                          \
        else HFEM_DINV(NTV, 1, 1);                                                                            \
        break;
        CASE(1, 4) CASE(2, 4) CASE(3, 4) CASE(4, 4) CASE(5, 4) CASE(6, 4) CASE(7, 4) CASE(8, 4) CASE(9, 8)
        CASE(10, 8) CASE(11, 8)
#undef CASE
    }
#undef HFEM_DINV
    return hipGetLastError();
}

hipError_t launch_dinv(int na, int e0, int n, const double *D, double *Dinv, int *info, hipStream_t s, int variant)
{
    return launch_dinv_t<double, double>(na, e0, n, D, Dinv, info, variant, s);
}

// inv(D) by the MFMA Gauss-Jordan in FP32, stored in FP32 (the mixed
// condensation's factorisation).
hipError_t launch_dinv32(int na, int e0, int n, const double *D, float *Dinv, int *info, hipStream_t s, int variant)
{
    return launch_dinv_t<float, float>(na, e0, n, D, Dinv, info, variant, s);
}

// ---------------------------------------------------------------------------
// Right-hand-side condensation with stored operands (schurElementVector,
// assembly/elementschur.cpp:345-458, schurImplementation 0): per element
//   y = inv(D) Ru,   FH = Rh - K y
// with inv(D) (column-major na x na) and K (column-major nb x na) kept from the
// last Jacobian in T = double or float (quasiNewtonAccuracy 1 / 0; the float
// copies are widened to double before use, as the reference does).  y goes to
// column nb of X (inv(D) Ru, row-major X), FH to column nb of Hx.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
schur_vector_kernel(int na, int nb, int ldx, int ne, const T *__restrict__ Dinv, const T *__restrict__ K,
                    const double *__restrict__ Ru, const double *__restrict__ Rh, double *__restrict__ X,
                    double *__restrict__ Hx)
{
    extern __shared__ double sv[];
    double *ru = sv, *y = sv + na;
    const int tid = threadIdx.x, nth = blockDim.x;
    const long es = (long)nb * (nb + 1);
    for (int e = blockIdx.x; e < ne; e += gridDim.x) {
        for (int i = tid; i < na; i += nth) ru[i] = Ru[(long)e * na + i];
        __syncthreads();
        const T *De = Dinv + (long)e * na * na;
        for (int i = tid; i < na; i += nth) {
            double s = 0.0;
            for (int k = 0; k < na; k++) s += (double)De[i + (long)na * k] * ru[k];
            y[i] = s;
            X[((long)e * na + i) * ldx + nb] = s;
        }
        __syncthreads();
        const T *Ke = K + (long)e * nb * na;
        for (int r = tid; r < nb; r += nth) {
            double s = Rh[(long)e * nb + r];
            for (int k = 0; k < na; k++) s -= (double)Ke[r + (long)nb * k] * y[k];
            Hx[e * es + (long)nb * nb + r] = s;
        }
        __syncthreads();
    }
}

hipError_t launch_schur_vector(int na, int nb, int ldx, int ne, int fp32, const void *Dinv, const void *K,
                               const double *Ru, const double *Rh, double *X, double *Hx, hipStream_t s)
{
    const int grid = std::min(ne, 4096);
    const size_t lds = 2 * (size_t)na * sizeof(double);
    if (fp32)
        hipLaunchKernelGGL(schur_vector_kernel<float>, dim3(grid), dim3(256), lds, s, na, nb, ldx, ne,
                           (const float *)Dinv, (const float *)K, Ru, Rh, X, Hx);
    else
        hipLaunchKernelGGL(schur_vector_kernel<double>, dim3(grid), dim3(256), lds, s, na, nb, ldx, ne,
                           (const double *)Dinv, (const double *)K, Ru, Rh, X, Hx);
    return hipGetLastError();
}

__global__ void to_f32_kernel(long n, const double *__restrict__ x, float *__restrict__ y)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = (float)x[i];
}

hipError_t launch_to_f32(long n, const double *x, float *y, hipStream_t s)
{
    const long grid = std::min<long>((n + 255) / 256, 8192);
    if (grid > 0) hipLaunchKernelGGL(to_f32_kernel, dim3(grid), dim3(256), 0, s, n, x, y);
    return hipGetLastError();
}

}  // namespace hfem

namespace hfem {

// Refinement bookkeeping of the mixed condensation.  reset clears the per-element
// maxima of the next correction and the unconverged count (a kernel, not a
// memset, so that the sequence can be captured in a graph).
__global__ void refine_reset_kernel(int n, unsigned long long *dmax, unsigned long long *cmax, int *count)
{
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        dmax[t] = 0;
        cmax[t] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *count = 0;
}

// Per element, after correction k: d = max |dX_k|, c = max |X|, p = max |dX_{k-1}|
// (for k = 1, p = c: X0 is the correction from 0).  The error left after the
// sweep is estimated as d (d / p) (linear convergence with factor d / p); an
// element whose estimate exceeds tol c, or is NaN, counts as unconverged.
__global__ void refine_count_kernel(int n, const unsigned long long *dmax, const unsigned long long *cmax,
                                    unsigned long long *dprev, int first, double tol, int *count)
{
    int local = 0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const double d = __longlong_as_double((long long)dmax[t]), c = __longlong_as_double((long long)cmax[t]);
        const double p = first ? c : __longlong_as_double((long long)dprev[t]);
        const double est = p > 0.0 ? d * (d / p) : d;
        if (!(est <= tol * c)) local++;
        dprev[t] = dmax[t];
    }
    if (local) atomicAdd(count, local);
}

hipError_t launch_refine_reset(int n, unsigned long long *dmax, unsigned long long *cmax, int *count, hipStream_t s)
{
    const int grid = std::max(1, std::min((n + 255) / 256, 1024));
    hipLaunchKernelGGL(refine_reset_kernel, dim3(grid), dim3(256), 0, s, n, dmax, cmax, count);
    return hipGetLastError();
}

hipError_t launch_refine_count(int n, const unsigned long long *dmax, const unsigned long long *cmax,
                               unsigned long long *dprev, int first, double tol, int *count, hipStream_t s)
{
    const int grid = std::max(1, std::min((n + 255) / 256, 1024));
    hipLaunchKernelGGL(refine_count_kernel, dim3(grid), dim3(256), 0, s, n, dmax, cmax, dprev, first, tol, count);
    return hipGetLastError();
}

}  // namespace hfem
