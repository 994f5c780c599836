// Static condensation of element blocks (assembly/elementschur.cpp:6-329), FP64.
//
// One workgroup per element (grid-stride).  D is factored in place by LU with
// partial pivoting (LAPACK getrf semantics: first maximal |pivot|, scale by the
// reciprocal, rank-1 trailing update), then
//   F  <- inv(D) F,  Ru <- inv(D) Ru                (getrs)
//   AH  = H - K inv(D) F,  FH = Rh - K inv(D) Ru    (written to H and FH)
// Rh itself is kept (the uncondensed trace residual enters rNorm).
#include "hfem_device.h"

namespace hfem {

__global__ void __launch_bounds__(256)
condense_fp64_kernel(int na, int nb, int e0, int n, double *__restrict__ Dm, double *__restrict__ Fm,
                     const double *__restrict__ Km, double *__restrict__ Hm, double *__restrict__ Ru,
                     const double *__restrict__ Rh, double *__restrict__ FH, int *__restrict__ info)
{
    __shared__ double rv[256];
    __shared__ int ri[256];
    __shared__ int piv[1024];
    const int tid = threadIdx.x, nth = blockDim.x;
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const long e = e0 + t;
        double *D = Dm + e * (long)na * na;
        double *F = Fm + e * (long)na * nb;
        const double *K = Km + e * (long)nb * na;
        double *H = Hm + e * (long)nb * nb;
        double *u = Ru + e * na;
        int sing = 0;
        // ---- LU with partial pivoting ----
        for (int j = 0; j < na; j++) {
            double best = -1.0;
            int bi = na;
            for (int i = j + tid; i < na; i += nth) {
                double v = fabs(D[i + (long)na * j]);
                if (v > best) { best = v; bi = i; }
            }
            rv[tid] = best;
            ri[tid] = bi;
            __syncthreads();
            for (int s = nth / 2; s > 0; s >>= 1) {
                if (tid < s) {
                    double o = rv[tid + s];
                    int oi = ri[tid + s];
                    if (o > rv[tid] || (o == rv[tid] && oi < ri[tid])) { rv[tid] = o; ri[tid] = oi; }
                }
                __syncthreads();
            }
            const int p = ri[0];
            if (tid == 0) piv[j] = p;
            const double pvt = D[p + (long)na * j];
            if (pvt == 0.0) sing = sing ? sing : j + 1;
            __syncthreads();
            if (p != j)
                for (int k = tid; k < na; k += nth) {
                    double a = D[j + (long)na * k];
                    D[j + (long)na * k] = D[p + (long)na * k];
                    D[p + (long)na * k] = a;
                }
            __syncthreads();
            if (pvt != 0.0) {
                const double rinv = 1.0 / pvt;
                for (int i = j + 1 + tid; i < na; i += nth) D[i + (long)na * j] *= rinv;
            }
            __syncthreads();
            const int w = na - j - 1;
            for (int idx = tid; idx < w * w; idx += nth) {
                int i = j + 1 + idx % w, k = j + 1 + idx / w;
                D[i + (long)na * k] -= D[i + (long)na * j] * D[j + (long)na * k];
            }
            __syncthreads();
        }
        if (tid == 0 && info) info[e] = sing;
        // ---- solve: columns of F and Ru, one column per thread ----
        for (int c = tid; c < nb + 1; c += nth) {
            double *b = c < nb ? F + (long)na * c : u;
            for (int j = 0; j < na; j++) {
                int p = piv[j];
                if (p != j) { double a = b[j]; b[j] = b[p]; b[p] = a; }
            }
            for (int j = 0; j < na; j++) {
                double bj = b[j];
                for (int i = j + 1; i < na; i++) b[i] -= D[i + (long)na * j] * bj;
            }
            for (int j = na - 1; j >= 0; j--) {
                b[j] /= D[j + (long)na * j];
                double bj = b[j];
                for (int i = 0; i < j; i++) b[i] -= D[i + (long)na * j] * bj;
            }
        }
        __syncthreads();
        // ---- Schur complement: H -= K F, FH = Rh - K u ----
        for (long idx = tid; idx < (long)nb * nb + nb; idx += nth) {
            if (idx < (long)nb * nb) {
                int r = idx % nb, c = idx / nb;
                double s = 0;
                for (int k = 0; k < na; k++) s += K[r + (long)nb * k] * F[k + (long)na * c];
                H[idx] -= s;
            } else {
                int r = idx - (long)nb * nb;
                double s = 0;
                for (int k = 0; k < na; k++) s += K[r + (long)nb * k] * u[k];
                FH[e * nb + r] = Rh[e * nb + r] - s;
            }
        }
        __syncthreads();
    }
}

hipError_t launch_condense_fp64(int na, int nb, int e0, int n, double *D, double *F, const double *K, double *H,
                                double *Ru, const double *Rh, double *FH, int *info, int grid, hipStream_t stream)
{
    hipLaunchKernelGGL(condense_fp64_kernel, dim3(grid), dim3(256), 0, stream, na, nb, e0, n, D, F, K, H, Ru, Rh, FH,
                       info);
    return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Mixed precision: FP32 LU (partial pivoting) of a copy of D, FP64 iterative
// refinement of every right-hand side against the FP64 D (residual F - D X in
// FP64), FP64 Schur product.  Work area per workgroup (floats):
//   D32[na*na] | X[na*(nb+1)] doubles | R[na*(nb+1)] doubles | T[na*(nb+1)] floats
// ---------------------------------------------------------------------------
long mixed_work_floats(int na, int nb)
{
    long w = ((long)na * na + 63) / 64 * 64 + 5L * na * (nb + 1);
    return (w + 63) / 64 * 64;
}

__device__ static void solve32(int na, const float *__restrict__ L, const int *piv, const double *rhs,
                               double *x, float *__restrict__ tmp)
{
    // tmp: na floats (thread-private slice of the work area)
    for (int j = 0; j < na; j++) tmp[j] = (float)rhs[j];
    for (int j = 0; j < na; j++) {
        int p = piv[j];
        if (p != j) { float a = tmp[j]; tmp[j] = tmp[p]; tmp[p] = a; }
    }
    for (int j = 0; j < na; j++) {
        float bj = tmp[j];
        for (int i = j + 1; i < na; i++) tmp[i] -= L[i + (long)na * j] * bj;
    }
    for (int j = na - 1; j >= 0; j--) {
        tmp[j] /= L[j + (long)na * j];
        float bj = tmp[j];
        for (int i = 0; i < j; i++) tmp[i] -= L[i + (long)na * j] * bj;
    }
    for (int j = 0; j < na; j++) x[j] = (double)tmp[j];
}

__global__ void __launch_bounds__(256)
condense_mixed_kernel(int na, int nb, int e0, int n, const double *__restrict__ Dm, double *__restrict__ Fm,
                      const double *__restrict__ Km, double *__restrict__ Hm, double *__restrict__ Ru,
                      const double *__restrict__ Rh, double *__restrict__ FH, int *__restrict__ info,
                      float *__restrict__ work, long wstride, int sweeps)
{
    __shared__ float rv[256];
    __shared__ int ri[256];
    __shared__ int piv[1024];
    const int tid = threadIdx.x, nth = blockDim.x;
    float *L = work + (long)blockIdx.x * wstride;
    double *X = (double *)(L + ((long)na * na + 63) / 64 * 64);
    double *R = X + (long)na * (nb + 1);
    float *T = (float *)(R + (long)na * (nb + 1));
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const long e = e0 + t;
        const double *D = Dm + e * (long)na * na;
        double *F = Fm + e * (long)na * nb;
        const double *K = Km + e * (long)nb * na;
        double *H = Hm + e * (long)nb * nb;
        double *u = Ru + e * na;
        int sing = 0;
        for (long i = tid; i < (long)na * na; i += nth) L[i] = (float)D[i];
        __syncthreads();
        for (int j = 0; j < na; j++) {
            float best = -1.0f;
            int bi = na;
            for (int i = j + tid; i < na; i += nth) {
                float v = fabsf(L[i + (long)na * j]);
                if (v > best) { best = v; bi = i; }
            }
            rv[tid] = best;
            ri[tid] = bi;
            __syncthreads();
            for (int s = nth / 2; s > 0; s >>= 1) {
                if (tid < s) {
                    float o = rv[tid + s];
                    int oi = ri[tid + s];
                    if (o > rv[tid] || (o == rv[tid] && oi < ri[tid])) { rv[tid] = o; ri[tid] = oi; }
                }
                __syncthreads();
            }
            const int p = ri[0];
            if (tid == 0) piv[j] = p;
            const float pvt = L[p + (long)na * j];
            if (pvt == 0.0f) sing = sing ? sing : j + 1;
            __syncthreads();
            if (p != j)
                for (int k = tid; k < na; k += nth) {
                    float a = L[j + (long)na * k];
                    L[j + (long)na * k] = L[p + (long)na * k];
                    L[p + (long)na * k] = a;
                }
            __syncthreads();
            if (pvt != 0.0f) {
                const float rinv = 1.0f / pvt;
                for (int i = j + 1 + tid; i < na; i += nth) L[i + (long)na * j] *= rinv;
            }
            __syncthreads();
            const int w = na - j - 1;
            for (int idx = tid; idx < w * w; idx += nth) {
                int i = j + 1 + idx % w, k = j + 1 + idx / w;
                L[i + (long)na * k] -= L[i + (long)na * j] * L[j + (long)na * k];
            }
            __syncthreads();
        }
        if (tid == 0 && info) info[e] = sing;
        for (int c = tid; c < nb + 1; c += nth) {
            const double *b = c < nb ? F + (long)na * c : u;
            double *x = X + (long)na * c, *r = R + (long)na * c;
            float *tmp = T + (long)na * c;
            solve32(na, L, piv, b, x, tmp);
            for (int s = 0; s < sweeps; s++) {
                // FP64 residual against the FP64 D, FP32 correction
                for (int i = 0; i < na; i++) r[i] = b[i];
                for (int k = 0; k < na; k++) {
                    const double xk = x[k];
                    for (int i = 0; i < na; i++) r[i] -= D[i + (long)na * k] * xk;
                }
                solve32(na, L, piv, r, r, tmp);
                for (int i = 0; i < na; i++) x[i] += r[i];
            }
        }
        __syncthreads();
        for (int c = tid; c < nb + 1; c += nth) {
            double *b = c < nb ? F + (long)na * c : u;
            const double *x = X + (long)na * c;
            for (int i = 0; i < na; i++) b[i] = x[i];
        }
        __syncthreads();
        for (long idx = tid; idx < (long)nb * nb + nb; idx += nth) {
            if (idx < (long)nb * nb) {
                int r = idx % nb, c = idx / nb;
                double s = 0;
                for (int k = 0; k < na; k++) s += K[r + (long)nb * k] * F[k + (long)na * c];
                H[idx] -= s;
            } else {
                int r = idx - (long)nb * nb;
                double s = 0;
                for (int k = 0; k < na; k++) s += K[r + (long)nb * k] * u[k];
                FH[e * nb + r] = Rh[e * nb + r] - s;
            }
        }
        __syncthreads();
    }
}

hipError_t launch_condense_mixed(int npv, int ncu, int na, int nb, int e0, int n, double *D, double *F,
                                 const double *K, double *H, double *Ru, const double *Rh, double *FH, int *info,
                                 float *work32, long wstride, int grid, int sweeps, hipStream_t s)
{
    (void)npv;
    (void)ncu;
    hipLaunchKernelGGL(condense_mixed_kernel, dim3(grid), dim3(256), 0, s, na, nb, e0, n, D, F, K, H, Ru, Rh, FH,
                       info, work32, wstride, sweeps);
    return hipGetLastError();
}

}  // namespace hfem
