// Static condensation of element blocks (assembly/elementschur.cpp:6-329) on gfx950.
//
// FP64 path: inv(D) per element by blocked Gauss-Jordan on the FP64 MFMA
// (dinv_mfma_kernel); then X = inv(D) [F | Ru] and [AH | FH] = [H | Rh] - K X as
// strided-batched DGEMMs over all elements (host side, hfem_condense).
// Mixed path: inv(D) by Gauss-Jordan in FP32 (dinv_kernel<R, float>, widened to
// FP64 on the way out), X0 = inv32 [F | Ru], then refinement sweeps
// X += inv32 ([F | Ru] - D X) with the residual in FP64, then the same FP64
// Schur product, all products as batched DGEMMs.
// Internal layouts: D column-major (na x na); Fx row-major (na x ldx) holding
// F then Ru in column nb; X the same layout; K column-major (nb x na); Hx
// column-major (nb x (nb+1)) holding H then Rh in column nb.
#include "hfem_device.h"

#include <algorithm>

namespace hfem {

typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f64x4 mfma64(double a, double b, f64x4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Gauss-Jordan inverse in registers with partial pivoting and implicit row
// exchanges.  Thread (ti, tj) = (tid % 16, tid / 16) owns rows ti + 16 r and
// columns tj + 16 c (r, c < R).  Step j takes as pivot the row p_j of largest
// |a(i, j)| among the rows not yet used (lowest index on ties) and applies the
// exchange step to every row; rows are never moved.  On exit a holds G with
//   inv(D)[label[i], pof[c]] = G[i][c],   pof[j] = p_j, label[p_j] = j.
// The column loop is unrolled by 16-column tile so that the owner of column j
// reads a compile-time register index.  T is the element type.
// ---------------------------------------------------------------------------
template <int R, typename T>
__device__ __forceinline__ int gauss_jordan(int na, T (&a)[R][R], T *colbuf, T *rowbuf, int *pof, int *label)
{
    const int tid = threadIdx.x, ti = tid & 15, tj = tid >> 4, lane = tid & 63;
    for (int i = tid; i < na; i += blockDim.x) label[i] = -1;
    __syncthreads();
    int sing = 0;
#pragma unroll
    for (int c0 = 0; c0 < R; c0++) {
        for (int jj = 0; jj < 16; jj++) {
            const int j = 16 * c0 + jj;
            if (j >= na) break;
            T *cb = colbuf + (j & 1) * 192, *rb = rowbuf + (j & 1) * 192;
            if (tj == jj) {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int i = ti + 16 * r;
                    if (i < na) cb[i] = a[r][c0];
                }
            }
            __syncthreads();
            // pivot among the rows not yet used (every wave computes it)
            T best = (T)-1;
            int bi = na;
            for (int i = lane; i < na; i += 64) {
                if (label[i] >= 0) continue;
                const T v = fabs(cb[i]);
                if (v > best) { best = v; bi = i; }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const T ov = __shfl_xor(best, off);
                const int oi = __shfl_xor(bi, off);
                if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
            }
            const int p = bi;
            if (ti == (p & 15)) {
                const int rp = p >> 4;
#pragma unroll
                for (int c = 0; c < R; c++) {
                    T v = (T)0;
#pragma unroll
                    for (int r = 0; r < R; r++) v = (r == rp) ? a[r][c] : v;
                    const int k = tj + 16 * c;
                    if (k < na) rb[k] = v;
                }
            }
            if (tid == 0) { label[p] = j; pof[j] = p; }
            __syncthreads();
            const T pv = cb[p];
            if (pv == (T)0 && !sing) sing = j + 1;
            const T pinv = pv == (T)0 ? (T)0 : (T)1 / pv;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int i = ti + 16 * r;
                const T f = cb[i < na ? i : 0];
#pragma unroll
                for (int c = 0; c < R; c++) {
                    const int k = tj + 16 * c;
                    const T prow = rb[k < na ? k : 0] * pinv;
                    const bool kj = (c == c0) && (tj == jj);
                    T v;
                    if (i == p) v = kj ? pinv : prow;
                    else v = kj ? -f * pinv : a[r][c] - f * prow;
                    a[r][c] = v;
                }
            }
        }
    }
    __syncthreads();
    return sing;
}

// LDS bytes of dinv_kernel: the double-buffered pivot column and row (192 each)
// and the pivot order and row labels.
size_t condense_lds_bytes(int elt)
{
    size_t head = (size_t)6 * 192 * elt + 192 * sizeof(int);
    return (head + 15) / 16 * 16;
}

// inv(D) of every element of [e0, e0+n) by Gauss-Jordan with partial pivoting
// (schurImplementation 0 semantics) in arithmetic T (D rounded to T on the way
// in), written column-major (ld na) to Dinv in FP64.  One workgroup per element.
// The mixed condensation runs it with T = float.
template <int R, typename T>
__global__ void __launch_bounds__(256)
dinv_kernel(int na, int e0, int n, const double *__restrict__ Dm, double *__restrict__ Dinv, int *__restrict__ info)
{
    extern __shared__ double lds[];
    T *colbuf = (T *)lds, *rowbuf = colbuf + 2 * 192;
    int *pof = (int *)(colbuf + 4 * 192), *label = pof + 192;
    const int tid = threadIdx.x, ti = tid & 15, tj = tid >> 4;
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const long e = e0 + t;
        const double *D = Dm + e * (long)na * na;
        T a[R][R];
#pragma unroll
        for (int c = 0; c < R; c++)
#pragma unroll
            for (int r = 0; r < R; r++) {
                int i = ti + 16 * r, k = tj + 16 * c;
                a[r][c] = (i < na && k < na) ? (T)D[i + (long)na * k] : (T)0;
            }
        int sing = gauss_jordan<R, T>(na, a, colbuf, rowbuf, pof, label);
        if (tid == 0 && info) info[e] = sing;
        double *Mi = Dinv + e * (long)na * na;
#pragma unroll
        for (int c = 0; c < R; c++)
#pragma unroll
            for (int r = 0; r < R; r++) {
                int i = ti + 16 * r, m = tj + 16 * c;
                if (i < na && m < na) Mi[label[i] + (long)na * pof[m]] = (double)a[r][c];
            }
        __syncthreads();
    }
}

// Maximum of a 64-bit unsigned value over the wave, returned to every lane.  DPP
// steps (quad exchanges, row shifts by 4 and 8, row broadcasts 15 and 31) leave
// the maximum in lane 63; lanes a step does not write keep their own value.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v)
{
#define HFEM_DPP_MAX(ctrl, rm, bm)                                                                           \
    {                                                                                                        \
        const int hi = (int)(v >> 32), lo = (int)(unsigned)v;                                                \
        const int h2 = __builtin_amdgcn_update_dpp(hi, hi, ctrl, rm, bm, false);                             \
        const int l2 = __builtin_amdgcn_update_dpp(lo, lo, ctrl, rm, bm, false);                             \
        const unsigned long long o = ((unsigned long long)(unsigned)h2 << 32) | (unsigned)l2;                \
        v = o > v ? o : v;                                                                                   \
    }
    HFEM_DPP_MAX(0xb1, 0xf, 0xf)    // quad_perm [1,0,3,2]
    HFEM_DPP_MAX(0x4e, 0xf, 0xf)    // quad_perm [2,3,0,1]
    HFEM_DPP_MAX(0x114, 0xf, 0xe)   // row_shr:4
    HFEM_DPP_MAX(0x118, 0xf, 0xc)   // row_shr:8
    HFEM_DPP_MAX(0x142, 0xa, 0xf)   // row_bcast:15
    HFEM_DPP_MAX(0x143, 0xc, 0xf)   // row_bcast:31
#undef HFEM_DPP_MAX
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), 63), lo = __builtin_amdgcn_readlane((int)(unsigned)v, 63);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

// ---------------------------------------------------------------------------
// inv(D) by blocked Gauss-Jordan on the FP64 MFMA.  Same pivoting and implicit
// row exchanges as gauss_jordan above (partial pivoting, lowest index on ties,
// rows never moved), so inv(D)[label[i], pof[c]] = G[i][c] holds on exit.
// The na x na matrix (padded to NN = 16 NT) lives in registers as 16 x 16 MFMA
// result tiles spread over the W waves (tile t = s W + wave in slot s; lane l,
// element v holds row l/16 + 4v, column l%16 of its tile).  Per 16-column panel:
//   1. wave 0 runs the 16 unblocked steps on the panel alone, in registers
//      (two or three rows per lane, no barrier); the panel then holds Gp, whose column k is T e_{P[k]} for the product T of
//      the panel's elementary steps and its pivot rows P;
//   2. every other column becomes T a = Abar + Gp a(P, :), with Abar = a with
//      the rows P zeroed: a rank-16 update done as four MFMAs per tile straight
//      into the register tiles (A operand Gp, B operand the old pivot rows,
//      both from LDS).
// The result is written column by column of inv(D) through an LDS staging tile
// so that the permutation (output column pof[c], row label^-1) is applied with
// coalesced stores.
// ---------------------------------------------------------------------------
template <int NT, int W, int MW>
__global__ void __launch_bounds__(64 * W, MW)
dinv_mfma_kernel(int na, int e0, int n, const double *__restrict__ D, double *__restrict__ Dinv, int *__restrict__ info)
{
    constexpr int NN = 16 * NT, NTILE = NT * NT, TPW = (NTILE + W - 1) / W, RPL = (NN + 63) / 64;
    constexpr int PS = 17;                // panel row stride: odd, so a column walk is bank-conflict free
    __shared__ double pan[NN * PS];       // panel / staging of 16 columns, row-major: pan[i*PS + k]
    __shared__ double Z[NN * PS];         // pivot rows of the panel, Z[k*NN + c]; output staging
    __shared__ int label[NN], pof[NN], rowk[NN];
    __shared__ int sing_s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lr = lane & 15, lk = lane >> 4;
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const long e = e0 + t;
        const double *De = D + e * (long)na * na;
        // D into the register tiles: all loads of a wave issued together
        f64x4 a[TPW];
        {
            int wv = __builtin_amdgcn_readfirstlane(wave);
            asm volatile("" : "+s"(wv));
#pragma unroll
            for (int s = 0; s < TPW; s++) {
                const int T = s * W + wv, r = T / NT, q = T % NT;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = r * 16 + lk + 4 * v, c = q * 16 + lr;
                    const bool ok = T < NTILE && i < na && c < na;
                    const double d = De[ok ? i + (long)na * c : 0];
                    a[s][v] = ok ? d : 0.0;
                }
            }
        }
        for (int i = tid; i < NN; i += 64 * W) label[i] = i < na ? -1 : NN;
        if (tid == 0) sing_s = 0;
        __syncthreads();
        for (int p = 0; p < NT; p++) {
            // 1. the panel to LDS.  (wv: the wave index made opaque at each step, so
            //    that per-tile addresses are formed where used, not held in registers
            //    across the panel loop)
            int wv = __builtin_amdgcn_readfirstlane(wave);
            asm volatile("" : "+s"(wv));
#pragma unroll
            for (int s = 0; s < TPW; s++) {
                const int T = s * W + wv, r = T / NT, q = T % NT;
                if (T < NTILE && q == p)
#pragma unroll
                    for (int v = 0; v < 4; v++) pan[(r * 16 + lk + 4 * v) * PS + lr] = a[s][v];
            }
            __syncthreads();
            // 2. unblocked steps on the panel (wave 0) in registers: lane l holds rows
            //    l, l+64, ...; the pivot row is broadcast by readlane, and the pivot is
            //    found from an exact wave maximum of the |value| bit patterns (DPP) with
            //    the lowest row among equal keys taken from a ballot
            if (wave == 0) {
                double x[RPL][16];
                bool used[RPL];
                int pk[RPL];
#pragma unroll
                for (int h = 0; h < RPL; h++) {
                    const int i = lane + 64 * h;
                    used[h] = i >= NN || label[i] >= 0;
                    pk[h] = -1;
#pragma unroll
                    for (int k = 0; k < 16; k++) x[h][k] = i < NN ? pan[i * PS + k] : 0.0;
                }
                int sing = 0;
#pragma unroll
                for (int jj = 0; jj < 16; jj++) {
                    // fully unrolled so that x[h][jj] is a register (no early exit)
                    const int j = p * 16 + jj;
                    if (j < na) {
                        // key 0 for used rows, 1 + |value| bits otherwise (NaN ranks highest)
                        unsigned long long key[RPL], m = 0;
#pragma unroll
                        for (int h = 0; h < RPL; h++) {
                            key[h] = used[h] ? 0ull
                                             : ((unsigned long long)__double_as_longlong(x[h][jj]) & 0x7fffffffffffffffull) + 1;
                            m = key[h] > m ? key[h] : m;
                        }
                        m = wave_max_u64(m);
                        int pr = NN;
#pragma unroll
                        for (int h = RPL - 1; h >= 0; h--) {
                            const unsigned long long b = __ballot(key[h] == m);
                            if (b) pr = 64 * h + __ffsll((long long)b) - 1;
                        }
                        const int owner = pr & 63, hp = pr >> 6;
                        double prow[16];
#pragma unroll
                        for (int k = 0; k < 16; k++) {
                            double v = x[0][k];
#pragma unroll
                            for (int h = 1; h < RPL; h++) v = hp == h ? x[h][k] : v;
                            const int lo = __builtin_amdgcn_readlane(__double2loint(v), owner);
                            const int hi = __builtin_amdgcn_readlane(__double2hiint(v), owner);
                            prow[k] = __hiloint2double(hi, lo);
                        }
                        const double pv = prow[jj];
                        if (pv == 0.0 && !sing) sing = j + 1;
                        const double pinv = pv == 0.0 ? 0.0 : 1.0 / pv;
                        // scaled pivot row s_k = a(p, k) / a(p, j); every row takes one FMA per
                        // column, then the pivot row (one lane) is overwritten
                        double sk[16];
#pragma unroll
                        for (int k = 0; k < 16; k++) sk[k] = prow[k] * pinv;
#pragma unroll
                        for (int h = 0; h < RPL; h++) {
                            const int i = lane + 64 * h;
                            const double f = x[h][jj];
#pragma unroll
                            for (int k = 0; k < 16; k++) x[h][k] = k == jj ? -f * pinv : fma(-f, sk[k], x[h][k]);
                            if (i == pr) {
#pragma unroll
                                for (int k = 0; k < 16; k++) x[h][k] = k == jj ? pinv : sk[k];
                                used[h] = true;
                                pk[h] = jj;
                            }
                        }
                        if (lane == 0) pof[j] = pr;
                    }
                }
#pragma unroll
                for (int h = 0; h < RPL; h++) {
                    const int i = lane + 64 * h;
                    if (i < NN) {
#pragma unroll
                        for (int k = 0; k < 16; k++) pan[i * PS + k] = x[h][k];
                        rowk[i] = pk[h];
                        if (pk[h] >= 0) label[i] = p * 16 + pk[h];
                    }
                }
                if (lane == 0 && sing && !sing_s) sing_s = sing;
            }
            __syncthreads();
            // 3. the old pivot rows of the other columns
            wv = __builtin_amdgcn_readfirstlane(wave);
            asm volatile("" : "+s"(wv));
#pragma unroll
            for (int s = 0; s < TPW; s++) {
                const int T = s * W + wv, r = T / NT, q = T % NT;
                if (T < NTILE && q != p)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int k = rowk[r * 16 + lk + 4 * v];
                        if (k >= 0) Z[k * NN + q * 16 + lr] = a[s][v];
                    }
            }
            __syncthreads();
            // 4. the rank-16 update (the panel columns take Gp)
            wv = __builtin_amdgcn_readfirstlane(wave);
            asm volatile("" : "+s"(wv));
#pragma unroll
            for (int s = 0; s < TPW; s++) {
                const int T = s * W + wv, r = T / NT, q = T % NT;
                if (T >= NTILE) continue;
                if (q == p) {
#pragma unroll
                    for (int v = 0; v < 4; v++) a[s][v] = pan[(r * 16 + lk + 4 * v) * PS + lr];
                } else {
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        if (rowk[r * 16 + lk + 4 * v] >= 0) a[s][v] = 0.0;
#pragma unroll
                    for (int ks = 0; ks < 4; ks++)
                        a[s] = mfma64(pan[(r * 16 + lr) * PS + ks * 4 + lk], Z[(ks * 4 + lk) * NN + q * 16 + lr], a[s]);
                }
            }
            __syncthreads();
        }
        // inv(D)(a, pof[c]) = G(pof[a], c), two 16-column tiles of G at a time
        // (staged in pan and Z)
        double *Oe = Dinv + e * (long)na * na;
        for (int q0 = 0; q0 < NT; q0 += 2) {
            int wv = __builtin_amdgcn_readfirstlane(wave);
            asm volatile("" : "+s"(wv));
#pragma unroll
            for (int s = 0; s < TPW; s++) {
                const int T = s * W + wv, r = T / NT, q = T % NT;
                if (T < NTILE && (q == q0 || q == q0 + 1)) {
                    double *st = q == q0 ? pan : Z;
#pragma unroll
                    for (int v = 0; v < 4; v++) st[(r * 16 + lk + 4 * v) * PS + lr] = a[s][v];
                }
            }
            __syncthreads();
            const int ncols = min(32, na - q0 * 16);
            for (int idx = tid; idx < ncols * na; idx += 64 * W) {
                const int cc = idx / na, ar = idx % na, c = q0 * 16 + cc;
                const double *st = cc < 16 ? pan : Z;
                Oe[ar + (long)na * pof[c]] = st[pof[ar] * PS + (cc & 15)];
            }
            __syncthreads();
        }
        if (tid == 0) info[e] = sing_s;
    }
}

int condense_rows_per_thread(int na) { return na <= 32 ? 2 : (na + 15) / 16; }

#define HFEM_R_CASES(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11)

hipError_t launch_dinv(int na, int e0, int n, const double *D, double *Dinv, int *info, hipStream_t s)
{
    const int NT = (na + 15) / 16;
    const int grid = std::min(n, 32768);
    switch (NT) {
// (tiles per side, waves, minimum waves per SIMD): two 8-wave workgroups per CU
// where the register tiles fit in 128 VGPRs
#define CASE(T, W, MW)                                                                                             \
    case T:                                                                                                        \
        hipLaunchKernelGGL((dinv_mfma_kernel<T, W, MW>), dim3(grid), dim3(64 * W), 0, s, na, e0, n, D, Dinv,      \
                           info);                                                                                  \
        break;
        CASE(1, 4, 2) CASE(2, 4, 2) CASE(3, 4, 2) CASE(4, 4, 2) CASE(5, 4, 2) CASE(6, 4, 2) CASE(7, 4, 2)
        CASE(8, 4, 2) CASE(9, 8, 2) CASE(10, 8, 2) CASE(11, 8, 2)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// inv(D) by the register-grid Gauss-Jordan in FP32, widened to FP64 (the mixed
// condensation's factorisation).
hipError_t launch_dinv_fp32(int na, int e0, int n, const double *D, double *Dinv, int *info, hipStream_t s)
{
    const int R = condense_rows_per_thread(na);
    const size_t lds = condense_lds_bytes((int)sizeof(float));
    const int grid = std::min(n, 32768);
    switch (R) {
#define CASE(RR)                                                                                                  \
    case RR:                                                                                                      \
        hipLaunchKernelGGL((dinv_kernel<RR, float>), dim3(grid), dim3(256), lds, s, na, e0, n, D, Dinv, info);     \
        break;
        HFEM_R_CASES(CASE)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
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
