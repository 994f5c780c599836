// Trace-system solver kernels (gfx950): inverses of BSR diagonal blocks, MDF
// block weights, BILU0 factorisation and its triangular sweeps one level of the
// ordering at a time, block-Jacobi, the vector kernels of GMRES, the solution
// update (getU) and DIRK stage sources.
//
// BSR blocks are bsz x bsz column-major, rows = faces, the diagonal block first in
// each row (the reference's Hg layout, assembly/globalassembly.cpp:42-86).
#include "hfem_device.h"

#include <algorithm>

namespace hfem {

constexpr int SV_THREADS = 256;
constexpr int ELIM_BATCH = 8;   // block-inverse entries a thread updates per batch

// A lane's value from lane l - s (DPP control CTRL, rows in ROWS); lanes without
// a source keep idv.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_take(double idv, double v)
{
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(idv), __double2hiint(v), CTRL, ROWS, 0xf, false);
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(idv), __double2loint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_take(float idv, float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(idv), __float_as_int(v), CTRL, ROWS, 0xf, false));
}

// One step of the pivot choice's scan: (b, i) becomes the better of itself and
// the pair DPP brings in -- the larger value, the smaller row on ties -- with
// (-1, k) where no pair comes in (the search's starting pair: no row of column k
// has a magnitude below 0, and a NaN never replaces it).
template <int CTRL, int ROWS, typename T>
__device__ __forceinline__ void argmax_step(T &b, int &i, int k)
{
    const T yb = dpp_take<CTRL, ROWS>((T)-1, b);
    const int yi = __builtin_amdgcn_update_dpp(k, i, CTRL, ROWS, 0xf, false);
    const bool t = yb > b || (yb == b && yi < i);
    b = t ? yb : b;
    i = t ? yi : i;
}

// In-place Gauss-Jordan inverse of an n x n column-major block in LDS with
// partial pivoting: rows are interchanged as pivots are chosen and the columns of
// the result unscrambled in reverse order at the end.  All threads of the
// workgroup call it; piv (n ints), rowk/colk (n values each) and red (1 value)
// are LDS; T is the arithmetic (double, or float for FP32 preconditioner factors).
// LATENCY picks the elimination for an inverse on a dependency chain (batched
// reads, no branches per entry: the BILU0 factorisations) over the one for many
// independent blocks per CU (one entry at a time, fewer registers: the block-Jacobi
// and MDF inverses); both compute the same values.
// Returns 0, or 1 + the step of the first zero pivot (workgroup-uniform).
template <typename T, bool LATENCY = true>
__device__ int lds_inverse(T *A, int n, int *piv, T *rowk, T *colk, T *red)
{
    const int tid = threadIdx.x, nth = blockDim.x;
    // idx / n as a multiply-high by ceil(2^32 / n): exact for every idx < n * n
    // with n <= 128 (checked exhaustively); n = 1 gives 0, and idx is 0 there
    const unsigned ndiv = (unsigned)((0x100000000ull + (unsigned)n - 1) / (unsigned)n);
    for (int k = 0; k < n; k++) {
        if (tid < 64) {
            T best = (T)-1;
            int bi = k;
            for (int i = k + tid; i < n; i += 64) {
                const T v = fabs(A[i + n * k]);
                if (v > best) { best = v; bi = i; }
            }
            // inclusive scan over the wave by DPP: shifts by 1, 2, 4, 8 within each
            // row of 16 lanes, then row 0's total into row 1 and row 2's into row 3,
            // then lanes 0-31's total into rows 2 and 3: lane 63 ends with the choice
            argmax_step<0x111, 0xf>(best, bi, k);   // row_shr:1
            argmax_step<0x112, 0xf>(best, bi, k);   // row_shr:2
            argmax_step<0x114, 0xf>(best, bi, k);   // row_shr:4
            argmax_step<0x118, 0xf>(best, bi, k);   // row_shr:8
            argmax_step<0x142, 0xa>(best, bi, k);   // row_bcast:15
            argmax_step<0x143, 0xc>(best, bi, k);   // row_bcast:31
            if (tid == 63) { piv[k] = bi; red[0] = best; }
        }
        __syncthreads();
        if (red[0] == (T)0) return k + 1;
        const int p = piv[k];
        for (int j = tid; j < n; j += nth) {
            T a = A[k + n * j], b = A[p + n * j];
            rowk[j] = b;                         // pivot row after the interchange
            if (p != k) { A[k + n * j] = b; A[p + n * j] = a; }
        }
        __syncthreads();
        for (int i = tid; i < n; i += nth) colk[i] = A[i + n * k];
        __syncthreads();
        const T pinv = (T)1 / rowk[k];
        if (!LATENCY) {
            // many blocks per CU (throughput): one entry at a time, fewer registers
            for (int idx = tid; idx < n * n; idx += nth) {
                const int j = idx / n, i = idx - j * n;
                T v;
                if (i == k) v = (j == k) ? pinv : rowk[j] * pinv;
                else if (j == k) v = -colk[i] * pinv;
                else v = A[idx] - colk[i] * (rowk[j] * pinv);
                A[idx] = v;
            }
        }
        // a thread's entries idx = tid + u * nth in batches of ELIM_BATCH: every read
        // of a batch is issued before its first write, so the LDS latency is paid
        // once per batch rather than once per entry
        for (int base = tid; LATENCY && base < n * n; base += ELIM_BATCH * nth) {
            T a[ELIM_BATCH], c[ELIM_BATCH], rw[ELIM_BATCH];
            bool rk[ELIM_BATCH], ck[ELIM_BATCH];
#pragma unroll
            for (int u = 0; u < ELIM_BATCH; u++) {
                // (an entry past the block reads entry 0's operands and is not stored)
                const int idx = base + u * nth < n * n ? base + u * nth : 0;
                const int j = (int)__umulhi((unsigned)idx, ndiv), i = idx - j * n;
                rk[u] = i == k;
                ck[u] = j == k;
                a[u] = A[idx];
                c[u] = colk[i];
                rw[u] = rowk[j];
            }
            // the three cases computed alike and selected (no branches per entry)
#pragma unroll
            for (int u = 0; u < ELIM_BATCH; u++) {
                const int idx = base + u * nth;
                const T rp = rw[u] * pinv;
                const T other = a[u] - c[u] * rp;
                const T colv = -c[u] * pinv;
                const T v = rk[u] ? (ck[u] ? pinv : rp) : (ck[u] ? colv : other);
                if (idx < n * n) A[idx] = v;
            }
        }
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; k--) {
        const int p = piv[k];
        if (p != k)
            for (int i = tid; i < n; i += nth) {
                T t = A[i + n * k];
                A[i + n * k] = A[i + n * p];
                A[i + n * p] = t;
            }
        __syncthreads();
    }
    return 0;
}

template <typename T> struct LdsInv {
    T *rowk, *colk, *red;
    int *piv;
};

template <typename T>
__device__ LdsInv<T> lds_inv_area(T *p, int n)
{
    LdsInv<T> w;
    w.rowk = p;
    w.colk = p + n;
    w.red = p + 2 * n;
    w.piv = (int *)(p + 2 * n + 2);
    return w;
}

size_t lds_inv_bytes(int n) { return sizeof(double) * (2 * n + 2) + sizeof(int) * (n + 2); }

// Inverses of the diagonal blocks of Hg (block-Jacobi, mkBlockJacobi.m:9-11, and
// the MDF weights).  One workgroup per row.
__global__ void __launch_bounds__(SV_THREADS)
block_inverse_kernel(int nrows, int bsz, const int *__restrict__ rowpts, const double *__restrict__ Hg,
                     double *__restrict__ Dinv, int *__restrict__ info)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x, nth = blockDim.x;
    const long b2 = (long)bsz * bsz;
    double *A = lds;
    LdsInv<double> w = lds_inv_area(A + b2, bsz);
    for (int row = blockIdx.x; row < nrows; row += gridDim.x) {
        const double *src = Hg + (long)rowpts[row] * b2;
        for (long i = tid; i < b2; i += nth) A[i] = src[i];
        __syncthreads();
        const int r = lds_inverse<double, false>(A, bsz, w.piv, w.rowk, w.colk, w.red);
        for (long i = tid; i < b2; i += nth) Dinv[row * b2 + i] = A[i];
        if (tid == 0 && info) info[row] = r;
        __syncthreads();
    }
}

// MDF block weights C(n) = ||inv(H_ii) H_n||_F for every block n of row i
// (kernel/computeMDFordering.m:14-20).  One workgroup per row.
__global__ void __launch_bounds__(SV_THREADS)
mdf_weight_kernel(int nrows, int bsz, const int *__restrict__ rowpts, const double *__restrict__ Hg,
                  const double *__restrict__ Dinv, double *__restrict__ Cw)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x, nth = blockDim.x;
    const long b2 = (long)bsz * bsz;
    double *A = lds, *B = lds + b2, *red = lds + 2 * b2;
    for (int row = blockIdx.x; row < nrows; row += gridDim.x) {
        for (long i = tid; i < b2; i += nth) A[i] = Dinv[row * b2 + i];
        for (int n = rowpts[row]; n < rowpts[row + 1]; n++) {
            for (long i = tid; i < b2; i += nth) B[i] = Hg[(long)n * b2 + i];
            __syncthreads();
            double s2 = 0.0;
            for (long idx = tid; idx < b2; idx += nth) {
                const int i = idx % bsz, j = idx / bsz;
                double s = 0.0;
                for (int k = 0; k < bsz; k++) s += A[i + bsz * k] * B[k + bsz * j];
                s2 += s * s;
            }
            red[tid] = s2;
            __syncthreads();
            for (int st = nth / 2; st > 0; st >>= 1) {
                if (tid < st) red[tid] += red[tid + st];
                __syncthreads();
            }
            if (tid == 0) Cw[n] = sqrt(red[0]);
            __syncthreads();
        }
    }
}

// BILU0 factorisation of the rows of one level (kernel/computeBILU0.m), in the
// row-oriented form: row ri takes, for each neighbour rk numbered before it in
// increasing order, L_ik = A_ik inv(U_kk), then A_ii -= L_ik A_ki and A_il -=
// L_ik A_kl for the later common neighbours rl; finally A_ii is inverted.  Each
// block of row ri receives the same updates in the same order as in the
// reference's right-looking loop; rows of one level do not depend on each other.
// The column lists of row ri and of the current neighbour row sit in LDS, so the
// neighbour order and the target search cost no memory round trips; blocks are
// staged into LDS with eight loads in flight per thread and multiplied on the
// FP64 MFMA.  The blocks of row ri that are read again (the targets and the
// diagonal) are always read and written with the fixed entry -> thread map of
// wg_load_block / wg_subtract_block, so a thread only ever re-reads global
// entries it wrote itself; L_ik, stored from the MFMA result layout, is not read
// again by the row.
constexpr int FACTOR_MAXROW = 64;     // longest BSR row the factorisation takes

typedef double f64x4 __attribute__((ext_vector_type(4)));

// The n x n product C = A B of two column-major LDS blocks on the MFMA (FP64, or
// FP32 for FP32 preconditioner factors): the 16 x 16 tiles of C are spread over
// the waves, rows and columns past n are masked to zero; out(i, j, c) receives
// each entry of C (i, j < n).  MFMA 16x16x4 operands: lane l supplies A(l%16,
// l/16) and B(l/16, l%16); result lane l, element v holds C(l/16 + 4v, l%16)
// (FP64) or C(4(l/16) + v, l%16) (FP32).
template <typename T> struct BlkMfma;
template <> struct BlkMfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 run(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};
template <> struct BlkMfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 run(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};

template <typename T, typename OUT>
__device__ __forceinline__ void lds_block_product(const T *A, const T *B, int n, OUT out)
{
    typedef BlkMfma<T> M;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nw = blockDim.x >> 6;
    const int lr = lane & 15, lk = lane >> 4, nt = (n + 15) / 16, ntt = nt * nt;
    // four of the wave's tiles at a time (t = tb + nw u), their MFMA chains
    // interleaved; each tile's k order is that of one tile alone
    constexpr int U = 4;
    for (int tb = wave; tb < ntt; tb += U * nw) {
        int ia[U], jb[U], i0[U], j0[U];
        bool va[U], vb[U], tv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t0 = tb + nw * u;
            tv[u] = t0 < ntt;
            const int t = tv[u] ? t0 : tb;
            i0[u] = (t % nt) * 16;
            j0[u] = (t / nt) * 16;
            ia[u] = min(i0[u] + lr, n - 1);
            jb[u] = min(j0[u] + lr, n - 1);
            va[u] = i0[u] + lr < n;
            vb[u] = j0[u] + lr < n;
        }
        typename M::v4 acc[U];
#pragma unroll
        for (int u = 0; u < U; u++) acc[u] = (typename M::v4){(T)0, (T)0, (T)0, (T)0};
        for (int k0 = 0; k0 < n; k0 += 4) {
            const int k = k0 + lk, kc = min(k, n - 1);
            T a[U], b[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                a[u] = (va[u] && k < n) ? A[ia[u] + n * kc] : (T)0;
                b[u] = (vb[u] && k < n) ? B[kc + n * jb[u]] : (T)0;
            }
#pragma unroll
            for (int u = 0; u < U; u++) acc[u] = M::run(a[u], b[u], acc[u]);
        }
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = i0[u] + M::row(lk, v), j = j0[u] + lr;
                if (tv[u] && i < n && j < n) out(i, j, acc[u][v]);
            }
    }
}

// dst[i] = src[i] for i < n (n >= 1) by the workgroup, eight loads in flight per
// thread (addresses clamped into the block).  Entry i is always handled by thread
// i % blockDim.x.
template <typename T>
__device__ __forceinline__ void wg_load_block(T *dst, const T *__restrict__ src, long n)
{
    const int tid = threadIdx.x, nth = blockDim.x;
    for (long b = tid; b < n; b += 8L * nth) {
        T v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = src[min(b + (long)j * nth, n - 1)];
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (b + (long)j * nth < n) dst[b + (long)j * nth] = v[j];
    }
}

// dst[i] -= d[i] for i < n, global dst, LDS d; same entry -> thread map as
// wg_load_block.
template <typename T>
__device__ __forceinline__ void wg_subtract_block(T *dst, const T *d, long n)
{
    const int tid = threadIdx.x, nth = blockDim.x;
    for (long b = tid; b < n; b += 8L * nth) {
        T v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = dst[min(b + (long)j * nth, n - 1)];
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (b + (long)j * nth < n) dst[b + (long)j * nth] = v[j] - d[b + (long)j * nth];
    }
}

// T: the factors' type (double, or float: the reference's single-precision
// preconditioner, precPrecision 0, preprocessing/initializestructs.cpp:476-486,
// factored here on the FP32 MFMA).
template <typename T>
__global__ void __launch_bounds__(SV_THREADS)
bilu0_factor_kernel(int bsz, const int *__restrict__ rowpts, const int *__restrict__ colind,
                    const int *__restrict__ u2o, const int *__restrict__ rows, int nlev, T *__restrict__ Mg,
                    int *__restrict__ info)
{
    extern __shared__ double lds_raw[];
    T *lds = (T *)lds_raw;
    const int tid = threadIdx.x, nth = blockDim.x;
    const long b2 = (long)bsz * bsz;
    T *X = lds, *B = lds + b2, *L = lds + 2 * b2;
    LdsInv<T> w = lds_inv_area(lds + 3 * b2, bsz);
    // row ri: columns and their order numbers; the lower neighbours' positions by
    // increasing order number; the current neighbour row's columns
    int *rcol = w.piv + bsz + 2, *rord = rcol + FACTOR_MAXROW, *lst = rord + FACTOR_MAXROW,
        *ncol = lst + FACTOR_MAXROW + 1;
    for (int t = blockIdx.x; t < nlev; t += gridDim.x) {
        const int ri = rows[t], oi = u2o[ri], d0 = rowpts[ri], len = rowpts[ri + 1] - d0;
        if (len > FACTOR_MAXROW) {
            // workgroup-uniform: reported through info, the row is left as it is
            if (tid == 0) info[ri] = -len;
            continue;
        }
        if (tid < len) {
            const int c = colind[d0 + tid];
            rcol[tid] = c;
            rord[tid] = u2o[c];
        }
        __syncthreads();
        if (tid == 0) {
            int cnt = 0;
            for (int q = 1; q < len; q++) {
                if (rord[q] >= oi) continue;
                int m = cnt++;                    // insertion by increasing order number
                while (m > 0 && rord[lst[1 + m - 1]] > rord[q]) {
                    lst[1 + m] = lst[1 + m - 1];
                    m--;
                }
                lst[1 + m] = q;
            }
            lst[0] = cnt;
        }
        __syncthreads();
        const int cnt = lst[0];
        for (int qq = 0; qq < cnt; qq++) {
            const int q = lst[1 + qq], ik = d0 + q, rk = rcol[q], ork = rord[q];
            const int k0 = rowpts[rk], klen = rowpts[rk + 1] - k0;
            if (tid < klen && tid < FACTOR_MAXROW) ncol[tid] = colind[k0 + tid];
            wg_load_block(X, Mg + (long)ik * b2, b2);
            wg_load_block(B, Mg + (long)k0 * b2, b2);     // inv(U_kk)
            __syncthreads();
            // L_ik = A_ik inv(U_kk), kept in LDS and stored in place of A_ik
            T *Mik = Mg + (long)ik * b2;
            lds_block_product(X, B, bsz, [&](int i, int j, T c) {
                L[i + bsz * j] = c;
                Mik[i + bsz * j] = c;
            });
            __syncthreads();
            // targets in row ri: the diagonal (source: ri's position in row rk) and the
            // later common neighbours rl (source: rl's position in row rk)
            for (int l = 0; l < len; l++) {
                if (l != 0 && rord[l] <= ork) continue;
                const int rl = rcol[l];
                int km = -1;
                for (int m = 1; m < klen && m < FACTOR_MAXROW; m++)
                    if (ncol[m] == rl) { km = m; break; }
                if (km < 0) continue;
                wg_load_block(B, Mg + (long)(k0 + km) * b2, b2);
                __syncthreads();
                lds_block_product(L, B, bsz, [&](int i, int j, T c) { X[i + bsz * j] = c; });
                __syncthreads();
                wg_subtract_block(Mg + (long)(d0 + l) * b2, X, b2);
                __syncthreads();
            }
            // targets skipped above run no barrier: every wave must be done scanning
            // ncol before the next neighbour's column list overwrites it
            __syncthreads();
        }
        wg_load_block(X, Mg + (long)d0 * b2, b2);
        __syncthreads();
        const int r = lds_inverse(X, bsz, w.piv, w.rowk, w.colk, w.red);
        for (long i = tid; i < b2; i += nth) Mg[(long)d0 * b2 + i] = X[i];
        if (tid == 0) info[ri] = r;
        __syncthreads();
    }
}

// Broadcast lane `src` (wave-uniform) of a double to the whole wave.
__device__ __forceinline__ double wave_bcast(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// acc_i += sum_cc B[i + bsz cc] z[cc] for the lane's rows i = lane, lane + 64, with z
// held one entry per lane (z0 = z[lane], z1 = z[lane + 64]) and broadcast by readlane.
// The columns go in chunks of CH whose loads are all issued before the first
// multiply (addresses clamped into the block; columns past bsz are multiplied by
// zero), so a block costs ceil(bsz / CH) memory round trips instead of one per
// column pair.
template <int CH, bool TWO, typename TM>
__device__ __forceinline__ void wave_block_matvec_ch(int bsz, const TM *__restrict__ B, double z0, double z1,
                                                     double &acc0, double &acc1)
{
    const int lane = threadIdx.x & 63;
    const int r0 = lane < bsz ? lane : 0, r1 = lane + 64 < bsz ? lane + 64 : 0;
    double s0 = 0.0, s1 = 0.0;
    for (int c0 = 0; c0 < bsz; c0 += CH) {
        double v0[CH], v1[CH];
#pragma unroll
        for (int j = 0; j < CH; j++) {
            const long col = (long)bsz * min(c0 + j, bsz - 1);
            v0[j] = B[r0 + col];
            if (TWO) v1[j] = B[r1 + col];
        }
#pragma unroll
        for (int j = 0; j < CH; j++) {
            const int cc = c0 + j;
            double zj = wave_bcast(cc < 64 ? z0 : z1, cc & 63);
            zj = cc < bsz ? zj : 0.0;
            s0 += v0[j] * zj;
            if (TWO) s1 += v1[j] * zj;
        }
    }
    if (lane < bsz) acc0 += s0;
    if (TWO && lane + 64 < bsz) acc1 += s1;
}

template <typename TM>
__device__ __forceinline__ void wave_block_matvec(int bsz, const TM *__restrict__ B, double z0, double z1,
                                                  double &acc0, double &acc1)
{
    if (bsz > 64) wave_block_matvec_ch<16, true>(bsz, B, z0, z1, acc0, acc1);
    else wave_block_matvec_ch<64, false>(bsz, B, z0, z1, acc0, acc1);
}

// One BILU0 sweep over the rows of one level (kernel/applyBILU0.m).  For row r,
// t = b_r - sum over the neighbours k taken by the sweep (numbered before r in the
// forward sweep L y = x, after r in the backward sweep U v = y) of M_rk z_k; the
// forward sweep stores z_r = t, the backward sweep z_r = inv(U_rr) t.  One
// 256-thread workgroup per row: the four waves take the neighbour blocks in turn
// (lane = block row, the neighbour vector broadcast from registers), their partial
// sums meet in LDS.
// The backward sweep's inv(U_rr) b_r, this wave's share: wave w takes the columns
// cc = w, w + NW, ... of inv(U_rr) (preloaded in dv for blocks up to 64 wide, else
// read from the block in chunks of 8 columns), b_r staged in LDS (tb); rows lane
// (a0) and lane + 64 (a1).  Shared by both sweep kernels so that they compute the
// same bits.
template <bool SMALL, int ND, int NW, typename TM>
__device__ __forceinline__ void inv_diag_share(int bsz, int wv, int lane, int r0, int r1, const TM *Ui,
                                               const double (&dv)[ND], const double *tb, double &a0, double &a1)
{
    a0 = 0.0;
    a1 = 0.0;
    if (SMALL) {
#pragma unroll
        for (int j = 0; j < ND; j++) {
            const int cc = wv + NW * j;
            a0 += dv[j] * (cc < bsz ? tb[min(cc, bsz - 1)] : 0.0);
        }
    } else {
        // (loads of 8 columns issued together, addresses clamped into the block)
        for (int c0 = wv; c0 < bsz; c0 += 8 * NW) {
            double v0[8], v1[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const long col = (long)bsz * min(c0 + j * NW, bsz - 1);
                v0[j] = Ui[r0 + col];
                v1[j] = Ui[r1 + col];
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int cc = c0 + j * NW;
                const double tc = cc < bsz ? tb[min(cc, bsz - 1)] : 0.0;
                a0 += v0[j] * tc;
                a1 += v1[j] * tc;
            }
        }
    }
    if (lane >= bsz) a0 = 0.0;
    if (lane + 64 >= bsz) a1 = 0.0;
}

// Backward sweeps read upper blocks scaled by the factorisation's last step
// (bilu0_scale_upper_kernel): U~_rl = inv(U_rr) U_rl, so that
//   z_r = inv(U_rr) b_r - sum_l U~_rl z_l
// and inv(U_rr) b_r is formed before the neighbours' z are needed.
template <bool BACKWARD, bool SMALL, typename TM>
__global__ void __launch_bounds__(SV_THREADS)
bilu0_sweep_kernel(int bsz, const int *__restrict__ rowpts, const int *__restrict__ colind,
                   const int *__restrict__ u2o, const int *__restrict__ rows, int nlev, const TM *__restrict__ Mg,
                   const double *__restrict__ b, double *__restrict__ z)
{
    constexpr int NW = SV_THREADS / 64;
    __shared__ double red[NW][128];
    __shared__ double tb[128];
    __shared__ int scol[FACTOR_MAXROW], stake[FACTOR_MAXROW];
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const long b2 = (long)bsz * bsz;
    const int r0 = lane < bsz ? lane : 0, r1 = lane + 64 < bsz ? lane + 64 : 0;
    for (int t = blockIdx.x; t < nlev; t += gridDim.x) {
        const int r = rows[t], o = u2o[r], k0 = rowpts[r], len = min(rowpts[r + 1] - k0, FACTOR_MAXROW);
        const TM *Ui = Mg + (long)k0 * b2;
        // the loads that depend on nothing else go out together: the row's column
        // list (rows have at most FACTOR_MAXROW blocks: the factorisation rejects
        // longer ones), its right-hand side and, backward, this wave's columns
        // wv + NW j of inv(U_rr)
        if (tid < len) {
            const int c = colind[k0 + tid];
            scol[tid] = c;
            stake[tid] = BACKWARD ? (u2o[c] > o) : (u2o[c] < o);
        }
        const double bv = tid < bsz ? b[(long)r * bsz + tid] : 0.0;
        if (BACKWARD && tid < bsz) tb[tid] = bv;
        constexpr int ND = (SMALL && BACKWARD) ? 64 / NW : 1;
        double dv[ND];
        if (SMALL && BACKWARD) {
#pragma unroll
            for (int j = 0; j < ND; j++) dv[j] = Ui[r0 + (long)bsz * min(wv + NW * j, bsz - 1)];
        }
        __syncthreads();
        double p0 = 0.0, p1 = 0.0;
        if (BACKWARD) inv_diag_share<SMALL, ND, NW>(bsz, wv, lane, r0, r1, Ui, dv, tb, p0, p1);
        double acc0 = 0.0, acc1 = 0.0;
        int sel = 0;
        for (int q = 1; q < len; q++) {
            if (!stake[q]) continue;
            if ((sel++ % NW) != wv) continue;
            const int c = scol[q];
            const double *zc = z + (long)c * bsz;
            const double z0 = lane < bsz ? zc[lane] : 0.0, z1 = lane + 64 < bsz ? zc[lane + 64] : 0.0;
            if (SMALL) wave_block_matvec_ch<64, false>(bsz, Mg + (long)(k0 + q) * b2, z0, z1, acc0, acc1);
            else wave_block_matvec_ch<16, true>(bsz, Mg + (long)(k0 + q) * b2, z0, z1, acc0, acc1);
        }
        red[wv][lane] = BACKWARD ? p0 - acc0 : acc0;
        red[wv][lane + 64] = BACKWARD ? p1 - acc1 : acc1;
        __syncthreads();
        if (tid < bsz) {
            double sum = 0.0;
#pragma unroll
            for (int w = 0; w < NW; w++) sum += red[w][tid];
            z[(long)r * bsz + tid] = BACKWARD ? sum : bv - sum;
        }
        __syncthreads();
    }
}

// Wait until *f holds `epoch`.  The flags and the vector entries they guard are
// read and written only by agent-scope atomics (coherent across the XCDs without
// whole-L2 write-backs or invalidations, which an acquire / release pair costs
// here per row).  At most spin_limit polls; a wait that runs out, or any wait once
// another has (*err set), sets *err and returns.
__device__ __forceinline__ void wait_ready(const unsigned *f, unsigned epoch, int spin_limit, int *err)
{
    // (the polled values go through readfirstlane: a uniform loop)
    int spins = 0;
    while ((unsigned)__builtin_amdgcn_readfirstlane(
               __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != epoch) {
        if (++spins > spin_limit ||
            __builtin_amdgcn_readfirstlane(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// One BILU0 sweep without level launches (kernel/applyBILU0.m), with the per-row
// arithmetic of bilu0_sweep_kernel (the same bits).  The rows go in `order` -- the
// level lists concatenated, so every row comes after the rows it reads -- to
// persistent workgroups through an atomic ticket.  A wave issues the loads of a
// neighbour block (bsz <= 64: all 64 columns into registers) before it waits for
// that neighbour's ready flag, so the block's latency overlaps the wait.  Vector
// entries written in the sweep are stored and read by agent-scope atomics; the row's
// flag is stored once those stores have completed (workgroup fence, then barrier).
// Flags hold the
// epoch of the current application (sweep_reset_kernel bumps it), so they are never
// cleared.  A wait that gives up (wait_ready) leaves a wrong result and *err set,
// which the host reports; every workgroup still runs to the end of the tickets.
// No deadlock: a workgroup waits only on rows of smaller tickets, which were drawn
// by workgroups already running.
template <bool BACKWARD, bool SMALL, typename TM>
__global__ void __launch_bounds__(SV_THREADS)
bilu0_sweep_syncfree_kernel(int bsz, const int *__restrict__ rowpts, const int *__restrict__ colind,
                            const int *__restrict__ u2o, const int *__restrict__ order, int n,
                            const TM *__restrict__ Mg, const double *__restrict__ b, double *z, unsigned *flags,
                            const unsigned *epoch_p, unsigned *counter, int spin_limit, int *err)
{
    constexpr int NW = SV_THREADS / 64;
    __shared__ double red[NW][128];
    __shared__ double tb[128];
    __shared__ int scol[FACTOR_MAXROW], stake[FACTOR_MAXROW];
    __shared__ int sticket;
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const long b2 = (long)bsz * bsz;
    const int r0 = lane < bsz ? lane : 0, r1 = lane + 64 < bsz ? lane + 64 : 0;
    const unsigned epoch = __hip_atomic_load(epoch_p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // No lane-conditional code may sit between the last barrier of one row and the
    // first of the next: with the flag store and the ticket draw both under
    // "tid == 0", the structurizer merged them across the back edge into an inner
    // loop that only lane 0 leaves, so the other lanes re-ran the old row forever.
    // Both are therefore done by the whole of wave 0 (a wave-uniform condition):
    // the draw adds 1 from lane 0 and 0 from the others, and lane 0's old value is
    // the ticket; every lane stores the same flag.
    while (true) {
        if (wv == 0) sticket = __builtin_amdgcn_readfirstlane(atomicAdd(counter, lane == 0 ? 1u : 0u));
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(sticket);
        if (t >= n) break;
        const int r = order[t], o = u2o[r], k0 = rowpts[r], len = min(rowpts[r + 1] - k0, FACTOR_MAXROW);
        const TM *Ui = Mg + (long)k0 * b2;
        if (tid < len) {
            const int c = colind[k0 + tid];
            scol[tid] = c;
            stake[tid] = BACKWARD ? (u2o[c] > o) : (u2o[c] < o);
        }
        // the right-hand side of entries lane and lane + 64, for wave 0, which
        // stores the row (so that its own wait covers every store of the row)
        const double bv0 = wv == 0 && lane < bsz ? b[(long)r * bsz + lane] : 0.0;
        const double bv1 = wv == 0 && lane + 64 < bsz ? b[(long)r * bsz + lane + 64] : 0.0;
        if (BACKWARD && tid < bsz) tb[tid] = b[(long)r * bsz + tid];
        constexpr int ND = (SMALL && BACKWARD) ? 64 / NW : 1;
        double dv[ND];
        if (SMALL && BACKWARD) {
#pragma unroll
            for (int j = 0; j < ND; j++) dv[j] = Ui[r0 + (long)bsz * min(wv + NW * j, bsz - 1)];
        }
        __syncthreads();
        // backward: inv(U_rr) b_r before any wait on a neighbour
        double p0 = 0.0, p1 = 0.0;
        if (BACKWARD) inv_diag_share<SMALL, ND, NW>(bsz, wv, lane, r0, r1, Ui, dv, tb, p0, p1);
        double acc0 = 0.0, acc1 = 0.0;
        int sel = 0;
        for (int q = 1; q < len; q++) {
            if (!__builtin_amdgcn_readfirstlane(stake[q])) continue;
            if ((sel++ % NW) != wv) continue;
            const int c = __builtin_amdgcn_readfirstlane(scol[q]);
            const TM *B = Mg + (long)(k0 + q) * b2;
            const double *zc = z + (long)c * bsz;
            if (SMALL) {
                double v0[64];
#pragma unroll
                for (int j = 0; j < 64; j++) v0[j] = B[r0 + (long)bsz * min(j, bsz - 1)];
                wait_ready(flags + c, epoch, spin_limit, err);
                const double z0 = lane < bsz ? __hip_atomic_load(zc + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                             : 0.0;
                double s0 = 0.0;
#pragma unroll
                for (int j = 0; j < 64; j++) {
                    double zj = wave_bcast(z0, j);
                    zj = j < bsz ? zj : 0.0;
                    s0 += v0[j] * zj;
                }
                if (lane < bsz) acc0 += s0;
            } else {
                wait_ready(flags + c, epoch, spin_limit, err);
                const double z0 =
                    lane < bsz ? __hip_atomic_load(zc + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
                const double z1 =
                    lane + 64 < bsz ? __hip_atomic_load(zc + lane + 64, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                    : 0.0;
                wave_block_matvec_ch<16, true>(bsz, B, z0, z1, acc0, acc1);
            }
        }
        red[wv][lane] = BACKWARD ? p0 - acc0 : acc0;
        red[wv][lane + 64] = BACKWARD ? p1 - acc1 : acc1;
        __syncthreads();
        if (wv == 0) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int i = lane + 64 * h;
                if (i < bsz) {
                    double sum = 0.0;
#pragma unroll
                    for (int w = 0; w < NW; w++) sum += red[w][i];
                    const double v = BACKWARD ? sum : (h ? bv1 : bv0) - sum;
                    __hip_atomic_store(z + (long)r * bsz + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        // publish row r: wave 0 stored it; once its stores have completed (vmcnt 0:
        // device-scope stores are acknowledged at the coherence point) the flag goes
        // out.  (The barrier also frees red, tb and the column lists for the next row.)
        if (wv == 0) __builtin_amdgcn_s_waitcnt(0x0F70);   // s_waitcnt vmcnt(0)
        __syncthreads();
        if (wv == 0) __hip_atomic_store(flags + r, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// A new epoch for the flags of the next application and fresh tickets for its two
// sweeps (a kernel, not a memset, so that the application can be captured in a graph).
// The factorisation's last step for the sweeps: every upper block of row r becomes
// U~_rl = inv(U_rr) U_rl (inv(U_rr) is the row's diagonal block), in place, after
// the whole factorisation (the other rows' eliminations read the unscaled U_rl).
// One workgroup per row: inv(U_rr) in LDS (FP64), the upper blocks a few columns
// at a time, each thread one entry of inv(U_rr) times a staged column.
template <typename TM>
__global__ void __launch_bounds__(SV_THREADS)
bilu0_scale_upper_kernel(int bsz, int n, const int *__restrict__ rowpts, const int *__restrict__ colind,
                         const int *__restrict__ u2o, TM *__restrict__ Mg)
{
    extern __shared__ double su[];
    const long b2 = (long)bsz * bsz;
    const int tid = threadIdx.x, cpp = max(1, SV_THREADS / bsz);
    double *Dl = su, *Cl = su + b2;
    for (int r = blockIdx.x; r < n; r += gridDim.x) {
        // (row bounds and the upper test through readfirstlane: loop control with
        // barriers inside must be uniform)
        const int k0 = __builtin_amdgcn_readfirstlane(rowpts[r]);
        const int len = __builtin_amdgcn_readfirstlane(rowpts[r + 1]) - k0;
        const int o = __builtin_amdgcn_readfirstlane(u2o[r]);
        const TM *Ui = Mg + (long)k0 * b2;
        for (long i = tid; i < b2; i += SV_THREADS) Dl[i] = (double)Ui[i];
        __syncthreads();
        for (int q = 1; q < len; q++) {
            if (__builtin_amdgcn_readfirstlane(u2o[colind[k0 + q]]) <= o) continue;
            TM *B = Mg + (long)(k0 + q) * b2;
            for (int c0 = 0; c0 < bsz; c0 += cpp) {
                const int nc = min(cpp, bsz - c0);
                for (int i = tid; i < nc * bsz; i += SV_THREADS) Cl[i] = (double)B[(long)c0 * bsz + i];
                __syncthreads();
                if (tid < nc * bsz) {
                    const int c = tid / bsz, i = tid - c * bsz;
                    double acc = 0.0;
                    for (int k = 0; k < bsz; k++) acc += Dl[i + (long)bsz * k] * Cl[c * bsz + k];
                    B[(long)(c0 + c) * bsz + i] = (TM)acc;
                }
                __syncthreads();
            }
        }
    }
}

__global__ void sweep_reset_kernel(unsigned *sync)
{
    if (threadIdx.x == 0) {
        unsigned e = sync[0] + 1;
        sync[0] = e == 0 ? 1 : e;
        sync[1] = 0;
        sync[2] = 0;
    }
}

// Block copy and subtraction through agent-scope atomics (coherent across the
// XCDs), with the entry -> thread map of wg_load_block / wg_subtract_block: the
// sync-free factorisation reads and writes the factor blocks only this way, since
// rows are factored by workgroups on any XCD.
template <typename T>
__device__ __forceinline__ void wg_load_block_coh(T *dst, const T *src, long n)
{
    const int tid = threadIdx.x, nth = blockDim.x;
    for (long b = tid; b < n; b += 8L * nth) {
        T v[8];
#pragma unroll
        for (int j = 0; j < 8; j++)
            v[j] = __hip_atomic_load(src + min(b + (long)j * nth, n - 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (b + (long)j * nth < n) dst[b + (long)j * nth] = v[j];
    }
}

template <typename T>
__device__ __forceinline__ void wg_subtract_block_coh(T *dst, const T *d, long n)
{
    const int tid = threadIdx.x, nth = blockDim.x;
    for (long b = tid; b < n; b += 8L * nth) {
        T v[8];
#pragma unroll
        for (int j = 0; j < 8; j++)
            v[j] = __hip_atomic_load(dst + min(b + (long)j * nth, n - 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (b + (long)j * nth < n)
                __hip_atomic_store(dst + b + (long)j * nth, v[j] - d[b + (long)j * nth], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The BILU0 factorisation without level launches: bilu0_factor_kernel's per-row
// arithmetic (the same bits) with the rows taken in `order` (the level lists
// concatenated) by persistent workgroups through an atomic ticket, as in
// bilu0_sweep_syncfree_kernel.  Before reading row k's blocks (inv(U_kk) and U_kl)
// every wave waits for k's ready flag, set once all of k's blocks are final; the
// blocks go through agent-scope atomics (wg_load_block_coh / wg_subtract_block_coh)
// and the flag is stored after every wave's stores have completed.  A row too long
// for the factorisation is reported through info and still published.
template <typename T>
__global__ void __launch_bounds__(SV_THREADS)
bilu0_factor_syncfree_kernel(int bsz, const int *__restrict__ rowpts, const int *__restrict__ colind,
                             const int *__restrict__ u2o, const int *__restrict__ order, int n, T *Mg,
                             int *__restrict__ info, unsigned *flags, const unsigned *epoch_p, unsigned *counter,
                             int spin_limit, int *err)
{
    extern __shared__ double lds_raw[];
    T *lds = (T *)lds_raw;
    const int tid = threadIdx.x, nth = blockDim.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const long b2 = (long)bsz * bsz;
    T *X = lds, *B = lds + b2, *L = lds + 2 * b2;
    LdsInv<T> w = lds_inv_area(lds + 3 * b2, bsz);
    int *rcol = w.piv + bsz + 2, *rord = rcol + FACTOR_MAXROW, *lst = rord + FACTOR_MAXROW,
        *ncol = lst + FACTOR_MAXROW + 1, *sticket = ncol + FACTOR_MAXROW;
    const unsigned epoch = __hip_atomic_load(epoch_p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // (loop control wave-uniform throughout: see bilu0_sweep_syncfree_kernel)
    while (true) {
        if (wv == 0) *sticket = __builtin_amdgcn_readfirstlane(atomicAdd(counter, lane == 0 ? 1u : 0u));
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(*sticket);
        if (t >= n) break;
        const int ri = order[t], oi = u2o[ri], d0 = rowpts[ri], len = rowpts[ri + 1] - d0;
        if (len <= FACTOR_MAXROW) {
            if (tid < len) {
                const int c = colind[d0 + tid];
                rcol[tid] = c;
                rord[tid] = u2o[c];
            }
            __syncthreads();
            if (tid == 0) {
                int cnt = 0;
                for (int q = 1; q < len; q++) {
                    if (rord[q] >= oi) continue;
                    int m = cnt++;
                    while (m > 0 && rord[lst[1 + m - 1]] > rord[q]) {
                        lst[1 + m] = lst[1 + m - 1];
                        m--;
                    }
                    lst[1 + m] = q;
                }
                lst[0] = cnt;
            }
            __syncthreads();
            const int cnt = __builtin_amdgcn_readfirstlane(lst[0]);
            for (int qq = 0; qq < cnt; qq++) {
                const int q = __builtin_amdgcn_readfirstlane(lst[1 + qq]);
                const int ik = d0 + q, rk = __builtin_amdgcn_readfirstlane(rcol[q]);
                const int ork = __builtin_amdgcn_readfirstlane(rord[q]);
                const int k0 = rowpts[rk], klen = rowpts[rk + 1] - k0;
                if (tid < klen && tid < FACTOR_MAXROW) ncol[tid] = colind[k0 + tid];
                wait_ready(flags + rk, epoch, spin_limit, err);
                wg_load_block_coh(X, Mg + (long)ik * b2, b2);
                wg_load_block_coh(B, Mg + (long)k0 * b2, b2);     // inv(U_kk)
                __syncthreads();
                T *Mik = Mg + (long)ik * b2;
                lds_block_product(X, B, bsz, [&](int i, int j, T c) {
                    L[i + bsz * j] = c;
                    __hip_atomic_store(Mik + i + bsz * j, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                });
                __syncthreads();
                for (int l = 0; l < len; l++) {
                    if (l != 0 && __builtin_amdgcn_readfirstlane(rord[l]) <= ork) continue;
                    const int rl = __builtin_amdgcn_readfirstlane(rcol[l]);
                    int km = -1;
                    for (int m = 1; m < klen && m < FACTOR_MAXROW; m++)
                        if (__builtin_amdgcn_readfirstlane(ncol[m]) == rl) {
                            km = m;
                            break;
                        }
                    if (km < 0) continue;
                    wg_load_block_coh(B, Mg + (long)(k0 + km) * b2, b2);
                    __syncthreads();
                    lds_block_product(L, B, bsz, [&](int i, int j, T c) { X[i + bsz * j] = c; });
                    __syncthreads();
                    wg_subtract_block_coh(Mg + (long)(d0 + l) * b2, X, b2);
                    __syncthreads();
                }
                __syncthreads();
            }
            wg_load_block_coh(X, Mg + (long)d0 * b2, b2);
            __syncthreads();
            const int r = lds_inverse(X, bsz, w.piv, w.rowk, w.colk, w.red);
            for (long i = tid; i < b2; i += nth)
                __hip_atomic_store(Mg + (long)d0 * b2 + i, X[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (tid == 0) info[ri] = r;
        } else {
            if (tid == 0) info[ri] = -len;
        }
        // publish row ri once every wave's stores of it have completed
        __builtin_amdgcn_s_waitcnt(0x0F70);   // s_waitcnt vmcnt(0)
        __syncthreads();
        if (wv == 0) __hip_atomic_store(flags + ri, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Block-Jacobi application y_i = inv(H_ii) x_i.  One wave per row.
template <typename TM>
__global__ void __launch_bounds__(SV_THREADS)
block_jacobi_kernel(int nrows, int bsz, const TM *__restrict__ Dinv, const double *__restrict__ x,
                    double *__restrict__ y)
{
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), wpb = blockDim.x >> 6;
    const long b2 = (long)bsz * bsz;
    for (int r = blockIdx.x * wpb + wv; r < nrows; r += gridDim.x * wpb) {
        const double *xr = x + (long)r * bsz;
        const double z0 = lane < bsz ? xr[lane] : 0.0, z1 = lane + 64 < bsz ? xr[lane + 64] : 0.0;
        double a0 = 0.0, a1 = 0.0;
        wave_block_matvec(bsz, Dinv + (long)r * b2, z0, z1, a0, a1);
        if (lane < bsz) y[(long)r * bsz + lane] = a0;
        if (lane + 64 < bsz) y[(long)r * bsz + lane + 64] = a1;
    }
}

// ---- GMRES vector kernels ----

constexpr int MDOT_GROUP = 8;

// partial[blk*k + i] = sum over the workgroup's slice of V_i . w  (V_i at V + i*ldv)
__global__ void __launch_bounds__(SV_THREADS)
multidot_kernel(long n, int k, const double *__restrict__ V, long ldv, const double *__restrict__ w,
                double *__restrict__ partial)
{
    __shared__ double red[SV_THREADS * MDOT_GROUP];
    const int tid = threadIdx.x, nth = blockDim.x;
    const long chunk = (n + gridDim.x - 1) / gridDim.x;
    const long q0 = blockIdx.x * chunk, q1 = q0 + chunk < n ? q0 + chunk : n;
    for (int ig = 0; ig < k; ig += MDOT_GROUP) {
        const int g = k - ig < MDOT_GROUP ? k - ig : MDOT_GROUP;
        double acc[MDOT_GROUP];
#pragma unroll
        for (int j = 0; j < MDOT_GROUP; j++) acc[j] = 0.0;
        for (long q = q0 + tid; q < q1; q += nth) {
            const double wq = w[q];
#pragma unroll
            for (int j = 0; j < MDOT_GROUP; j++)
                if (j < g) acc[j] += V[(ig + j) * ldv + q] * wq;
        }
#pragma unroll
        for (int j = 0; j < MDOT_GROUP; j++) red[j * SV_THREADS + tid] = acc[j];
        __syncthreads();
        for (int st = nth / 2; st > 0; st >>= 1) {
            if (tid < st)
#pragma unroll
                for (int j = 0; j < MDOT_GROUP; j++) red[j * SV_THREADS + tid] += red[j * SV_THREADS + tid + st];
            __syncthreads();
        }
        if (tid < g) partial[(long)blockIdx.x * k + ig + tid] = red[tid * SV_THREADS];
        __syncthreads();
    }
}

// out[i] = sum_b partial[b*k + i]
__global__ void multidot_reduce_kernel(int nblk, int k, const double *__restrict__ partial, double *__restrict__ out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < nblk; b++) s += partial[(long)b * k + i];
        out[i] = s;
    }
}

// w += sign * sum_i coef[i] V_i
__global__ void __launch_bounds__(SV_THREADS)
multi_axpy_kernel(long n, int k, const double *__restrict__ V, long ldv, const double *__restrict__ coef,
                  double sign, double *__restrict__ w)
{
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int i = 0; i < k; i++) s += coef[i] * V[i * ldv + q];
        w[q] += sign * s;
    }
}

// y = a*x + b*y
__global__ void axpby_kernel(long n, double a, const double *__restrict__ x, double b, double *__restrict__ y)
{
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        y[q] = a * x[q] + b * y[q];
}

__global__ void fill_kernel(long n, double v, double *__restrict__ y)
{
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) y[q] = v;
}

// ---- solution update (getU, solvers/solutionupdate.cpp:6-67) ----
// u_e <- u_e - alpha inv(D)F duh_e + alpha inv(D)Ru, with X = [inv(D)F  inv(D)Ru]
// row-major (na x ldx).  One workgroup per element.
__global__ void __launch_bounds__(SV_THREADS)
update_u_kernel(int ne, int npv, int nc, int nch, int ndf, int na, int nb, int ldx, const int *__restrict__ elcon,
                const double *__restrict__ X, const double *__restrict__ duh, double alpha, double *__restrict__ UDG)
{
    extern __shared__ double de[];
    const int tid = threadIdx.x, nth = blockDim.x;
    for (int e = blockIdx.x; e < ne; e += gridDim.x) {
        for (int q = tid; q < nb; q += nth) {
            const int j = q / nch, k = q % nch;
            de[q] = duh[(long)elcon[(long)e * ndf + j] * nch + k];
        }
        __syncthreads();
        const double *Xe = X + (long)e * na * ldx;
        for (int i = tid; i < na; i += nth) {
            const double *row = Xe + (long)i * ldx;
            double s = 0.0;
            for (int c = 0; c < nb; c++) s += row[c] * de[c];
            double *u = UDG + (long)e * npv * nc + i;
            *u = (*u - alpha * s) + alpha * row[nb];
        }
        __syncthreads();
    }
}

// DIRK stage source (solvers/dgsolvers.cpp:316-349): SH = a*UDGi + b*UDG + c*SH on the
// u components, 0 on the q components.
__global__ void stage_source_kernel(long ne, int npv, int nc, int ncu, double a, const double *__restrict__ UDGi,
                                    double b, const double *__restrict__ UDG, double c, double *__restrict__ SH)
{
    const long n = ne * npv * nc;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int comp = (int)((q / npv) % nc);
        SH[q] = comp < ncu ? (a * UDGi[q] + b * UDG[q]) + c * SH[q] : 0.0;
    }
}

// ---- launchers ----

static int grid_for(long n, int per = SV_THREADS)
{
    long g = (n + per - 1) / per;
    return (int)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

static hipError_t set_lds(const void *fn, size_t bytes)
{
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

size_t block_inverse_lds(int bsz) { return sizeof(double) * bsz * bsz + lds_inv_bytes(bsz); }
size_t bilu0_factor_lds(int bsz)
{
    return sizeof(double) * 3 * bsz * bsz + lds_inv_bytes(bsz) + sizeof(int) * (4 * FACTOR_MAXROW + 1);
}

hipError_t launch_block_inverse(int nrows, int bsz, const int *rowpts, const double *Hg, double *Dinv, int *info,
                                hipStream_t s)
{
    const size_t lds = block_inverse_lds(bsz);
    hipError_t e = set_lds((const void *)block_inverse_kernel, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(block_inverse_kernel, dim3(std::min(nrows, 8192)), dim3(SV_THREADS), lds, s, nrows, bsz,
                       rowpts, Hg, Dinv, info);
    return hipGetLastError();
}

hipError_t launch_mdf_weights(int nrows, int bsz, const int *rowpts, const double *Hg, const double *Dinv,
                              double *Cw, hipStream_t s)
{
    const size_t lds = sizeof(double) * (2 * bsz * bsz + SV_THREADS);
    hipError_t e = set_lds((const void *)mdf_weight_kernel, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mdf_weight_kernel, dim3(std::min(nrows, 8192)), dim3(SV_THREADS), lds, s, nrows, bsz, rowpts,
                       Hg, Dinv, Cw);
    return hipGetLastError();
}

template <typename T>
static hipError_t factor_level(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *rows, int nlev,
                               T *Mg, int *info, hipStream_t s)
{
    const size_t lds = bilu0_factor_lds(bsz);
    hipError_t e = set_lds((const void *)bilu0_factor_kernel<T>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bilu0_factor_kernel<T>, dim3(std::min(nlev, 8192)), dim3(SV_THREADS), lds, s, bsz, rowpts,
                       colind, u2o, rows, nlev, Mg, info);
    return hipGetLastError();
}

template <typename T>
static hipError_t factor_syncfree(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *order, int n,
                                  T *Mg, int *info, unsigned *flags, unsigned *sync, int nwg, int spin_limit, int *err,
                                  hipStream_t s)
{
    if (n <= 0 || nwg <= 0) return hipErrorInvalidValue;
    const size_t lds = bilu0_factor_lds(bsz) + 16;   // and the ticket
    hipError_t e = set_lds((const void *)bilu0_factor_syncfree_kernel<T>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sweep_reset_kernel, dim3(1), dim3(64), 0, s, sync);
    hipLaunchKernelGGL(bilu0_factor_syncfree_kernel<T>, dim3(std::min(n, nwg)), dim3(SV_THREADS), lds, s, bsz, rowpts,
                       colind, u2o, order, n, Mg, info, flags, sync, sync + 1, spin_limit, err);
    return hipGetLastError();
}

hipError_t launch_bilu0_factor_syncfree(int bsz, const int *rowpts, const int *colind, const int *u2o,
                                        const int *order, int n, double *Mg, int *info, unsigned *flags,
                                        unsigned *sync, int nwg, int spin_limit, int *err, hipStream_t s)
{
    return factor_syncfree(bsz, rowpts, colind, u2o, order, n, Mg, info, flags, sync, nwg, spin_limit, err, s);
}
hipError_t launch_bilu0_factor_syncfree(int bsz, const int *rowpts, const int *colind, const int *u2o,
                                        const int *order, int n, float *Mg, int *info, unsigned *flags,
                                        unsigned *sync, int nwg, int spin_limit, int *err, hipStream_t s)
{
    return factor_syncfree(bsz, rowpts, colind, u2o, order, n, Mg, info, flags, sync, nwg, spin_limit, err, s);
}

hipError_t launch_bilu0_factor_level(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *rows,
                                     int nlev, double *Mg, int *info, hipStream_t s)
{
    return factor_level(bsz, rowpts, colind, u2o, rows, nlev, Mg, info, s);
}

hipError_t launch_bilu0_factor_level(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *rows,
                                     int nlev, float *Mg, int *info, hipStream_t s)
{
    return factor_level(bsz, rowpts, colind, u2o, rows, nlev, Mg, info, s);
}

template <typename TM>
static hipError_t scale_upper(int bsz, int n, const int *rowpts, const int *colind, const int *u2o, TM *Mg,
                              hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const size_t lds = sizeof(double) * ((size_t)bsz * bsz + (size_t)std::max(1, SV_THREADS / bsz) * bsz);
    hipError_t e = set_lds((const void *)bilu0_scale_upper_kernel<TM>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bilu0_scale_upper_kernel<TM>, dim3(std::min(n, 16384)), dim3(SV_THREADS), lds, s, bsz, n,
                       rowpts, colind, u2o, Mg);
    return hipGetLastError();
}
hipError_t launch_bilu0_scale_upper(int bsz, int n, const int *rowpts, const int *colind, const int *u2o,
                                    double *Mg, hipStream_t s)
{
    return scale_upper(bsz, n, rowpts, colind, u2o, Mg, s);
}
hipError_t launch_bilu0_scale_upper(int bsz, int n, const int *rowpts, const int *colind, const int *u2o,
                                    float *Mg, hipStream_t s)
{
    return scale_upper(bsz, n, rowpts, colind, u2o, Mg, s);
}

template <bool BACKWARD, typename TM>
static hipError_t sweep_level(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *rows,
                              int nlev, const TM *Mg, const double *x, double *y, hipStream_t s)
{
    if (bsz > 128) return hipErrorInvalidValue;
    if (bsz <= 64)
        hipLaunchKernelGGL((bilu0_sweep_kernel<BACKWARD, true, TM>), dim3(std::min(nlev, 8192)), dim3(SV_THREADS), 0,
                           s, bsz, rowpts, colind, u2o, rows, nlev, Mg, x, y);
    else
        hipLaunchKernelGGL((bilu0_sweep_kernel<BACKWARD, false, TM>), dim3(std::min(nlev, 8192)), dim3(SV_THREADS), 0,
                           s, bsz, rowpts, colind, u2o, rows, nlev, Mg, x, y);
    return hipGetLastError();
}

hipError_t launch_bilu0_lsolve_level(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *rows,
                                     int nlev, const double *Mg, const double *x, double *y, hipStream_t s)
{
    return sweep_level<false>(bsz, rowpts, colind, u2o, rows, nlev, Mg, x, y, s);
}
hipError_t launch_bilu0_lsolve_level(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *rows,
                                     int nlev, const float *Mg, const double *x, double *y, hipStream_t s)
{
    return sweep_level<false>(bsz, rowpts, colind, u2o, rows, nlev, Mg, x, y, s);
}
hipError_t launch_bilu0_usolve_level(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *rows,
                                     int nlev, const double *Mg, const double *y, double *v, hipStream_t s)
{
    return sweep_level<true>(bsz, rowpts, colind, u2o, rows, nlev, Mg, y, v, s);
}
hipError_t launch_bilu0_usolve_level(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *rows,
                                     int nlev, const float *Mg, const double *y, double *v, hipStream_t s)
{
    return sweep_level<true>(bsz, rowpts, colind, u2o, rows, nlev, Mg, y, v, s);
}

// Both sweeps of one BILU0 application without level launches: x -> t (forward,
// rows in lorder) -> y (backward, rows in uorder).  sync: [0] epoch, [1], [2] the
// tickets of the two sweeps; flags: n forward then n backward ready flags.
template <typename TM>
static hipError_t sweeps_syncfree(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *lorder,
                                  const int *uorder, int n, const TM *Mg, const double *x, double *t, double *y,
                                  unsigned *flags, unsigned *sync, int nwg, int spin_limit, int *err, hipStream_t s)
{
    if (bsz > 128 || n <= 0 || nwg <= 0) return hipErrorInvalidValue;
    const int g = std::min(n, nwg);
    hipLaunchKernelGGL(sweep_reset_kernel, dim3(1), dim3(64), 0, s, sync);
    if (bsz <= 64) {
        hipLaunchKernelGGL((bilu0_sweep_syncfree_kernel<false, true, TM>), dim3(g), dim3(SV_THREADS), 0, s, bsz, rowpts,
                           colind, u2o, lorder, n, Mg, x, t, flags, sync, sync + 1, spin_limit, err);
        hipLaunchKernelGGL((bilu0_sweep_syncfree_kernel<true, true, TM>), dim3(g), dim3(SV_THREADS), 0, s, bsz, rowpts,
                           colind, u2o, uorder, n, Mg, t, y, flags + n, sync, sync + 2, spin_limit, err);
    } else {
        hipLaunchKernelGGL((bilu0_sweep_syncfree_kernel<false, false, TM>), dim3(g), dim3(SV_THREADS), 0, s, bsz,
                           rowpts, colind, u2o, lorder, n, Mg, x, t, flags, sync, sync + 1, spin_limit, err);
        hipLaunchKernelGGL((bilu0_sweep_syncfree_kernel<true, false, TM>), dim3(g), dim3(SV_THREADS), 0, s, bsz,
                           rowpts, colind, u2o, uorder, n, Mg, t, y, flags + n, sync, sync + 2, spin_limit, err);
    }
    return hipGetLastError();
}

hipError_t launch_bilu0_syncfree(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *lorder,
                                 const int *uorder, int n, const double *Mg, const double *x, double *t, double *y,
                                 unsigned *flags, unsigned *sync, int nwg, int spin_limit, int *err, hipStream_t s)
{
    return sweeps_syncfree(bsz, rowpts, colind, u2o, lorder, uorder, n, Mg, x, t, y, flags, sync, nwg, spin_limit,
                           err, s);
}
hipError_t launch_bilu0_syncfree(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *lorder,
                                 const int *uorder, int n, const float *Mg, const double *x, double *t, double *y,
                                 unsigned *flags, unsigned *sync, int nwg, int spin_limit, int *err, hipStream_t s)
{
    return sweeps_syncfree(bsz, rowpts, colind, u2o, lorder, uorder, n, Mg, x, t, y, flags, sync, nwg, spin_limit,
                           err, s);
}

template <typename TM>
static hipError_t bj(int nrows, int bsz, const TM *Dinv, const double *x, double *y, hipStream_t s)
{
    if (bsz > 128) return hipErrorInvalidValue;
    const int wpb = SV_THREADS / 64;
    hipLaunchKernelGGL(block_jacobi_kernel<TM>, dim3(std::min((nrows + wpb - 1) / wpb, 8192)), dim3(SV_THREADS), 0,
                       s, nrows, bsz, Dinv, x, y);
    return hipGetLastError();
}
hipError_t launch_block_jacobi(int nrows, int bsz, const double *Dinv, const double *x, double *y, hipStream_t s)
{
    return bj(nrows, bsz, Dinv, x, y, s);
}
hipError_t launch_block_jacobi(int nrows, int bsz, const float *Dinv, const double *x, double *y, hipStream_t s)
{
    return bj(nrows, bsz, Dinv, x, y, s);
}

__global__ void to_float_kernel(long n, const double *__restrict__ x, float *__restrict__ y)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = (float)x[i];
}

hipError_t launch_to_float(long n, const double *x, float *y, hipStream_t s)
{
    hipLaunchKernelGGL(to_float_kernel, dim3(grid_for(n)), dim3(SV_THREADS), 0, s, n, x, y);
    return hipGetLastError();
}

int multidot_blocks(long n) { return (int)std::min<long>(512, std::max<long>(1, n / 2048)); }

hipError_t launch_multidot(long n, int k, const double *V, long ldv, const double *w, double *partial, double *out,
                           hipStream_t s)
{
    const int nblk = multidot_blocks(n);
    hipLaunchKernelGGL(multidot_kernel, dim3(nblk), dim3(SV_THREADS), 0, s, n, k, V, ldv, w, partial);
    hipLaunchKernelGGL(multidot_reduce_kernel, dim3(1), dim3(128), 0, s, nblk, k, partial, out);
    return hipGetLastError();
}

hipError_t launch_multi_axpy(long n, int k, const double *V, long ldv, const double *coef, double sign, double *w,
                             hipStream_t s)
{
    hipLaunchKernelGGL(multi_axpy_kernel, dim3(grid_for(n)), dim3(SV_THREADS), 0, s, n, k, V, ldv, coef, sign, w);
    return hipGetLastError();
}

hipError_t launch_axpby(long n, double a, const double *x, double b, double *y, hipStream_t s)
{
    hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(SV_THREADS), 0, s, n, a, x, b, y);
    return hipGetLastError();
}

hipError_t launch_fill(long n, double v, double *y, hipStream_t s)
{
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(SV_THREADS), 0, s, n, v, y);
    return hipGetLastError();
}

hipError_t launch_update_u(int ne, int npv, int nc, int nch, int ndf, int na, int nb, int ldx, const int *elcon,
                           const double *X, const double *duh, double alpha, double *UDG, hipStream_t s)
{
    hipLaunchKernelGGL(update_u_kernel, dim3(std::min(ne, 8192)), dim3(SV_THREADS), sizeof(double) * nb, s, ne, npv,
                       nc, nch, ndf, na, nb, ldx, elcon, X, duh, alpha, UDG);
    return hipGetLastError();
}

hipError_t launch_stage_source(long ne, int npv, int nc, int ncu, double a, const double *UDGi, double b,
                               const double *UDG, double c, double *SH, hipStream_t s)
{
    hipLaunchKernelGGL(stage_source_kernel, dim3(grid_for(ne * npv * nc)), dim3(SV_THREADS), 0, s, ne, npv, nc, ncu,
                       a, UDGi, b, UDG, c, SH);
    return hipGetLastError();
}

}  // namespace hfem
