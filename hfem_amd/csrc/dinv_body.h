// Device body of the per-element inverse inv(D) by blocked Gauss-Jordan on the
// MFMA (assembly/elementschur.cpp:94-105 factor and solve), shared by the
// per-family kernels and the ragged batch kernel.
#pragma once
#include "hfem_device.h"

namespace hfem {

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

// Per-type pieces of the MFMA Gauss-Jordan: the 16x16x4 MFMA and the row that
// element v of lane group lk holds in a 16x16 result tile (measured on gfx950:
// FP64 C(lk + 4v, lane%16), FP32 C(4lk + v, lane%16); both take A(lane%16,
// lane/16) and B(lane/16, lane%16)), the pivot key (0 for used rows, else 1 +
// the |value| bit pattern, so that NaN ranks highest) with its wave maximum,
// and the broadcast of one lane's value.
template <typename T> struct GJ;
template <> struct GJ<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    typedef unsigned long long key_t;
    static __device__ __forceinline__ v4 mfma(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
    static __device__ __forceinline__ key_t key(double x)
    {
        return ((unsigned long long)__double_as_longlong(x) & 0x7fffffffffffffffull) + 1;
    }
    static __device__ __forceinline__ key_t wave_max(key_t k) { return wave_max_u64(k); }
    static __device__ __forceinline__ double bcast(double v, int lane)
    {
        const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
        const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
        return __hiloint2double(hi, lo);
    }
    // 1/x for a normal nonzero pivot: the hardware reciprocal refined by two
    // Newton-Raphson steps r <- r + r (1 - x r) (each squares the relative error),
    // three dependent FMAs shorter than an IEEE division on the pivot chain
    static __device__ __forceinline__ double recip(double x)
    {
        double r = __builtin_amdgcn_rcp(x);
        r = fma(r, fma(-x, r, 1.0), r);
        return fma(r, fma(-x, r, 1.0), r);
    }
};
template <> struct GJ<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    typedef unsigned key_t;
    static __device__ __forceinline__ v4 mfma(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
    static __device__ __forceinline__ key_t key(float x) { return (__float_as_uint(x) & 0x7fffffffu) + 1u; }
    static __device__ __forceinline__ key_t wave_max(key_t v)
    {
#define HFEM_DPP_MAX32(ctrl, rm, bm)                                                                         \
    {                                                                                                        \
        const unsigned o = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, ctrl, rm, bm, false);       \
        v = o > v ? o : v;                                                                                   \
    }
        HFEM_DPP_MAX32(0xb1, 0xf, 0xf)    // quad_perm [1,0,3,2]
        HFEM_DPP_MAX32(0x4e, 0xf, 0xf)    // quad_perm [2,3,0,1]
        HFEM_DPP_MAX32(0x114, 0xf, 0xe)   // row_shr:4
        HFEM_DPP_MAX32(0x118, 0xf, 0xc)   // row_shr:8
        HFEM_DPP_MAX32(0x142, 0xa, 0xf)   // row_bcast:15
        HFEM_DPP_MAX32(0x143, 0xc, 0xf)   // row_bcast:31
#undef HFEM_DPP_MAX32
        return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
    }
    static __device__ __forceinline__ float bcast(float v, int lane)
    {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
    }
    static __device__ __forceinline__ float recip(float x)
    {
        const float r = __builtin_amdgcn_rcpf(x);
        return fmaf(r, fmaf(-x, r, 1.0f), r);
    }
};

// ---------------------------------------------------------------------------
// inv(D) by blocked Gauss-Jordan on the MFMA, in arithmetic T (FP64, or FP32
// for the mixed condensation: D is rounded to FP32 on the way in), written in
// TO.  Partial pivoting with implicit row exchanges: step j takes as pivot the
// unused row p_j of largest |a(i, j)| (lowest index on ties, as IDAMAX) and
// applies the exchange step to every row; rows are never moved, so
// inv(D)[label[i], pof[c]] = G[i][c] holds on exit (pof[j] = p_j, label[p_j] = j).  The na x na matrix (padded to NN = 16 NT)
// lives in registers as 16 x 16 MFMA result tiles spread over the W waves (tile
// t = s W + wave in slot s).  Per 16-column panel:
//   1. wave 0 runs the 16 unblocked steps on the panel alone, in registers
//      (two or three rows per lane, no barrier); the panel then holds Gp, whose
//      column k is T e_{P[k]} for the product T of the panel's elementary steps
//      and its pivot rows P;
//   2. every other column becomes T a = Abar + Gp a(P, :), with Abar = a with
//      the rows P zeroed: a rank-16 update done as four MFMAs per tile straight
//      into the register tiles (A operand Gp, B operand the old pivot rows,
//      both from LDS).
// The panel steps of one wave are the critical path of an element, so the
// launch shapes trade waves per element for elements per CU: with W = 1 the
// whole matrix sits in one wave's registers and no barrier waits on a panel.
// The result is written column by column of inv(D) through an LDS staging tile
// so that the permutation (output column pof[c], row label^-1) is applied with
// coalesced stores.
// ---------------------------------------------------------------------------
// LDS bytes of one element's group: pan and Z (NN x 17 of T each), label, pof,
// rowk (NN ints each) and the singularity flag, 16-byte aligned.
__host__ __device__ constexpr int dinv_group_lds(int esize, int NT)
{
    return ((2 * 16 * NT * 17 * esize + 3 * 16 * NT * 4 + 4) + 15) / 16 * 16;
}
template <typename T, int NT>
__host__ __device__ constexpr int dinv_group_lds()
{
    return dinv_group_lds((int)sizeof(T), NT);
}

// The body for one workgroup of GR groups of W waves: group g inverts element
// r GR + g of every round r that the source src hands out (src.first(), then
// src.next(r) until a value >= the number of rounds ceil(n / GR)).  The groups
// of a workgroup run in lockstep through the same barriers: a group whose
// element lies beyond n repeats group 0's element and stores nothing.  lds:
// GR group regions of dinv_group_lds<T, NT>() bytes.
template <typename T, typename TO, int NT, int W, int GR, class Src>
__device__ __forceinline__ void dinv_body(int na, int e0, int n, const double *__restrict__ D, TO *__restrict__ Dinv,
                                          int *__restrict__ info, char *lds, Src &src)
{
    typedef GJ<T> G;
    typedef typename G::v4 v4;
    typedef typename G::key_t key_t;
    constexpr int NN = 16 * NT, NTILE = NT * NT, TPW = (NTILE + W - 1) / W, RPL = (NN + 63) / 64;
    constexpr int PMW = (4 * TPW + 31) / 32;   // words of the per-lane pivot-row mask
    constexpr int PS = 17;                // panel row stride: odd, so a column walk is bank-conflict free
    const int gw = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), grp = GR == 1 ? 0 : gw / W;
    char *gb = lds + grp * dinv_group_lds<T, NT>();
    T *pan = (T *)gb;                     // panel / staging of 16 columns, row-major: pan[i*PS + k]
    T *Z = pan + NN * PS;                 // pivot rows of the panel, Z[k*NN + c]; output staging
    int *label = (int *)(Z + NN * PS), *pof = label + NN, *rowk = pof + NN;
    int &sing_s = rowk[NN];
    const int tid = threadIdx.x - 64 * W * grp, lane = tid & 63, wave = gw - W * grp, lr = lane & 15,
              lk = lane >> 4;
    const int nr = (n + GR - 1) / GR;
    // A last panel with fewer than 16 pivots leaves rows of Z unwritten that its
    // rank-16 update multiplies by the panel's zero columns: they must hold finite
    // values (LDS comes with whatever an earlier kernel left), so Z starts at zero
    // (later it only ever holds finite pivot rows and output tiles)
    for (int i = tid; i < NN * PS; i += 64 * W) Z[i] = (T)0;
    for (int rd = src.first(); rd < nr; rd = src.next(rd)) {
        const int t = rd * GR + grp;
        const bool live = t < n;
        const long e = e0 + (live ? t : rd * GR);
        const double *De = D + e * (long)na * na;
        // D into the register tiles: all loads of a wave issued together
        v4 a[TPW];
        {
            int wv = __builtin_amdgcn_readfirstlane(wave);
            asm volatile("" : "+s"(wv));
#pragma unroll
            for (int s = 0; s < TPW; s++) {
                const int T_ = s * W + wv, r = T_ / NT, q = T_ % NT;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = r * 16 + G::row(lk, v), c = q * 16 + lr;
                    const bool ok = T_ < NTILE && i < na && c < na;
                    const double d = De[ok ? i + (long)na * c : 0];
                    a[s][v] = ok ? (T)d : (T)0;
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
                const int T_ = s * W + wv, r = T_ / NT, q = T_ % NT;
                if (T_ < NTILE && q == p)
#pragma unroll
                    for (int v = 0; v < 4; v++) pan[(r * 16 + G::row(lk, v)) * PS + lr] = a[s][v];
            }
            __syncthreads();
            // 2. unblocked steps on the panel (wave 0) in registers: lane l holds rows
            //    l, l+64, ...; the pivot row is broadcast by readlane, and the pivot is
            //    found from an exact wave maximum of the |value| bit patterns (DPP) with
            //    the lowest row among equal keys taken from a ballot
            if (wave == 0) {
                T x[RPL][16];
                bool used[RPL];
                int pk[RPL];
#pragma unroll
                for (int h = 0; h < RPL; h++) {
                    const int i = lane + 64 * h;
                    used[h] = i >= NN || label[i] >= 0;
                    pk[h] = -1;
#pragma unroll
                    for (int k = 0; k < 16; k++) x[h][k] = i < NN ? pan[i * PS + k] : (T)0;
                }
                int sing = 0;
#pragma unroll
                for (int jj = 0; jj < 16; jj++) {
                    // fully unrolled so that x[h][jj] is a register (no early exit)
                    const int j = p * 16 + jj;
                    if (j < na) {
                        key_t key[RPL], m = 0;
#pragma unroll
                        for (int h = 0; h < RPL; h++) {
                            key[h] = used[h] ? (key_t)0 : G::key(x[h][jj]);
                            m = key[h] > m ? key[h] : m;
                        }
                        m = G::wave_max(m);
                        int pr = NN;
#pragma unroll
                        for (int h = RPL - 1; h >= 0; h--) {
                            const unsigned long long b = __ballot(key[h] == m);
                            if (b) pr = 64 * h + __ffsll((long long)b) - 1;
                        }
                        const int owner = pr & 63, hp = pr >> 6;
                        T prow[16];
#pragma unroll
                        for (int k = 0; k < 16; k++) {
                            T v = x[0][k];
#pragma unroll
                            for (int h = 1; h < RPL; h++) v = hp == h ? x[h][k] : v;
                            prow[k] = G::bcast(v, owner);
                        }
                        const T pv = prow[jj];
                        if (pv == (T)0 && !sing) sing = j + 1;
                        const T pinv = pv == (T)0 ? (T)0 : G::recip(pv);
                        // scaled pivot row s_k = a(p, k) / a(p, j); every row takes one FMA per
                        // column, then the pivot row (one lane) is overwritten
                        T sk[16];
#pragma unroll
                        for (int k = 0; k < 16; k++) sk[k] = prow[k] * pinv;
#pragma unroll
                        for (int h = 0; h < RPL; h++) {
                            const int i = lane + 64 * h;
                            const T f = x[h][jj];
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
            // 3. the old pivot rows of the other columns.  A tile's four pivot
            //    numbers are read together (one LDS round trip per tile, not per
            //    row), and which of the lane's rows are pivot rows is kept as a
            //    bitmask (bit 4s + v) for step 4
            wv = __builtin_amdgcn_readfirstlane(wave);
            asm volatile("" : "+s"(wv));
            unsigned pm[PMW];
#pragma unroll
            for (int i = 0; i < PMW; i++) pm[i] = 0u;
#pragma unroll
            for (int s = 0; s < TPW; s++) {
                const int T_ = s * W + wv, r = T_ / NT, q = T_ % NT;
                if (T_ < NTILE && q != p) {
                    int k[4];
#pragma unroll
                    for (int v = 0; v < 4; v++) k[v] = rowk[r * 16 + G::row(lk, v)];
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        if (k[v] >= 0) {
                            Z[k[v] * NN + q * 16 + lr] = a[s][v];
                            pm[(4 * s + v) / 32] |= 1u << ((4 * s + v) % 32);
                        }
                }
            }
            __syncthreads();
            // 4. the rank-16 update (the panel columns take Gp)
            wv = __builtin_amdgcn_readfirstlane(wave);
            asm volatile("" : "+s"(wv));
#pragma unroll
            for (int s = 0; s < TPW; s++) {
                const int T_ = s * W + wv, r = T_ / NT, q = T_ % NT;
                if (T_ >= NTILE) continue;
                if (q == p) {
#pragma unroll
                    for (int v = 0; v < 4; v++) a[s][v] = pan[(r * 16 + G::row(lk, v)) * PS + lr];
                } else {
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        if ((pm[(4 * s + v) / 32] >> ((4 * s + v) % 32)) & 1u) a[s][v] = (T)0;
#pragma unroll
                    for (int ks = 0; ks < 4; ks++)
                        a[s] = G::mfma(pan[(r * 16 + lr) * PS + ks * 4 + lk], Z[(ks * 4 + lk) * NN + q * 16 + lr], a[s]);
                }
            }
            __syncthreads();
        }
        // inv(D)(a, pof[c]) = G(pof[a], c), two 16-column tiles of G at a time
        // (staged in pan and Z).  A wave writes whole columns, lanes along the rows
        // a (coalesced stores); the staging rows pof[a] of a lane's rows are read
        // once per element
        TO *Oe = Dinv + e * (long)na * na;
        int pa[RPL];
#pragma unroll
        for (int h = 0; h < RPL; h++) pa[h] = lane + 64 * h < na ? pof[lane + 64 * h] : 0;
        for (int q0 = 0; q0 < NT; q0 += 2) {
            int wv = __builtin_amdgcn_readfirstlane(wave);
            asm volatile("" : "+s"(wv));
#pragma unroll
            for (int s = 0; s < TPW; s++) {
                const int T_ = s * W + wv, r = T_ / NT, q = T_ % NT;
                if (T_ < NTILE && (q == q0 || q == q0 + 1)) {
                    T *st = q == q0 ? pan : Z;
#pragma unroll
                    for (int v = 0; v < 4; v++) st[(r * 16 + G::row(lk, v)) * PS + lr] = a[s][v];
                }
            }
            __syncthreads();
            const int ncols = min(32, na - q0 * 16);
            for (int cc = wave; cc < ncols; cc += W) {
                const int col = pof[q0 * 16 + cc];
                const T *st = cc < 16 ? pan : Z;
#pragma unroll
                for (int h = 0; h < RPL; h++)
                    if (live && lane + 64 * h < na) Oe[lane + 64 * h + (long)na * col] = (TO)st[pa[h] * PS + (cc & 15)];
            }
            __syncthreads();
        }
        if (tid == 0 && live) info[e] = sing_s;
    }
}

}  // namespace hfem
