// Batched per-element products of the static condensation on the gfx950 MFMA
// (assembly/elementschur.cpp:94-105 and :217-222 for the FP64 path; the mixed
// path's FP32 solve and its FP64 refinement residual).
//
// For every element e of a range and every output entry (i, j), i < M, j < N:
//   C[e](i, j) = beta C0[e](i, j) + alpha sum_{k < K} A[e](k, i) B[e](k, j)
// with A(k, i) at A + e sa + k lda + i and B(k, j) at B + e sb + k ldb + j
// (i and j contiguous), C(i, j) at C + e sc + i ldc + j, C0 likewise.  All the
// condensation's operands have this form in the library's layouts:
//   X  = inv(D) [F Ru]      A = inv(D) (column-major), B = [F Ru] (row-major)
//   AH^T = H^T - X^T K^T    A = X (row-major), B = K (column-major), C = H
//                           (column-major, i = trace column incl. FH, j = row)
//   R  = [F Ru] - D X       A = D, B = X, C0 = [F Ru]
// so every operand stage is a run of contiguous rows.
//
// One workgroup computes one strip of an element's output: the strip's 16x16
// tiles (tile rows [r0, r1) x all tile columns) are dealt to the W waves (tile
// t = wave + W s, s < TPW), each tile's accumulator held in registers.  The
// K dimension is walked in chunks of KC: a chunk of A (the strip's rows) and of
// B (all columns) is staged in LDS, converted to the MFMA type TM, while the
// next chunk's global loads are in flight (double buffering, one barrier per
// chunk).  Strips of one element map to the same XCD (block b runs on XCD
// b mod 8 and the strips of element e are the work items with the same
// residue), so the operand shared by the strips is fetched from HBM once.
// Optionally the epilogue records per element the largest |alpha A B| and the
// largest |C| (atomic maxima of the bit patterns), the refinement's
// convergence test.
#include "ctx.h"

#include <algorithm>

namespace hfem {

template <typename TM> struct Mf;
template <> struct Mf<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mfma(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row of element v of lane group lk in a 16x16 result tile (measured on gfx950)
    static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};
template <> struct Mf<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mfma(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};

__device__ __forceinline__ unsigned long long wave_max_bits(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    return v;
}

// LDS row stride (elements of TM) of a staged chunk of width w: w rounded so that
// the four lane groups of an MFMA operand read (k rows apart) start on
// different banks.
__host__ __device__ inline int bgemm_ld(int w, int esize)
{
    const int P = 256 / esize;
    return w + ((16 - w) % P + P) % P;
}

template <typename TM, typename TA, typename TB, int W, int TPW, int KC, int PF>
__global__ void __launch_bounds__(64 * W)
bgemm_kernel(BgemmArgs g)
{
    typedef Mf<TM> F;
    typedef typename F::v4 v4;
    extern __shared__ double bg_lds[];
    constexpr int NTH = 64 * W;
    const int tid = threadIdx.x, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int TR = (g.M + 15) >> 4, TC = (g.N + 15) >> 4;
    const int rows_max = (TR + g.nstrip - 1) / g.nstrip;
    const int MS = 16 * rows_max, NS = 16 * TC;
    const int lda_s = bgemm_ld(MS, sizeof(TM)), ldb_s = bgemm_ld(NS, sizeof(TM));
    const int bufsz = KC * (lda_s + ldb_s);
    // tile table: staged offsets (16 r, 16 c) of tile t = r TC + c, after the two
    // staging buffers.  Slots past the strip's tiles point at tile (0, 0): their
    // products are computed unpredicated and never stored.
    TM *lds = (TM *)bg_lds;
    int2 *ttab = (int2 *)(lds + 2 * bufsz);
    for (int tt = tid; tt < W * TPW; tt += 64 * W)
        ttab[tt] = tt / TC < rows_max ? make_int2(16 * (tt / TC), 16 * (tt % TC)) : make_int2(0, 0);
    __syncthreads();
    const int nchunk = (g.K + KC - 1) / KC;
    // staging: KC*MS values of A then KC*NS values of B per chunk; value idx of a
    // chunk is (k, i) = divmod(idx, MS) (A) or divmod(idx - KC*MS, NS) (B), the
    // quotient taken in FP32 (exact: idx < 2^14 and the fractional part of
    // (idx + 1/2) / w is at least 1/(2w) away from an integer)
    const int na_stage = KC * MS, ntot = KC * (MS + NS);
    const float rms = 1.0f / (float)MS, rns = 1.0f / (float)NS;
    // per-wave tile offsets in the staged chunk (uniform: scalar registers)
    const int groups = (g.n + 7) / 8;
    const long items = (long)groups * 8 * g.nstrip;
    for (long it = blockIdx.x; it < items; it += gridDim.x) {
        // work item -> (element, strip): the strips of one element share it mod 8
        const int t = (int)((it / (8 * g.nstrip)) * 8 + it % 8), s = (int)((it / 8) % g.nstrip);
        if (t >= g.n) continue;
        const long e = (long)g.e0 + t;
        const int r0 = s * TR / g.nstrip, r1 = (s + 1) * TR / g.nstrip;
        const int nt = (r1 - r0) * TC, i0 = 16 * r0, mrows = min(g.M - i0, 16 * (r1 - r0));
        const TA *Ae = (const TA *)g.A + e * g.sa + i0;
        const TB *Be = (const TB *)g.B + e * g.sb;
        TM pre[PF];
        auto load = [&](int ch) {
            const int k0 = ch * KC;
#pragma unroll
            for (int u = 0; u < PF; u++) {
                const int idx = tid + NTH * u;
                TM v = (TM)0;
                if (idx < na_stage) {
                    const int k = (int)(((float)idx + 0.5f) * rms), i = idx - k * MS;
                    if (k0 + k < g.K && i < mrows) v = (TM)Ae[(long)(k0 + k) * g.lda + i];
                } else if (idx < ntot) {
                    const int q = idx - na_stage, k = (int)(((float)q + 0.5f) * rns), j = q - k * NS;
                    if (k0 + k < g.K && j < g.N) v = (TM)Be[(long)(k0 + k) * g.ldb + j];
                }
                pre[u] = v;
            }
        };
        auto store = [&](int buf) {
            TM *As = lds + buf * bufsz, *Bs = As + KC * lda_s;
#pragma unroll
            for (int u = 0; u < PF; u++) {
                const int idx = tid + NTH * u;
                if (idx < na_stage) {
                    const int k = (int)(((float)idx + 0.5f) * rms), i = idx - k * MS;
                    As[k * lda_s + i] = pre[u];
                } else if (idx < ntot) {
                    const int q = idx - na_stage, k = (int)(((float)q + 0.5f) * rns), j = q - k * NS;
                    Bs[k * ldb_s + j] = pre[u];
                }
            }
        };
        v4 acc[TPW];
#pragma unroll
        for (int q = 0; q < TPW; q++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[q][v] = (TM)0;
        load(0);
        store(0);
        __syncthreads();
        for (int ch = 0; ch < nchunk; ch++) {
            if (ch + 1 < nchunk) load(ch + 1);
            const TM *As = lds + (ch & 1) * bufsz, *Bs = As + KC * lda_s;
            // tiles in groups of four: four independent accumulators per k step, the
            // group's staged offsets read once per chunk from the tile table
#pragma unroll
            for (int q0 = 0; q0 < TPW; q0 += 4) {
                int ao[4], bo[4];
#pragma unroll
                for (int qq = 0; qq < 4; qq++) {
                    const int2 o = ttab[min(wave + W * (q0 + qq), W * TPW - 1)];   // (clamp: q0 + qq >= TPW is not used)
                    ao[qq] = o.x;
                    bo[qq] = o.y;
                }
#pragma unroll
                for (int ks = 0; ks < KC / 4; ks++) {
                    const TM *Ak = As + (ks * 4 + lk) * lda_s + lr, *Bk = Bs + (ks * 4 + lk) * ldb_s + lr;
#pragma unroll
                    for (int qq = 0; qq < 4; qq++)
                        if (q0 + qq < TPW) acc[q0 + qq] = F::mfma(Ak[ao[qq]], Bk[bo[qq]], acc[q0 + qq]);
                }
            }
            if (ch + 1 < nchunk) store((ch + 1) & 1);
            __syncthreads();
        }
        // epilogue in FP64: C = beta C0 + alpha acc
        double *Ce = g.C + e * g.sc;
        const double *C0e = g.C0 ? g.C0 + e * g.sc0 : nullptr;
        unsigned long long dm = 0, cm = 0;
#pragma unroll
        for (int q = 0; q < TPW; q++) {
            if (wave + W * q < nt) {
                const int2 o = ttab[wave + W * q];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = i0 + o.x + F::row(lk, v), j = o.y + lr;
                    if (i < g.M && j < g.N) {
                        const double d = g.alpha * (double)acc[q][v];
                        const double y = C0e ? fma(g.beta, C0e[(long)i * g.ldc0 + j], d) : d;
                        Ce[(long)i * g.ldc + j] = y;
                        const unsigned long long bd = (unsigned long long)__double_as_longlong(fabs(d));
                        const unsigned long long by = (unsigned long long)__double_as_longlong(fabs(y));
                        dm = bd > dm ? bd : dm;
                        cm = by > cm ? by : cm;
                    }
                }
            }
        }
        if (g.dmax) {
            dm = wave_max_bits(dm);
            cm = wave_max_bits(cm);
            if (lane == 0) {
                atomicMax(g.dmax + t, dm);
                atomicMax(g.cmax + t, cm);
            }
        }
    }
}

// The same products without LDS: every wave owns an RT x CT block of 16x16
// output tiles and reads its MFMA operands straight from memory (16 contiguous
// values per lane group, served by L2 after the first wave of an element touches
// them), with the next k step's operands loaded while the current ones are
// multiplied.  Small register footprint, so many waves per SIMD hide the
// latency that the staged form waits on.  A workgroup of W waves takes W
// consecutive blocks of one element; workgroups of one element run on one XCD
// (the mod-8 mapping of the staged form).
template <typename TM, typename TA, typename TB, int RT, int CT>
__global__ void __launch_bounds__(256)
bgemm_direct_kernel(BgemmArgs g)
{
    typedef Mf<TM> F;
    typedef typename F::v4 v4;
    constexpr int W = 4;
    const int tid = threadIdx.x, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int TR = (g.M + 15) >> 4, TC = (g.N + 15) >> 4;
    const int nbr = (TR + RT - 1) / RT, nbc = (TC + CT - 1) / CT, nblk = nbr * nbc;
    const int nwi = (nblk + W - 1) / W;          // workgroup items per element
    const int groups = (g.n + 7) / 8;
    const long items = (long)groups * 8 * nwi;
    const int nks = (g.K + 3) >> 2;
    for (long it = blockIdx.x; it < items; it += gridDim.x) {
        const int t = (int)((it / (8 * nwi)) * 8 + it % 8), wi = (int)((it / 8) % nwi);
        const int b = wi * W + wave;
        if (t >= g.n || b >= nblk) continue;
        const long e = (long)g.e0 + t;
        const int br = b / nbc, bc = b - br * nbc;
        const TA *Ae = (const TA *)g.A + e * g.sa;
        const TB *Be = (const TB *)g.B + e * g.sb;
        int ia[RT], jb[CT];
        bool ra[RT], cb[CT];
#pragma unroll
        for (int rr = 0; rr < RT; rr++) {
            ia[rr] = 16 * (br * RT + rr) + lr;
            ra[rr] = ia[rr] < g.M;
            ia[rr] = ra[rr] ? ia[rr] : 0;
        }
#pragma unroll
        for (int cc = 0; cc < CT; cc++) {
            jb[cc] = 16 * (bc * CT + cc) + lr;
            cb[cc] = jb[cc] < g.N;
            jb[cc] = cb[cc] ? jb[cc] : 0;
        }
        v4 acc[RT][CT];
#pragma unroll
        for (int rr = 0; rr < RT; rr++)
#pragma unroll
            for (int cc = 0; cc < CT; cc++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[rr][cc][v] = (TM)0;
        TM fa[2][RT], fb[2][CT];
        auto load = [&](int ks, int buf) {
            const int k = 4 * ks + lk;
            const bool kok = k < g.K;
            const long ka = (long)(kok ? k : 0) * g.lda, kb = (long)(kok ? k : 0) * g.ldb;
#pragma unroll
            for (int rr = 0; rr < RT; rr++) {
                const TA v = Ae[ka + ia[rr]];
                fa[buf][rr] = (kok && ra[rr]) ? (TM)v : (TM)0;
            }
#pragma unroll
            for (int cc = 0; cc < CT; cc++) {
                const TB v = Be[kb + jb[cc]];
                fb[buf][cc] = (kok && cb[cc]) ? (TM)v : (TM)0;
            }
        };
        load(0, 0);
        for (int ks = 0; ks < nks; ks += 2) {
            if (ks + 1 < nks) load(ks + 1, 1);
#pragma unroll
            for (int rr = 0; rr < RT; rr++)
#pragma unroll
                for (int cc = 0; cc < CT; cc++) acc[rr][cc] = F::mfma(fa[0][rr], fb[0][cc], acc[rr][cc]);
            if (ks + 1 < nks) {
                if (ks + 2 < nks) load(ks + 2, 0);
#pragma unroll
                for (int rr = 0; rr < RT; rr++)
#pragma unroll
                    for (int cc = 0; cc < CT; cc++) acc[rr][cc] = F::mfma(fa[1][rr], fb[1][cc], acc[rr][cc]);
            }
        }
        double *Ce = g.C + e * g.sc;
        const double *C0e = g.C0 ? g.C0 + e * g.sc0 : nullptr;
        unsigned long long dm = 0, cm = 0;
#pragma unroll
        for (int rr = 0; rr < RT; rr++)
#pragma unroll
            for (int cc = 0; cc < CT; cc++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int i = 16 * (br * RT + rr) + F::row(lk, v), j = 16 * (bc * CT + cc) + lr;
                    if (i < g.M && j < g.N) {
                        const double d = g.alpha * (double)acc[rr][cc][v];
                        const double y = C0e ? fma(g.beta, C0e[(long)i * g.ldc0 + j], d) : d;
                        Ce[(long)i * g.ldc + j] = y;
                        const unsigned long long bd = (unsigned long long)__double_as_longlong(fabs(d));
                        const unsigned long long by = (unsigned long long)__double_as_longlong(fabs(y));
                        dm = bd > dm ? bd : dm;
                        cm = by > cm ? by : cm;
                    }
                }
        if (g.dmax) {
            dm = wave_max_bits(dm);
            cm = wave_max_bits(cm);
            if (lane == 0) {
                atomicMax(g.dmax + t, dm);
                atomicMax(g.cmax + t, cm);
            }
        }
    }
}

template <typename TM, typename TA, typename TB, int RT, int CT>
static hipError_t bgemm_direct_go(const BgemmArgs &g, hipStream_t s)
{
    const int TR = (g.M + 15) / 16, TC = (g.N + 15) / 16;
    const int nblk = ((TR + RT - 1) / RT) * ((TC + CT - 1) / CT), nwi = (nblk + 3) / 4;
    const long items = (long)((g.n + 7) / 8) * 8 * nwi;
    const int grid = (int)std::min<long>(items, 1 << 20);
    hipLaunchKernelGGL((bgemm_direct_kernel<TM, TA, TB, RT, CT>), dim3(grid), dim3(256), 0, s, g);
    return hipGetLastError();
}

// Candidate launch shapes (waves, tiles per wave, K chunk).  The tuner times the
// ones that fit a shape; variant 0 is the default.
struct BgemmShape {
    int W, TPW, KC, PF;   // waves, tiles per wave, K chunk, staged values per thread
};
static constexpr BgemmShape kShapes[] = {{8, 22, 8, 7}, {8, 12, 8, 6}, {4, 24, 8, 10}, {16, 11, 8, 4}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

// Direct (LDS-free) variants follow the staged ones: wave blocks of RT x CT tiles.
struct DirectShape {
    int RT, CT;
};
static constexpr DirectShape kDirect[] = {{2, 2}, {1, 2}, {2, 4}};
constexpr int kNumDirect = sizeof(kDirect) / sizeof(kDirect[0]);

int bgemm_num_variants() { return kNumShapes + kNumDirect; }

// Strips needed for variant v at M x N (0 if the variant cannot run the shape).
static int bgemm_strips(int v, int M, int N, int esize, size_t maxlds)
{
    const BgemmShape &sh = kShapes[v];
    const int TR = (M + 15) / 16, TC = (N + 15) / 16;
    for (int ns = 1; ns <= TR; ns++) {
        const int rows = (TR + ns - 1) / ns;
        if (rows * TC > sh.W * sh.TPW) continue;
        const int MS = 16 * rows, NS = 16 * TC;
        if (sh.KC * (MS + NS) > sh.PF * 64 * sh.W) continue;   // staged values per thread
        const size_t lds = 2 * (size_t)sh.KC * (bgemm_ld(MS, esize) + bgemm_ld(NS, esize)) * esize +
                           8 * (size_t)sh.W * sh.TPW;
        if (lds > maxlds) continue;
        return ns;
    }
    return 0;
}

static size_t bgemm_lds(int v, int M, int N, int ns, int esize)
{
    const BgemmShape &sh = kShapes[v];
    const int TR = (M + 15) / 16, TC = (N + 15) / 16, rows = (TR + ns - 1) / ns;
    return 2 * (size_t)sh.KC * (bgemm_ld(16 * rows, esize) + bgemm_ld(16 * TC, esize)) * esize +
           8 * (size_t)sh.W * sh.TPW;
}

template <typename TM, typename TA, typename TB, int W, int TPW, int KC, int PF>
static hipError_t bgemm_go(const BgemmArgs &g, size_t lds, hipStream_t s)
{
    auto k = bgemm_kernel<TM, TA, TB, W, TPW, KC, PF>;
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const long items = (long)((g.n + 7) / 8) * 8 * g.nstrip;
    const int grid = (int)std::min<long>(items, 65536);
    hipLaunchKernelGGL(k, dim3(grid), dim3(64 * W), lds, s, g);
    return hipGetLastError();
}

// mode 0: FP64 MFMA with A, B double; mode 1: FP32 MFMA with A float, B double.
hipError_t launch_bgemm(int mode, int variant, BgemmArgs g, size_t maxlds, hipStream_t s)
{
    if (variant < 0 || variant >= kNumShapes + kNumDirect) return hipErrorInvalidValue;
    if (g.n <= 0) return hipSuccess;
    if (variant >= kNumShapes) {
        g.nstrip = 1;
#define HFEM_BD(V)                                                                                                 \
    case V:                                                                                                        \
        return mode == 0 ? bgemm_direct_go<double, double, double, kDirect[V].RT, kDirect[V].CT>(g, s)             \
                         : bgemm_direct_go<float, float, double, kDirect[V].RT, kDirect[V].CT>(g, s);
        switch (variant - kNumShapes) {
            HFEM_BD(0) HFEM_BD(1) HFEM_BD(2)
        }
#undef HFEM_BD
        return hipErrorInvalidValue;
    }
    const int esize = mode == 0 ? 8 : 4;
    const int ns = bgemm_strips(variant, g.M, g.N, esize, maxlds);
    if (ns == 0) return hipErrorInvalidConfiguration;
    g.nstrip = ns;
    const size_t lds = bgemm_lds(variant, g.M, g.N, ns, esize);
#define HFEM_BG(V)                                                                                                 \
    case V:                                                                                                        \
        return mode == 0 ? bgemm_go<double, double, double, kShapes[V].W, kShapes[V].TPW, kShapes[V].KC, kShapes[V].PF>(g, lds, s) \
                         : bgemm_go<float, float, double, kShapes[V].W, kShapes[V].TPW, kShapes[V].KC, kShapes[V].PF>(g, lds, s);
    switch (variant) {
        HFEM_BG(0) HFEM_BG(1) HFEM_BG(2) HFEM_BG(3)
    }
#undef HFEM_BG
    return hipErrorInvalidValue;
}

// Whether variant v can run an M x N product (strips fit registers and LDS).
bool bgemm_fits(int mode, int variant, int M, int N, size_t maxlds)
{
    if (variant >= kNumShapes && variant < kNumShapes + kNumDirect) return M > 0 && N > 0;
    if (variant < 0 || variant >= kNumShapes) return false;
    return bgemm_strips(variant, M, N, mode == 0 ? 8 : 4, maxlds) > 0;
}

}  // namespace hfem
