// Global trace system on the device: BSR assembly of the condensed element
// blocks (assembly/globalassembly.cpp:8-161 mapElement2LinearSystem, HDG
// branch), right-hand sides, residual norm and the BSR matrix-vector product
// (spec kernel/matvec.m:1-12).
//
// Assembly is written in gather form: each BSR block (row face i, column face j)
// lists the (element, local face a, local face b) triples that touch it (at most
// two), so every block entry is summed by one thread in a fixed order -- no
// atomics, deterministic results.  Block storage is column-major bsz x bsz, as
// in the reference (Hg[k*bsz^2 + col*bsz + row]).
#include "hfem_device.h"

namespace hfem {

// einv[e*ndf + a*npf + m] = element-local node of face a that sits at canonical
// position m of the global face (the inverse of elcon - face*npf).
__global__ void bsr_gather_kernel(int nblocks, int bsz, int nch, int npf, int nb, long estride,
                                  const int *__restrict__ blkptr, const int *__restrict__ blklist,
                                  const int *__restrict__ einv, const double *__restrict__ AH, double *__restrict__ Hg)
{
    const long total = (long)nblocks * bsz * bsz;
    const int ndf = nb / nch;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const int k = idx / (bsz * bsz);
        const int rc = idx % (bsz * bsz), row = rc % bsz, col = rc / bsz;
        const int m1 = row / nch, n1 = row % nch, m2 = col / nch, n2 = col % nch;
        double s = 0.0;
        for (int q = blkptr[k]; q < blkptr[k + 1]; q++) {
            const int e = blklist[3 * q], a = blklist[3 * q + 1], b = blklist[3 * q + 2];
            const int li = einv[(long)e * ndf + a * npf + m1], lj = einv[(long)e * ndf + b * npf + m2];
            const long r = a * bsz + li * nch + n1, c = b * bsz + lj * nch + n2;
            s += AH[(long)e * estride + c * nb + r];
        }
        Hg[idx] = s;
    }
}

// Rg/r[i*bsz + m*nch + n] = sum over the (<=2) elements of face i of the element vector.
__global__ void face_gather_kernel(int nf, int bsz, int nch, int npf, int nb, long estride,
                                   const int *__restrict__ f2e, const int *__restrict__ einv,
                                   const double *__restrict__ V1, double *__restrict__ out1,
                                   const double *__restrict__ V2, double *__restrict__ out2)
{
    const long total = (long)nf * bsz;
    const int ndf = nb / nch;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const int i = idx / bsz, r = idx % bsz, m = r / nch, nn = r % nch;
        double s1 = 0.0, s2 = 0.0;
        for (int side = 0; side < 2; side++) {
            const int e = f2e[4 * i + 2 * side], a = f2e[4 * i + 2 * side + 1];
            if (e < 0) continue;
            const int li = einv[(long)e * ndf + a * npf + m];
            const long off = (long)e * estride + a * bsz + li * nch + nn;
            s1 += V1[off];
            if (V2) s2 += V2[off];
        }
        out1[idx] = s1;
        if (out2) out2[idx] = s2;
    }
}

// Partial sums of squares; partial[blockIdx.x] (deterministic two-stage reduction).
__global__ void sumsq_kernel(long n, const double *__restrict__ x, int square, double *__restrict__ partial)
{
    __shared__ double red[256];
    double s = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        s += square ? x[i] * x[i] : x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// y = Hg x with one row block per workgroup of four waves.  Blocks are stored
// column-major and a row's blocks are contiguous, so the row is one run of
// nblk*bsz columns of bsz doubles each: column q starts at Hg + (k0*bsz + q)*bsz.
// The four waves deal the columns among themselves (wave w takes q = w, w+4, ...),
// lane r reads entry r of a column (r + 64 for the second chunk when bsz > 64)
// and the wave-uniform x entry of that column (the wave index goes through
// readfirstlane so that column bookkeeping, cols[] and x[] are scalar loads; as
// vector loads they cost a third of the bandwidth); eight columns are in flight
// per wave, read nontemporal (the matrix is streamed once per product).  The four
// partial rows are summed through LDS.
template <int RCH>
__global__ __launch_bounds__(256) void bsr_matvec_rows_kernel(int bsz, const int *__restrict__ rowstart,
                                                              const int *__restrict__ cols,
                                                              const double *__restrict__ Hg,
                                                              const double *__restrict__ x, double *__restrict__ y)
{
    __shared__ double part[4][64 * RCH];
    const int i = blockIdx.x, lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int k0 = rowstart[i];
    const long ncol = (long)(rowstart[i + 1] - k0) * bsz;
    const double *base = Hg + (long)k0 * bsz * bsz;
    bool live[RCH];
    double acc[RCH];
#pragma unroll
    for (int h = 0; h < RCH; h++) {
        live[h] = lane + 64 * h < bsz;
        acc[h] = 0.0;
    }
    // (kq, cq): block within the row and column within the block of column q
    int kq = 0, cq = w;
    while (cq >= bsz) cq -= bsz, kq++;
    long q = w;
    auto step = [&]() {
        cq += 4;
        while (cq >= bsz) cq -= bsz, kq++;
    };
    for (; q + 28 < ncol; q += 32) {
        double v[8][RCH], xv[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            xv[u] = x[(long)cols[k0 + kq] * bsz + cq];
            const double *col = base + (q + 4 * u) * bsz + lane;
#pragma unroll
            for (int h = 0; h < RCH; h++) v[u][h] = live[h] ? __builtin_nontemporal_load(col + 64 * h) : 0.0;
            step();
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
#pragma unroll
            for (int h = 0; h < RCH; h++) acc[h] += v[u][h] * xv[u];
    }
    for (; q < ncol; q += 4) {
        const double xq = x[(long)cols[k0 + kq] * bsz + cq];
        const double *col = base + q * bsz + lane;
#pragma unroll
        for (int h = 0; h < RCH; h++)
            if (live[h]) acc[h] += __builtin_nontemporal_load(col + 64 * h) * xq;
        step();
    }
#pragma unroll
    for (int h = 0; h < RCH; h++) part[w][lane + 64 * h] = acc[h];
    __syncthreads();
    for (int r = threadIdx.x; r < bsz; r += 256)
        y[(long)i * bsz + r] = (part[0][r] + part[1][r]) + (part[2][r] + part[3][r]);
}

// y = Hg x: one thread per output entry of a row block (blocks wider than 128).
__global__ void bsr_matvec_kernel(int nrow, int bsz, const int *__restrict__ rowstart, const int *__restrict__ cols,
                                  const double *__restrict__ Hg, const double *__restrict__ x, double *__restrict__ y)
{
    const long total = (long)nrow * bsz;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const int i = idx / bsz, r = idx % bsz;
        double s = 0.0;
        for (int k = rowstart[i]; k < rowstart[i + 1]; k++) {
            const double *blk = Hg + (long)k * bsz * bsz + r;
            const double *xj = x + (long)cols[k] * bsz;
            for (int c = 0; c < bsz; c++) s += blk[(long)c * bsz] * xj[c];
        }
        y[idx] = s;
    }
}

static int grid_for(long total, int block = 256, int cap = 65536)
{
    long g = (total + block - 1) / block;
    return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

hipError_t launch_bsr_gather(int nblocks, int bsz, int nch, int npf, int nb, long estride, const int *blkptr,
                             const int *blklist, const int *einv, const double *AH, double *Hg, hipStream_t s)
{
    hipLaunchKernelGGL(bsr_gather_kernel, dim3(grid_for((long)nblocks * bsz * bsz)), dim3(256), 0, s, nblocks, bsz,
                       nch, npf, nb, estride, blkptr, blklist, einv, AH, Hg);
    return hipGetLastError();
}

hipError_t launch_face_gather(int nf, int bsz, int nch, int npf, int nb, long estride, const int *f2e,
                              const int *einv, const double *V1, double *o1, const double *V2, double *o2, hipStream_t s)
{
    hipLaunchKernelGGL(face_gather_kernel, dim3(grid_for((long)nf * bsz)), dim3(256), 0, s, nf, bsz, nch, npf, nb,
                       estride, f2e, einv, V1, o1, V2, o2);
    return hipGetLastError();
}

hipError_t launch_sumsq(long n, const double *x, int square, double *partial, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(sumsq_kernel, dim3(nblocks), dim3(256), 0, s, n, x, square, partial);
    return hipGetLastError();
}

hipError_t launch_bsr_matvec(int nrow, int bsz, const int *rowstart, const int *cols, const double *Hg,
                             const double *x, double *y, hipStream_t s)
{
    if (nrow <= 0) return hipSuccess;
    if (bsz <= 64)
        hipLaunchKernelGGL(bsr_matvec_rows_kernel<1>, dim3(nrow), dim3(256), 0, s, bsz, rowstart, cols, Hg, x, y);
    else if (bsz <= 128)
        hipLaunchKernelGGL(bsr_matvec_rows_kernel<2>, dim3(nrow), dim3(256), 0, s, bsz, rowstart, cols, Hg, x, y);
    else
        hipLaunchKernelGGL(bsr_matvec_kernel, dim3(grid_for((long)nrow * bsz)), dim3(256), 0, s, nrow, bsz, rowstart,
                           cols, Hg, x, y);
    return hipGetLastError();
}

}  // namespace hfem
