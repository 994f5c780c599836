// Element assembly for the compressible-NS HDG element path (gfx950).
//
// One workgroup assembles one element at a time (grid-stride over elements) and
// writes that element's D, F, K, H, Ru, Rh blocks (assembly/elementassembly.cpp:1481-1499
// and the faceint/elementint outputs).  D and K use the reference's column-major
// layouts; F is written row-major into Fx (na x ldx) with Ru as column nb, and H
// column-major into Hx (nb x (nb+1)) with Rh as column nb -- the layouts the
// condensation GEMMs stream.  Everything an element needs lives in LDS; global
// memory holds only the inputs, the pointwise fluxes and the outputs.
//
// Phases (barriers between them; inside a phase threads own disjoint outputs):
//   P0 element data (the master tables are loaded once per workgroup)
//   P1 volume/face geometry
//   P2 q operators M, C, E (qint) and the solution at volume points
//   P3 M^{-1} by Gauss-Jordan (one wave, in registers) next to C
//   P4 Q = [shapvt; face traces] M^{-1}: the point values of M^{-1}C and M^{-1}E
//      are Q C and Q E, never formed as tables
//   P8 element rows of D and F on the FP64 MFMA, one row block j at a time
//   P9 Ru and Rh
//   P9k trace rows of K and H on the FP64 MFMA, one face at a time
// (the pointwise fluxes and their Jacobians come from flux_kernel, launched before
// this kernel on the same element chunk)
#include "asm_body.h"

#include <algorithm>
#include <type_traits>

namespace hfem {

template <int ND, int NPV, int NPF, bool SPILL>
__global__ void __launch_bounds__(ASM_THREADS)
assemble_kernel(Dims d, Master m, Phys ph, const double *__restrict__ dgnodes, const double *__restrict__ UDG,
                const double *__restrict__ UH, const double *__restrict__ SH, const int *__restrict__ elcon,
                const int *__restrict__ bf, int e0, int n, int ldx, const double *__restrict__ fluxbuf, long fstride,
                double *__restrict__ Ru, double *__restrict__ Dm, double *__restrict__ Fx, double *__restrict__ Km,
                double *__restrict__ Hx, double *__restrict__ Rh, double *__restrict__ ru2, double *__restrict__ spillbuf,
                int res)
{
    StrideSource src;
    assemble_body<ND, NPV, NPF, SPILL>(d, m, ph, dgnodes, UDG, UH, SH, elcon, bf, e0, n, ldx, fluxbuf, fstride, Ru,
                                       Dm, Fx, Km, Hx, Rh, ru2, spillbuf, res, blockIdx.x, src);
}

template <int ND, int NPV, int NPF>
__global__ void __launch_bounds__(256)
flux_kernel(Dims d, Master m, Phys ph, const double *__restrict__ dgnodes, const double *__restrict__ UDG,
            const double *__restrict__ UH, const int *__restrict__ elcon, const int *__restrict__ bf, int e0, int n,
            double *__restrict__ fluxbuf, long fstride, int res)
{
    StrideSource src;
    flux_body<ND, NPV, NPF>(d, m, ph, dgnodes, UDG, UH, elcon, bf, e0, n, fluxbuf, fstride, res, src);
}


// getQ (assembly/elementassembly.cpp:1414-1479): q = M^{-1}(M sh_q + C u - E uh) for
// every element of [e0, e0+n), written into the q components of UDG.  C u and E uh
// are contracted at the quadrature points without forming C and E.  One
// workgroup per element.
template <int ND, int NPV, int NPF>
__global__ void __launch_bounds__(256)
getq_kernel(Dims d, Master m, Phys ph, const double *__restrict__ dgnodes, double *__restrict__ UDG,
            const double *__restrict__ UH, const double *__restrict__ SH, const int *__restrict__ elcon,
            const int *__restrict__ bf, int e0, int n)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x, nth = blockDim.x;
    constexpr int NCU = ND + 2, NC = NCU * (ND + 1);
    const int npv = NPV, ngv = d.ngv, npf = NPF, ngf = d.ngf, nfe = d.nfe, ndf = d.ndf, np = d.np;
    double *p = lds;
    double *pn = p; p += npv * ND;
    double *u = p; p += npv * NCU;
    double *uh = p; p += ndf * NCU;            // uh[k*ndf + j]
    double *jac = p; p += ngv;
    double *Xx = p; p += ngv * ND * ND;
    double *jacf = p; p += np;
    double *nl = p; p += np * ND;
    double *M = p; p += npv * npv;
    double *Mi = p; p += npv * npv;
    double *rowk = p; p += npv;
    double *colk = p; p += npv;
    double *ug = p; p += ngv * NCU;            // [g + ngv*c]
    double *uhg = p; p += np * NCU;            // [pp + np*c]
    double *rhs = p; p += npv * NCU * ND;      // [n + npv*(c + NCU*x)]
#define SVT(g, nn, dd) m.shapvt[(g) + ngv * ((nn) + npv * (dd))]
#define SVG(i, g, dd) m.shapvg[(i) + npv * ((g) + ngv * (dd))]
#define SFT(g, k, dd) m.shapft[(g) + ngf * ((k) + npf * (dd))]
#define SFG(k, g, dd) m.shapfg[(k) + npf * ((g) + ngf * (dd))]
#define PERM(k, a) m.perm[(k) + npf * (a)]
#define INVP(nn, a) m.invperm[(nn) + npv * (a)]
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const long e = e0 + t;
        double *ue = UDG + e * npv * NC;
        const double *she = SH + e * npv * NC;
        for (int i = tid; i < npv * ND; i += nth) pn[i] = dgnodes[e * npv * ND + i];
        for (int i = tid; i < npv * NCU; i += nth) u[i] = ue[i];
        for (int i = tid; i < ndf * NCU; i += nth) {
            const int j = i / NCU, k = i % NCU;
            uh[k * ndf + j] = UH[(long)elcon[e * ndf + j] * NCU + k];
        }
        __syncthreads();
        // geometry (elementgeom :74-124, facegeom :206-232)
        for (int i = tid; i < ngv + np; i += nth) {
            if (i < ngv) {
                const int g = i;
                double adj[ND][ND];
                const double det = volume_adjugate<ND>([&](int k, int r) { return SVT(g, k, 1 + r); }, pn, npv, adj);
                jac[g] = det;
                for (int x = 0; x < ND; x++)
                    for (int r = 0; r < ND; r++) Xx[g + ngv * (x + ND * r)] = adj[r][x];
            } else {
                const int pp = i - ngv, g = pp % ngf, a = pp / ngf;
                double tg[2][3];
                for (int r = 0; r < ND - 1; r++)
                    for (int x = 0; x < ND; x++) {
                        double s = 0;
                        for (int k = 0; k < npf; k++) s += SFT(g, k, 1 + r) * pn[x * npv + PERM(k, a)];
                        tg[r][x] = s;
                    }
                double nv[3];
                if (ND == 2) { nv[0] = tg[0][1]; nv[1] = -tg[0][0]; nv[2] = 0; }
                else {
                    nv[0] = tg[0][1] * tg[1][2] - tg[0][2] * tg[1][1];
                    nv[1] = tg[0][2] * tg[1][0] - tg[0][0] * tg[1][2];
                    nv[2] = tg[0][0] * tg[1][1] - tg[0][1] * tg[1][0];
                }
                double len = 0;
                for (int x = 0; x < ND; x++) len += nv[x] * nv[x];
                len = sqrt(len);
                jacf[pp] = len;
                for (int x = 0; x < ND; x++) nl[pp + np * x] = nv[x] / len;
            }
        }
        __syncthreads();
        // mass matrix (qint :1374) and the traces of u and uh
        for (int i = tid; i < npv * npv + ngv * NCU + np * NCU; i += nth) {
            if (i < npv * npv) {
                const int jj = i % npv, ii = i / npv;
                double s = 0;
                for (int g = 0; g < ngv; g++) s += (SVG(jj, g, 0) * SVT(g, ii, 0)) * jac[g];
                M[i] = ph.fc_q * s;
                Mi[i] = M[i];
            } else if (i < npv * npv + ngv * NCU) {
                const int r = i - npv * npv, g = r % ngv, c = r / ngv;
                double s = 0;
                for (int q = 0; q < npv; q++) s += SVT(g, q, 0) * u[c * npv + q];
                ug[r] = s;
            } else {
                const int r = i - npv * npv - ngv * NCU, pp = r % np, c = r / np, g = pp % ngf, a = pp / ngf;
                double s = 0;
                for (int q = 0; q < npf; q++) s += SFT(g, q, 0) * uh[c * ndf + a * npf + q];
                uhg[r] = s;
            }
        }
        __syncthreads();
        // rhs = M sh_q + C u - E uh
        for (int i = tid; i < npv * NCU * ND; i += nth) {
            const int nn = i % npv, r = i / npv, c = r % NCU, x = r / NCU;
            double s = 0;
            const double *shq = she + (long)(NCU + x * NCU + c) * npv;
            for (int l = 0; l < npv; l++) s += M[nn + npv * l] * shq[l];
            for (int g = 0; g < ngv; g++) {
                double cx = 0;
                for (int q = 0; q < ND; q++) cx += SVG(nn, g, q + 1) * Xx[g + ngv * (x + ND * q)];
                s += cx * ug[g + ngv * c];
            }
            for (int a = 0; a < nfe; a++) {
                const int k = INVP(nn, a);
                if (k < 0) continue;
                for (int g = 0; g < ngf; g++) {
                    const int pp = g + ngf * a;
                    s -= SFG(k, g, 0) * uhg[pp + np * c] * nl[pp + np * x] * jacf[pp];
                }
            }
            rhs[i] = s;
        }
        // M^{-1} by Gauss-Jordan (M is SPD)
        for (int k = 0; k < npv; k++) {
            __syncthreads();
            for (int i = tid; i < 2 * npv; i += nth) {
                if (i < npv) rowk[i] = Mi[k + npv * i];
                else colk[i - npv] = Mi[(i - npv) + npv * k];
            }
            __syncthreads();
            const double pinv = 1.0 / rowk[k];
            for (int idx = tid; idx < npv * npv; idx += nth) {
                const int i = idx % npv, j = idx / npv;
                double v;
                if (i == k) v = (j == k) ? pinv : rowk[j] * pinv;
                else if (j == k) v = -colk[i] * pinv;
                else v = Mi[idx] - colk[i] * rowk[j] * pinv;
                Mi[idx] = v;
            }
        }
        __syncthreads();
        for (int i = tid; i < npv * NCU * ND; i += nth) {
            const int nn = i % npv, r = i / npv;
            double s = 0;
            for (int l = 0; l < npv; l++) s += Mi[nn + npv * l] * rhs[l + npv * r];
            ue[NCU * npv + i] = s;
        }
        __syncthreads();
    }
#undef SVT
#undef SVG
#undef SFT
#undef SFG
#undef PERM
#undef INVP
}

size_t getq_lds_bytes(const Dims &d)
{
    const size_t n = (size_t)d.npv * d.nd + (size_t)d.npv * d.ncu + (size_t)d.ndf * d.nch + d.ngv +
                     (size_t)d.ngv * d.nd * d.nd + d.np + (size_t)d.np * d.nd + 2 * (size_t)d.npv * d.npv +
                     2 * (size_t)d.npv + (size_t)d.ngv * d.ncu + (size_t)d.np * d.nch + (size_t)d.npv * d.ncu * d.nd;
    return n * sizeof(double);
}

size_t flux_lds_bytes(const Dims &d)
{
    size_t n = (size_t)d.npv * d.nd + (size_t)d.npv * d.nc + (size_t)d.ndf * d.nch + (size_t)d.np * d.nd +
               (size_t)d.ngv * d.nc + (size_t)d.np * d.nc + (size_t)d.np * d.nch + (size_t)d.ngv * d.nd * d.nd;
    return n * sizeof(double) + 16 * sizeof(int);
}

size_t assemble_lds_bytes(const Dims &d)
{
    const AsmLayout L = asm_layout(d);
    return (size_t)L.total * sizeof(double) + (size_t)L.nints * sizeof(int);
}

// element families with compiled kernels: (nd, npv, npf)
#define HFEM_FAMILIES(X)                                                            \
    X(2, 3, 2) X(2, 6, 3) X(2, 10, 4) X(2, 15, 5) X(2, 21, 6)                       \
    X(3, 4, 3) X(3, 10, 6) X(3, 20, 10) X(3, 35, 15) X(3, 8, 4) X(3, 27, 9)

// families whose element data exceed the LDS: a spill-placement instantiation as well
#define HFEM_SPILL_FAMILIES(X) X(3, 35, 15)

// true when a kernel is compiled for the family in the placement its sizes need
bool assemble_supported(const Dims &d)
{
    const bool spill = asm_layout(d).spill != 0;
#define CHECK(a, b, c) if (!spill && d.nd == a && d.npv == b && d.npf == c) return true;
    HFEM_FAMILIES(CHECK)
#undef CHECK
#define CHECK(a, b, c) if (spill && d.nd == a && d.npv == b && d.npf == c) return true;
    HFEM_SPILL_FAMILIES(CHECK)
#undef CHECK
    return false;
}

// doubles per workgroup of the spill placement's global region (0 if none)
long assemble_spill_doubles(const Dims &d)
{
    const AsmLayout L = asm_layout(d);
    return L.spill ? L.stotal : 0;
}

hipError_t launch_assemble(const Dims &d, const Master &m, const Phys &ph, const double *dgnodes, const double *UDG,
                           const double *UH, const double *SH, const int *elcon, const int *bf, int e0, int n,
                           int grid, int ldx, double *fluxbuf, long fstride, int chunk,
                           double *Ru, double *D, double *Fx, double *K, double *Hx, double *Rh, double *ru2,
                           double *spill, hipStream_t stream, int res)
{
    const size_t lds = assemble_lds_bytes(d), flds = flux_lds_bytes(d);
    const bool sp = asm_layout(d).spill != 0;
    if (sp && !spill) return hipErrorInvalidValue;
    for (int c0 = e0; c0 < e0 + n; c0 += chunk) {
        const int cn = std::min(chunk, e0 + n - c0);
        const int fgrid = std::min(cn, 2048), agrid = std::min(cn, grid);
        bool done = false;
#define FLUX(a, b, c)                                                                                               \
        if (!done && d.nd == a && d.npv == b && d.npf == c) {                                                       \
            hipFuncSetAttribute((const void *)flux_kernel<a, b, c>, hipFuncAttributeMaxDynamicSharedMemorySize,    \
                                (int)flds);                                                                         \
            hipLaunchKernelGGL((flux_kernel<a, b, c>), dim3(fgrid), dim3(256), flds, stream, d, m, ph, dgnodes,     \
                               UDG, UH, elcon, bf, c0, cn, fluxbuf, fstride, res);                                  \
            done = true;                                                                                            \
        }
        HFEM_FAMILIES(FLUX)
#undef FLUX
        if (!done) return hipErrorInvalidValue;
        done = false;
#define LAUNCH(a, b, c, S)                                                                                          \
        if (!done && sp == S && d.nd == a && d.npv == b && d.npf == c) {                                            \
            hipFuncSetAttribute((const void *)assemble_kernel<a, b, c, S>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                (int)lds);                                                                          \
            hipLaunchKernelGGL((assemble_kernel<a, b, c, S>), dim3(agrid), dim3(ASM_THREADS), lds, stream, d, m, ph,\
                               dgnodes, UDG, UH, SH, elcon, bf, c0, cn, ldx, fluxbuf, fstride, Ru, D, Fx, K, Hx,   \
                               Rh, ru2, spill, res);                                                                \
            done = true;                                                                                            \
        }
#define LAUNCH_LDS(a, b, c) LAUNCH(a, b, c, false)
#define LAUNCH_SPILL(a, b, c) LAUNCH(a, b, c, true)
        HFEM_FAMILIES(LAUNCH_LDS)
        HFEM_SPILL_FAMILIES(LAUNCH_SPILL)
#undef LAUNCH_SPILL
#undef LAUNCH_LDS
#undef LAUNCH
        if (!done) return hipErrorInvalidValue;
        hipError_t err = hipGetLastError();
        if (err != hipSuccess) return err;
    }
    return hipSuccess;
}

hipError_t launch_getq(const Dims &d, const Master &m, const Phys &ph, const double *dgnodes, double *UDG,
                       const double *UH, const double *SH, const int *elcon, const int *bf, int e0, int n,
                       hipStream_t stream)
{
    const size_t lds = getq_lds_bytes(d);
    const int grid = std::min(n, 4096);
#define LAUNCH(a, b, c)                                                                                             \
    if (d.nd == a && d.npv == b && d.npf == c) {                                                                    \
        hipFuncSetAttribute((const void *)getq_kernel<a, b, c>, hipFuncAttributeMaxDynamicSharedMemorySize,         \
                            (int)lds);                                                                              \
        hipLaunchKernelGGL((getq_kernel<a, b, c>), dim3(grid), dim3(256), lds, stream, d, m, ph, dgnodes, UDG, UH,  \
                           SH, elcon, bf, e0, n);                                                                   \
        return hipGetLastError();                                                                                   \
    }
    HFEM_FAMILIES(LAUNCH)
#undef LAUNCH
    return hipErrorInvalidValue;
}

}  // namespace hfem
