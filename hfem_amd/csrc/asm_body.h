// Device bodies of the element assembly (assembly/elementassembly.cpp) and of the
// pointwise flux evaluation that feeds it, shared by the per-family kernels and
// the ragged batch kernels: LDS layouts, the per-element phases, and the element
// sources that hand a workgroup its elements.
#pragma once
#include "hfem_device.h"

#include <algorithm>
#include <type_traits>

namespace hfem {

constexpr int ASM_THREADS = 512;
constexpr int ASM_LDS_MAX = 160 * 1024;   // LDS per workgroup on gfx950

typedef double f64x4 __attribute__((ext_vector_type(4)));

// LDS layout of the assembly kernel in doubles (the int tables follow).  Master
// tables are transposed so that the thread-varying index is fastest:
// svt[n + npv*(g + ngv*dd)], svg[i + npv*(g + ngv*dd)] (as shapvg),
// sft[k + npf*(g + ngf*dd)], sfg[k + npf*(g + ngf*dd)] (as shapfg).
// The shapvt derivatives are dead after P1; their region then holds G (P2) and
// fu.Xx of one row block (P8).  The face operands of P9k (flux Jacobians of one
// face and its point values of M^{-1}C, M^{-1}E) get a region of their own when
// it fits, else they overlay the shapvt derivatives and shapvg, which P0 then
// reloads for the next element.
struct AsmLayout {
    int svt, svg, sft, sfg, pn, udg, uh, sh, jac, Xx, jacf, nl, M, Qm, C, E, udgg, fj, red;
    int xa, xha, cf, ef;
    int fuX2, fj2;          // second staging buffer of P8
    int total, nints, overlay, dbuf;
    int spill, stotal;      // spill: svt, svg, Qm, C, E in a per-workgroup global region of stotal doubles
};

// Offsets for one placement: spill = 0 puts everything in LDS, spill = 1 moves the
// large read-mostly tables (shapvt with its reused derivative region, shapvg, Q,
// C, E) to a per-workgroup region in global memory, for families whose element
// data exceed the LDS (p=4 tets).
__host__ __device__ inline AsmLayout asm_layout(const Dims &d, int spill)
{
    const int npv = d.npv, ngv = d.ngv, nd = d.nd, nc = d.nc, ncu = d.ncu, nch = d.nch, np = d.np;
    const int npf = d.npf, ngf = d.ngf, ndf = d.ndf, nd1 = d.nd + 1;
    AsmLayout L;
    int o = 0, s = 0;
    auto big = [&](int n) {
        int r;
        if (spill) { r = s; s += n; } else { r = o; o += n; }
        return r;
    };
    const int deriv = (npv > nc ? npv : nc) * ngv * nd;
    L.spill = spill;
    L.svt = big(npv * ngv + deriv);
    L.svg = big(npv * ngv * nd1);
    L.sft = o; o += npf * ngf * nd;
    L.sfg = o; o += npf * ngf * nd;
    L.pn = o; o += npv * nd;
    L.udg = o; o += npv * nc;
    L.uh = o; o += ndf * nch;
    L.sh = o; o += npv * ncu;
    L.jac = o; o += ngv;
    L.Xx = o; o += ngv * nd * nd;
    L.jacf = o; o += np;
    L.nl = o; o += np * nd;
    L.M = o; o += npv * npv;
    L.Qm = big((ngv + np) * (npv | 1));
    // Q rows and C_x, E_x columns at an odd stride (npv | 1): MFMA operand reads
    // walk them with the lane, and an even stride of doubles puts lanes 8 apart on
    // the same LDS bank
    const int cs = npv | 1;
    L.C = big(nd * cs * npv);
    L.E = big(nd * cs * ndf);
    L.udgg = o; o += ngv * nc;
    // fhu, fhuh of one row block (P8); the Ru/Rh tables (P9)
    const int fjn = np * (nc + nch), tn = ngv * ncu * nd1 + 2 * np * nch;
    L.fj = o;
    o += fjn > tn ? fjn : tn;
    L.red = o; o += 64;
    L.nints = npf * d.nfe + npv * d.nfe + d.nfe + 8;
    const int kh = ngf * nch * (nc + nch) + (npv + ndf) * ngf * nd;
    int base = o;
    L.overlay = 0;
    if (!spill && (long)(o + kh) * 8 + (long)L.nints * 4 > ASM_LDS_MAX && kh <= deriv + npv * ngv * nd1) {
        L.overlay = 1;
        base = L.svt + npv * ngv;
    } else {
        o += kh;
    }
    L.xa = base;
    L.xha = L.xa + ngf * nch * nc;
    L.cf = L.xha + ngf * nch * nch;
    L.ef = L.cf + npv * ngf * nd;
    const int sb = ngv * nd * nc + np * nc + np * nch;
    L.dbuf = (long)(o + sb) * 8 + (long)L.nints * 4 <= ASM_LDS_MAX;
    L.fuX2 = L.fj2 = 0;
    if (L.dbuf) {
        L.fuX2 = o;
        L.fj2 = o + ngv * nd * nc;
        o += sb;
    }
    L.total = o;
    L.stotal = s;
    return L;
}

// The placement used: everything in LDS when it fits, else the spill placement.
__host__ __device__ inline AsmLayout asm_layout(const Dims &d)
{
    const AsmLayout L = asm_layout(d, 0);
    if ((long)L.total * 8 + (long)L.nints * 4 <= ASM_LDS_MAX) return L;
    return asm_layout(d, 1);
}

// Volume geometry at one point: J(x,r) = sum_k dshape(k, r) pn[x*npv + k] with
// dshape(k, r) the derivative r of shape function k, and adj(J)(r,x), so that
// Xx(x,r) = adj[r][x] = jac d xi_r / d x_x.  Returns jac = det J.  Shared by the
// flux and assembly kernels so that both see the same Xx.
template <int ND, class DS>
__device__ __forceinline__ double volume_adjugate(DS dshape, const double *pn, int npv, double adj[ND][ND])
{
    double J[ND][ND];
    for (int x = 0; x < ND; x++)
        for (int r = 0; r < ND; r++) {
            double s = 0;
            for (int k = 0; k < npv; k++) s += dshape(k, r) * pn[x * npv + k];
            J[x][r] = s;
        }
    double det;
    if (ND == 2) {
        det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        adj[0][0] = J[1][1]; adj[0][1] = -J[0][1]; adj[1][0] = -J[1][0]; adj[1][1] = J[0][0];
    } else {
        for (int r = 0; r < ND; r++)
            for (int x = 0; x < ND; x++) {
                const int x1 = (x + 1) % 3, x2 = (x + 2) % 3, r1 = (r + 1) % 3, r2 = (r + 2) % 3;
                adj[r][x] = J[x1][r1] * J[x2][r2] - J[x1][r2] * J[x2][r1];
            }
        det = J[0][0] * adj[0][0] + J[1][0] * adj[0][1] + J[2][0] * adj[0][2];
    }
    return det;
}

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// In-place inverse of the NPV x NPV matrix M (column-major, M(i,j) at M[i + NPV*j])
// by Gauss-Jordan without pivoting, run by one wave: lane j holds column j in
// registers and each pivot column is broadcast with readlane, so no barrier is
// needed.  For symmetric positive definite M (the mass matrix).
template <int NPV>
__device__ __forceinline__ void gauss_jordan_wave(double *M, int lane)
{
    const int col = lane < NPV ? lane : 0;
    double m[NPV];
#pragma unroll
    for (int i = 0; i < NPV; i++) m[i] = M[i + NPV * col];
#pragma unroll
    for (int k = 0; k < NPV; k++) {
        double ck[NPV];                              // column k: M(i, k)
#pragma unroll
        for (int i = 0; i < NPV; i++) ck[i] = readlane_f64(m[i], k);
        const double pinv = 1.0 / ck[k];
        const double rk = m[k];                      // row k: M(k, col)
        const bool isk = lane == k;
#pragma unroll
        for (int i = 0; i < NPV; i++) {
            if (i == k) m[i] = isk ? pinv : rk * pinv;
            else m[i] = isk ? -ck[i] * pinv : m[i] - ck[i] * rk * pinv;
        }
    }
    if (lane < NPV)
#pragma unroll
        for (int i = 0; i < NPV; i++) M[i + NPV * lane] = m[i];
}

// The body of the assembly for one workgroup: elements t of [0, n) as the element
// source src hands them out (src.first(), then src.next(t) until a value >= n),
// element e0 + t at flux-buffer slot t; wg picks the workgroup's spill region.
template <int ND, int NPV, int NPF, bool SPILL, class Src>
__device__ __forceinline__ void
assemble_body(const Dims &d, const Master &m, const Phys &ph, const double *__restrict__ dgnodes,
              const double *__restrict__ UDG, const double *__restrict__ UH, const double *__restrict__ SH,
              const int *__restrict__ elcon, const int *__restrict__ bf, int e0, int n, int ldx,
              const double *__restrict__ fluxbuf, long fstride, double *__restrict__ Ru, double *__restrict__ Dm,
              double *__restrict__ Fx, double *__restrict__ Km, double *__restrict__ Hx, double *__restrict__ Rh,
              double *__restrict__ ru2, double *__restrict__ spillbuf, int res, int wg, Src &src)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x, nth = blockDim.x;
    const int lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    // wave-uniform: task and tile indices derived from it stay in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), nwave = nth >> 6;
    const FluxLayout FL = flux_layout(d);
    const AsmLayout LY = asm_layout(d, SPILL);
    constexpr int NCU = ND + 2, NC = NCU * (ND + 1), ND1 = ND + 1;
    // faces per element from the face size (tets: triangular faces, hexes: square faces)
    constexpr int NFE = ND == 2 ? 3 : ((NPF == 4 || NPF == 9 || NPF == 16) ? 6 : 4);
    constexpr int NDF = NPF * NFE;
    // 16-wide MFMA tiles: TD over the npv columns of D and K, TF over the ndf columns
    // of F and H, KT over the npv nodes contracted in the W products of P8
    constexpr int TD = (NPV + 15) / 16, TF = (NDF + 15) / 16, NT = TD + TF, KT = (NPV + 15) / 16;
    constexpr int CS = NPV | 1;           // row stride of Q, column stride of C_x and E_x (asm_layout)
    const int npv = NPV, ngv = d.ngv, npf = NPF, ngf = d.ngf, nfe = NFE, ndf = NDF;
    const int ncu = NCU, nch = NCU, nc = NC, na = d.na, nb = d.nb, nd1 = ND1, np = d.np;

    // the large tables: LDS, or this workgroup's global region in the spill placement
    double *big = SPILL ? spillbuf + (long)wg * LY.stotal : lds;
    double *svt = big + LY.svt, *svg = big + LY.svg, *sft = lds + LY.sft, *sfg = lds + LY.sfg;
    double *pn = lds + LY.pn, *udg = lds + LY.udg, *uh = lds + LY.uh, *sh = lds + LY.sh;
    double *jac = lds + LY.jac, *Xx = lds + LY.Xx, *jacf = lds + LY.jacf, *nl = lds + LY.nl;
    double *M = lds + LY.M, *Qm = big + LY.Qm;
    double *C = big + LY.C, *E = big + LY.E, *udgg = lds + LY.udgg, *fj = lds + LY.fj;
    double *red = lds + LY.red;
    double *xa = lds + LY.xa, *cf = lds + LY.cf, *ef = lds + LY.ef;
    double *G = svt + npv * ngv, *fuX = G;       // the shapvt-derivative region, reused
    int *perm = (int *)(lds + LY.total);
    int *invperm = perm + npf * nfe;
    int *bnd = invperm + npv * nfe;

    // master tables once per workgroup
    for (int i = tid; i < npv * ngv * nd1; i += nth) {
        const int nn = i % npv, r = i / npv, g = r % ngv, dd = r / ngv;
        svt[i] = m.shapvt[g + ngv * (nn + npv * dd)];
        svg[i] = m.shapvg[i];
    }
    for (int i = tid; i < npf * ngf * ND; i += nth) {
        const int k = i % npf, r = i / npf, g = r % ngf, dd = r / ngf;
        sft[i] = m.shapft[g + ngf * (k + npf * dd)];
        sfg[i] = m.shapfg[i];
    }
    for (int i = tid; i < npf * nfe; i += nth) perm[i] = m.perm[i];
    for (int i = tid; i < npv * nfe; i += nth) invperm[i] = m.invperm[i];
#define SVT(g, nn, dd) svt[(nn) + npv * ((g) + ngv * (dd))]
#define SVG(i, g, dd) svg[(i) + npv * ((g) + ngv * (dd))]
#define SFT(g, k, dd) sft[(k) + npf * ((g) + ngf * (dd))]
#define SFG(k, g, dd) sfg[(k) + npf * ((g) + ngf * (dd))]
#define PERM(k, a) perm[(k) + npf * (a)]
#define INVP(nn, a) invperm[(nn) + npv * (a)]

    bool fresh = true;
    for (int t = src.first(); t < n; t = src.next(t)) {
        const long e = e0 + t;
        const double *FB = fluxbuf + (long)t * fstride;
        const double *fX = FB + FL.fX, *FJ = FB + FL.J, *fh = FB + FL.fh;
        const double *fb = FB + FL.fb, *fbu = FB + FL.fbu, *fbuh = FB + FL.fbuh;
        const long ohu = (long)ngv * ND * NC, ohh = ohu + (long)np * NC;   // fhu, fhuh in a row block
        // ---------------- P0: gather ----------------
        // the previous element overwrote the shapvt derivatives (and, with the
        // overlay, shapvg)
        if (!fresh) {
            for (int i = npv * ngv + tid; i < npv * ngv * nd1; i += nth) {
                const int nn = i % npv, r = i / npv, g = r % ngv, dd = r / ngv;
                svt[i] = m.shapvt[g + ngv * (nn + npv * dd)];
            }
            if (LY.overlay)
                for (int i = tid; i < npv * ngv * nd1; i += nth) svg[i] = m.shapvg[i];
        }
        fresh = false;
        for (int i = tid; i < npv * ND; i += nth) pn[i] = dgnodes[e * npv * ND + i];
        for (int i = tid; i < npv * nc; i += nth) udg[i] = UDG[e * npv * nc + i];
        for (int i = tid; i < ndf * nch; i += nth) {
            int j = i / nch, k = i % nch;
            uh[k * ndf + j] = UH[(long)elcon[e * ndf + j] * nch + k];
        }
        if (ph.tdep)
            for (int i = tid; i < npv * ncu; i += nth) sh[i] = SH[e * npv * nc + i];
        for (int a = tid; a < nfe; a += nth) bnd[a] = bf[e * nfe + a] < 0;
        __syncthreads();

        // ---------------- P1: geometry ----------------
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

        // ---------------- P2: q operators (qint), solution at volume points ----------------
        // M, E, udgg and G(i,g,x) = sum_r shapvg(i,g,r+1) Xx(g,x,r), the jac-weighted
        // test-function gradients (C_x(i,k) = sum_g G(i,g,x) shapvt(g,k) follows in P3)
        {
            const int nG = npv * ngv * ND, nM = npv * npv, nU = ngv * nc, nE = ND * npv * ndf;
            // residual only: udgg is all the residual rows need
            const int i0 = res ? nG + nM : 0, i1 = res ? nG + nM + nU : nG + nM + nU + nE;
            for (int i = i0 + tid; i < i1; i += nth) {
                if (i < nG) {
                    const int nn = i % npv, r = i / npv, g = r % ngv, x = r / ngv;
                    double sg = 0;
#pragma unroll
                    for (int q = 0; q < ND; q++) sg += SVG(nn, g, q + 1) * Xx[g + ngv * (x + ND * q)];
                    G[i] = sg;
                } else if (i < nG + nM) {
                    // M(jj, ii) at M[jj + npv*ii]
                    const int r = i - nG, jj = r % npv, ii = r / npv;
                    double s = 0;
#pragma unroll 4
                    for (int g = 0; g < ngv; g++) s += (SVG(jj, g, 0) * SVT(g, ii, 0)) * jac[g];
                    M[r] = ph.fc_q * s;
                } else if (i < nG + nM + nU) {
                    const int r = i - nG - nM, g = r % ngv, c = r / ngv;
                    double s = 0;
                    for (int q = 0; q < npv; q++) s += SVT(g, q, 0) * udg[c * npv + q];
                    udgg[r] = s;
                } else {
                    // E_x(nn, l) at E[x*npv*ndf + l*npv + nn]
                    const int r = i - nG - nM - nU, x = r / (npv * ndf), rr = r % (npv * ndf), nn = rr % npv,
                              l = rr / npv;
                    const int a = l / npf, jf = l % npf, k = INVP(nn, a);
                    double s = 0;
                    if (k >= 0)
#pragma unroll 4
                        for (int g = 0; g < ngf; g++) {
                            const int pp = g + ngf * a;
                            s += (SFG(k, g, 0) * SFT(g, jf, 0)) * nl[pp + np * x] * jacf[pp];
                        }
                    E[x * CS * ndf + l * CS + nn] = s;
                }
            }
        }
        __syncthreads();

        // P3-P8 form the element matrices; the residual-only assembly
        // (assembleElementVector) goes from P2 straight to P9
        if (!res) {
        // ---------------- P3: M^{-1} by Gauss-Jordan (M is SPD: no pivoting), and C ----------------
        // Wave 0 inverts M in registers, one column per lane; the other waves form
        // C_x(jj, ii) at C[x*npv*npv + ii*npv + jj] meanwhile.
        if (wave == 0) {
            gauss_jordan_wave<NPV>(M, lane);
        } else {
            for (int r = tid - 64; r < ND * npv * npv; r += nth - 64) {
                const int x = r / (npv * npv), rr = r % (npv * npv), jj = rr % npv, ii = rr / npv;
                double s = 0;
#pragma unroll 4
                for (int g = 0; g < ngv; g++) s += G[jj + npv * (g + ngv * x)] * SVT(g, ii, 0);
                C[x * CS * npv + ii * CS + jj] = s;
            }
        }
        __syncthreads();

        // ---------------- P4: Q = [shapvt ; face traces] M^{-1} ----------------
        for (int idx = tid; idx < (ngv + np) * npv; idx += nth) {
            const int k = idx % npv, r = idx / npv;
            double s = 0;
            if (r < ngv) {
                for (int l = 0; l < npv; l++) s += SVT(r, l, 0) * M[l + npv * k];
            } else {
                const int pp = r - ngv, g = pp % ngf, a = pp / ngf;
                for (int q = 0; q < npf; q++) s += SFT(g, q, 0) * M[PERM(q, a) + npv * k];
            }
            Qm[k + CS * r] = s;
        }
        __syncthreads();

        // ---------------- P8: element rows of D and F on the FP64 MFMA ----------------
        // Row block j of D and F as one product over volume and face points.  The
        // q-folded volume integrand of elementint (:418-549), with
        // P_x'[(kc,i), g] = -sum_r shapvg(i,g,r+1) fuX(g, r, x' ncu + kc)
        // (+ fc_u jac shapvg(i,g,0) for x' = 0, kc = j, tdep) and
        // fuX(g,r,c) = sum_x fu(g,j,x,c) Xx(g,x,r), contributes
        //   D_j[(kc,i), cc] += sum_g P_0 shapvt(g,cc) + sum_{x,g} P_{x+1} (M^{-1}C_x)(g,cc)
        //   F_j[(kc,i), l]  += -sum_{x,g} P_{x+1} (M^{-1}E_x)(g,l);
        // the face integrand of faceint (:862-1072), the gather over the faces that
        // contain node i, is the same product over face points pp = (a, g) with
        //   A_c[(kc,i), pp] = [i on a] shapfg(k(i,a), g) jacf(pp) fhu(pp, j, c):
        //   D += A_kc [cc on a] shapft(g, k(cc,a)) + sum_x A_{(x+1)ncu+kc} (M^{-1}C_x)(pp,cc)
        //   F += A^h_kc [l on a] shapft(g, l) (fhuh in A^h) - sum_x A_{(x+1)ncu+kc} (M^{-1}E_x)(pp,l).
        // The point values of M^{-1}C_x are Q C_x (Q from P4), so the q columns are
        // contracted through the nodes: W_x = sum_g P_{x+1} Q(g,:) + sum_pp A_x Q(pp,:)
        // is accumulated as W_x^T (whose MFMA result layout is the A operand layout
        // of the next product), then D += sum_x W_x C_x, F -= sum_x W_x E_x.
        // One wave per (j, 16-row tile of (kc,i)), all column tiles of D (cc) and F (l)
        // accumulated together.  Operands are read at clamped indices and zeroed by
        // selects, never under a branch, so that the loads of a k-step issue together;
        // only zeros along the contracted index matter (rows and columns past the end
        // are never stored).
        {
            const int nrt = (ncu * npv + 15) / 16;
            // row block j of the flux Jacobians (one contiguous record of the flux kernel)
            // into uX[g + ngv*(r + ND*c)] = sum_x fu(g,j,x,c) Xx(g,x,r) and
            // hj[pp + np*c] = fhu(pp,j,c), then hj[np*NC + pp + np*kc] = fhuh(pp,j,kc), by
            // threads ti = 0 .. nt-1, U loads in flight per thread before the stores
            auto stage = [&](auto Ut, int j, int ti, int nt, double *uX, double *hj) {
                constexpr int U = decltype(Ut)::value;
                // opaque index: the index arithmetic is not hoisted into the callers' loops
                asm volatile("" : "+v"(ti));
                const double *src = FJ + (long)j * FL.sj;
                const int N = (int)FL.sj, NU3 = ngv * ND * NC;
                for (int i0 = ti; i0 < N; i0 += U * nt) {
                    double v[U];
#pragma unroll
                    for (int u = 0; u < U; u++) v[u] = src[min(i0 + u * nt, N - 1)];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int i = i0 + u * nt;
                        if (i < NU3) uX[i] = v[u];
                        else if (i < N) hj[i - NU3] = v[u];
                    }
                }
            };
            // With the second buffer the last wave stages row block j+1 while the
            // others contract row block j; without it all waves stage between blocks.
            const bool dbuf = LY.dbuf;
            const int ntw = dbuf ? nwave - 1 : nwave;
            // the two staging buffers, formed from the LDS base where they are both in
            // LDS (picked from an array of pointers they lose their address space, and
            // every read in the k-loops becomes a flat load followed by a full wait)
            auto uXb = [&](int b) -> double * {
                return SPILL ? (b ? lds + LY.fuX2 : fuX) : lds + (b ? LY.fuX2 : LY.svt + npv * ngv);
            };
            auto hjb = [&](int b) -> double * { return lds + (b ? LY.fj2 : LY.fj); };
            stage(std::integral_constant<int, 8>(), 0, tid, nth, uXb(0), hjb(0));
            for (int j = 0; j < ncu; j++) {
                __syncthreads();
                const int cb = dbuf ? (j & 1) : 0;
                if (!dbuf && j > 0) {
                    stage(std::integral_constant<int, 8>(), j, tid, nth, uXb(0), hjb(0));
                    __syncthreads();
                }
                if (dbuf && wave == nwave - 1) {
                    if (j + 1 < ncu) stage(std::integral_constant<int, 24>(), j + 1, lane, 64, uXb(cb ^ 1), hjb(cb ^ 1));
                    continue;
                }
                const double *uX = uXb(cb), *hj = hjb(cb), *hhj = hj + np * NC;
                for (int rt = wave; rt < nrt; rt += ntw) {
                    // this lane's A row (kc, i) and its accumulator rows
                    const int ra = rt * 16 + lr;
                    const bool va = ra < ncu * npv;
                    const int kca = va ? ra / npv : 0, ia = va ? ra % npv : 0;
                    const double tdj = (ph.tdep && kca == j) ? ph.fc_u : 0.0;
                    int ccl[TD], lcl[TF], kq[KT];
                    bool kqv[KT];
#pragma unroll
                    for (int t = 0; t < TD; t++) ccl[t] = min(t * 16 + lr, npv - 1);
#pragma unroll
                    for (int t = 0; t < TF; t++) lcl[t] = min(t * 16 + lr, ndf - 1);
#pragma unroll
                    for (int kt = 0; kt < KT; kt++) {
                        kq[kt] = min(kt * 16 + lr, npv - 1);
                        kqv[kt] = kt * 16 + lr < npv;
                    }
                    f64x4 acc[NT], w[ND][KT];
#pragma unroll
                    for (int t = 0; t < NT; t++) acc[t] = (f64x4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int x = 0; x < ND; x++)
#pragma unroll
                        for (int kt = 0; kt < KT; kt++) w[x][kt] = (f64x4){0.0, 0.0, 0.0, 0.0};
                    // volume points.  Every operand load of a k-step is issued before any
                    // of them is used (the sched_barrier): left to itself, the scheduler
                    // interleaves them with the arithmetic, one LDS round trip at a time.
                    for (int g0 = 0; g0 < ngv; g0 += 4) {
                        const int g = min(g0 + lk, ngv - 1);
                        const bool vg = g0 + lk < ngv;
                        double s[ND1], u0[ND], u1[ND][ND], b[TD], q[KT];
#pragma unroll
                        for (int r = 0; r < ND1; r++) s[r] = SVG(ia, g, r);
                        const double jg = jac[g];
#pragma unroll
                        for (int r = 0; r < ND; r++) u0[r] = uX[g + ngv * (r + ND * kca)];
#pragma unroll
                        for (int x = 0; x < ND; x++)
#pragma unroll
                            for (int r = 0; r < ND; r++) u1[x][r] = uX[g + ngv * (r + ND * ((x + 1) * ncu + kca))];
#pragma unroll
                        for (int t = 0; t < TD; t++) b[t] = SVT(g, ccl[t], 0);
#pragma unroll
                        for (int kt = 0; kt < KT; kt++) q[kt] = Qm[kq[kt] + CS * g];
                        __builtin_amdgcn_sched_barrier(0);
                        double a0 = tdj * jg * s[0];
#pragma unroll
                        for (int r = 0; r < ND; r++) a0 -= u0[r] * s[r + 1];
                        a0 = vg ? a0 : 0.0;
                        double ax[ND];
#pragma unroll
                        for (int x = 0; x < ND; x++) {
                            double v = 0.0;
#pragma unroll
                            for (int r = 0; r < ND; r++) v += u1[x][r] * s[r + 1];
                            ax[x] = vg ? -v : 0.0;
                        }
#pragma unroll
                        for (int kt = 0; kt < KT; kt++) q[kt] = kqv[kt] ? q[kt] : 0.0;
#pragma unroll
                        for (int t = 0; t < TD; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b[t], acc[t], 0, 0, 0);
#pragma unroll
                        for (int x = 0; x < ND; x++)
#pragma unroll
                            for (int kt = 0; kt < KT; kt++)
                                w[x][kt] = __builtin_amdgcn_mfma_f64_16x16x4f64(q[kt], ax[x], w[x][kt], 0, 0, 0);
                    }
                    // face points: the face-node positions of this lane's row and D
                    // columns on each face first (no dependent index loads in the loop),
                    // then the operand loads of a k-step issued together, as above
                    int kia[NFE], kcc[TD][NFE];
#pragma unroll
                    for (int a = 0; a < NFE; a++) {
                        kia[a] = INVP(ia, a);
#pragma unroll
                        for (int t = 0; t < TD; t++) kcc[t][a] = INVP(ccl[t], a);
                    }
                    for (int p0 = 0; p0 < np; p0 += 4) {
                        const int pp = min(p0 + lk, np - 1);
                        const int a = pp / ngf, gf = pp % ngf;
                        int ki = kia[0], kc[TD];
#pragma unroll
                        for (int t = 0; t < TD; t++) kc[t] = kcc[t][0];
#pragma unroll
                        for (int a2 = 1; a2 < NFE; a2++)
                            if (a == a2) {
                                ki = kia[a2];
#pragma unroll
                                for (int t = 0; t < TD; t++) kc[t] = kcc[t][a2];
                            }
                        double bu[NT], q[KT], hv[ND + 2];
                        const double sg = SFG(max(ki, 0), gf, 0), jf = jacf[pp];
#pragma unroll
                        for (int t = 0; t < TD; t++) bu[t] = SFT(gf, max(kc[t], 0), 0);
#pragma unroll
                        for (int t = 0; t < TF; t++) bu[TD + t] = SFT(gf, lcl[t] % npf, 0);
#pragma unroll
                        for (int kt = 0; kt < KT; kt++) q[kt] = Qm[kq[kt] + CS * (ngv + pp)];
                        hv[0] = hj[pp + np * kca];
                        hv[1] = hhj[pp + np * kca];
#pragma unroll
                        for (int x = 0; x < ND; x++) hv[2 + x] = hj[pp + np * ((x + 1) * ncu + kca)];
                        __builtin_amdgcn_sched_barrier(0);
                        const double af = (p0 + lk < np && ki >= 0) ? sg * jf : 0.0;
#pragma unroll
                        for (int t = 0; t < TD; t++) bu[t] = kc[t] >= 0 ? bu[t] : 0.0;
#pragma unroll
                        for (int t = 0; t < TF; t++) bu[TD + t] = lcl[t] / npf == a ? bu[TD + t] : 0.0;
#pragma unroll
                        for (int kt = 0; kt < KT; kt++) q[kt] = kqv[kt] ? q[kt] : 0.0;
                        const double ad = af * hv[0];
                        const double ah = af * hv[1];
#pragma unroll
                        for (int t = 0; t < TD; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ad, bu[t], acc[t], 0, 0, 0);
#pragma unroll
                        for (int t = TD; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ah, bu[t], acc[t], 0, 0, 0);
#pragma unroll
                        for (int x = 0; x < ND; x++) {
                            const double aq = af * hv[2 + x];
#pragma unroll
                            for (int kt = 0; kt < KT; kt++)
                                w[x][kt] = __builtin_amdgcn_mfma_f64_16x16x4f64(q[kt], aq, w[x][kt], 0, 0, 0);
                        }
                    }
                    // q columns through the nodes: element v of W_x^T tile kt is the A
                    // operand of the k-step over nodes kt*16 + 4v .. +3
#pragma unroll
                    for (int x = 0; x < ND; x++)
#pragma unroll
                        for (int kt = 0; kt < KT; kt++)
#pragma unroll
                            for (int v = 0; v < 4; v++) {
                                if (kt * 16 + 4 * v >= NPV) continue;
                                const int k = min(kt * 16 + 4 * v + lk, npv - 1);
                                double b[NT];
#pragma unroll
                                for (int t = 0; t < TD; t++) b[t] = C[x * CS * npv + ccl[t] * CS + k];
#pragma unroll
                                for (int t = 0; t < TF; t++) b[TD + t] = -E[x * CS * ndf + lcl[t] * CS + k];
#pragma unroll
                                for (int t = 0; t < NT; t++)
                                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(w[x][kt][v], b[t], acc[t], 0, 0, 0);
                                // keep the scheduler from hoisting every k-step's operand loads
                                // ahead of the products (they would not fit in registers)
                                __builtin_amdgcn_sched_barrier(0);
                            }
#pragma unroll
                    for (int t = 0; t < NT; t++)
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            const int r = rt * 16 + lk + 4 * v;
                            if (r >= ncu * npv) continue;
                            const int kcr = r / npv, ir = r % npv;
                            if (t < TD) {
                                const int cc = t * 16 + lr;
                                if (cc < npv) Dm[e * (long)na * na + (long)na * (kcr * npv + cc) + j * npv + ir] = acc[t][v];
                            } else {
                                const int l = (t - TD) * 16 + lr;
                                if (l < ndf) Fx[e * (long)na * ldx + (long)(j * npv + ir) * ldx + l * nch + kcr] = acc[t][v];
                            }
                        }
                }
            }
        }
        __syncthreads();

        }   // !res (P3-P8)

        // ---------------- P9: Ru, Rh ----------------
        // elementint's residual rows (:418-549) and faceint's (:862-1072, boundary
        // faces with fbou): tables T[g + ngv*(j + ncu*r)] (r = 0 the time term, r > 0
        // f . Xx from the flux kernel) and the face fluxes times jacf (interior fh for Ru; fh or fb for Rh)
        {
            double *T = fj, *Fi = fj + ngv * ncu * nd1, *Fo = Fi + np * nch;
            const int NT0 = ngv * ncu * nd1, NT1 = NT0 + np * nch, NT2 = NT1 + np * nch;
            for (int i = tid; i < NT2; i += nth) {
                if (i < NT0) {
                    const int g = i % ngv, r2 = i / ngv, jj = r2 % ncu, r = r2 / ncu;
                    double s = 0.0;
                    if (r == 0) {
                        if (ph.tdep) {
                            double sv = 0.0;
                            for (int q = 0; q < npv; q++) sv += SVT(g, q, 0) * sh[jj * npv + q];
                            s = (sv - udgg[g + ngv * jj] * ph.fc_u) * jac[g];
                        }
                    } else {
                        s = fX[g + ngv * (jj + ncu * (r - 1))];
                    }
                    T[i] = s;
                } else if (i < NT1) {
                    const int k = i - NT0, pp = k % np, jj = k / np;
                    Fi[k] = fh[pp + np * jj] * jacf[pp];
                } else {
                    const int k = i - NT1, pp = k % np, jj = k / np;
                    const double *fx = bnd[pp / ngf] ? fb : fh;
                    Fo[k] = fx[pp + np * jj] * jacf[pp];
                }
            }
            __syncthreads();
            const long nH = (long)nb * nb;
            double ru2loc = 0.0;
            for (int col = tid; col < na + nb; col += nth) {
                if (col < na) {
                    const int row = col, jj = row / npv, ii = row % npv;
                    double s = 0.0;
#pragma unroll 4
                    for (int g = 0; g < ngv; g++)
#pragma unroll
                        for (int r = 0; r < ND1; r++) s += SVG(ii, g, r) * T[g + ngv * (jj + ncu * r)];
                    for (int a = 0; a < nfe; a++) {
                        const int kk = INVP(ii, a);
                        if (kk < 0) continue;
#pragma unroll 4
                        for (int g = 0; g < ngf; g++) s -= SFG(kk, g, 0) * Fi[g + ngf * a + np * jj];
                    }
                    Ru[e * na + row] = s;
                    Fx[e * (long)na * ldx + (long)row * ldx + nb] = s;
                    ru2loc += s * s;
                } else {
                    const int row = col - na, jj = row % nch, tr = row / nch, a = tr / npf, kk = tr % npf;
                    double s = 0.0;
#pragma unroll 4
                    for (int g = 0; g < ngf; g++) s += SFG(kk, g, 0) * Fo[g + ngf * a + np * jj];
                    Rh[e * nb + row] = -s;
                    Hx[e * (nH + nb) + nH + row] = -s;
                }
            }
            // |Ru_e|^2 for the residual norm (assembly/globalassembly.cpp:501-502)
            for (int o = 32; o > 0; o >>= 1) ru2loc += __shfl_xor(ru2loc, o);
            if (lane == 0) red[wave] = ru2loc;
            __syncthreads();
            if (tid == 0 && ru2) {
                double s = 0.0;
                for (int w = 0; w < nwave; w++) s += red[w];
                ru2[e] = s;
            }
        }

        // ---------------- P9k: trace rows of K and H on the FP64 MFMA ----------------
        // faceint's trace rows (:1254-1318; boundary faces with the far-field flux,
        // :1074-1246): for face a, rows (a, kk, j) and columns (kc, cc) of K, (kc, l) of H,
        //   K^T[cc, (kc,kk,j)] = sum_{x',g} Bk[cc, (x',g)] Pk[(x',g), (kc,kk,j)],
        //   Pk = shapfg(kk,g) jacf(pp) xu(pp, j, x' ncu + kc)  (xuh for the u part of H),
        // with Bk / Bh the face-point operands of the face part of D / F and xu the
        // interior (fhu) or boundary (fbu) flux Jacobian of the face.  Computed
        // transposed so that a tile column runs along contiguous (kk, j) rows of K, H.
        // The face's point values of M^{-1}C_x and M^{-1}E_x (Q C_x, Q E_x on its
        // ngf points) are formed first, on the MFMA, next to the flux Jacobians.
        if (!res) {
            const int ncol = ncu * npf * nch, nct = (ncol + 15) / 16;
            const int nrg = (ngf + 15) / 16;
            const long nH = (long)nb * nb;
            for (int a = 0; a < nfe; a++) {
                const bool bd = bnd[a];
                const double *xu = bd ? fbu : FJ + ohu, *xuh = bd ? fbuh : FJ + ohh;
                const long sju = bd ? (long)np * NC : FL.sj, sjh = bd ? (long)np * NCU : FL.sj;
                __syncthreads();
                // this face's flux Jacobians: xa[g + ngf*(j + nch*c)] (c < nc), then
                // xha[g + ngf*(j + nch*kc)] right after them
                {
                    const int N = ngf * nch * (nc + nch);
                    for (int i0 = tid; i0 < N; i0 += 3 * nth) {
                        double v[3];
#pragma unroll
                        for (int u = 0; u < 3; u++) {
                            const int i = min(i0 + u * nth, N - 1);
                            const int g = i % ngf, r = i / ngf, jj = r % nch, c = r / nch, pp = g + ngf * a;
                            const double *src = c < nc ? xu + pp + np * c + sju * jj : xuh + pp + np * (c - nc) + sjh * jj;
                            v[u] = *src;
                        }
#pragma unroll
                        for (int u = 0; u < 3; u++)
                            if (i0 + u * nth < N) xa[i0 + u * nth] = v[u];
                    }
                }
                // cf[cc + npv*(g + ngf*x)] = sum_k Q(pp,k) C_x(k,cc),
                // ef[l + ndf*(g + ngf*x)] = sum_k Q(pp,k) E_x(k,l),   pp = g + ngf*a
                for (int task = wave; task < nrg * ND * NT; task += nwave) {
                    const int t = task % NT, r2 = task / NT, x = r2 % ND, rg = r2 / ND;
                    const int pp = min(rg * 16 + lr, ngf - 1) + ngf * a;
                    const bool isD = t < TD;
                    const int col = isD ? min(t * 16 + lr, npv - 1) : min((t - TD) * 16 + lr, ndf - 1);
                    const double *bsrc = isD ? C + x * CS * npv + col * CS : E + x * CS * ndf + col * CS;
                    f64x4 acc = (f64x4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int k0 = 0; k0 < NPV; k0 += 4) {
                        const int k = min(k0 + lk, npv - 1);
                        const double qv = Qm[k + CS * (ngv + pp)];
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(k0 + lk < npv ? qv : 0.0, bsrc[k], acc, 0, 0, 0);
                    }
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int gr = rg * 16 + lk + 4 * v;
                        if (gr >= ngf) continue;
                        if (isD) {
                            const int cc = t * 16 + lr;
                            if (cc < npv) cf[cc + npv * (gr + ngf * x)] = acc[v];
                        } else {
                            const int l = (t - TD) * 16 + lr;
                            if (l < ndf) ef[l + ndf * (gr + ngf * x)] = acc[v];
                        }
                    }
                }
                __syncthreads();
                for (int ct = wave; ct < nct; ct += nwave) {
                    const int colb = ct * 16 + lr;              // this lane's B column
                    const bool vb = colb < ncol;
                    const int kcb = vb ? colb / (npf * nch) : 0, rb = vb ? colb % (npf * nch) : 0;
                    const int kkb = rb / nch, jb = rb % nch;
                    // A rows of this lane (clamped; rows past the end are never stored) and
                    // their face-node positions on face a
                    int ccl[TD], kcl[TD], lcl[TF];
#pragma unroll
                    for (int t = 0; t < TD; t++) {
                        ccl[t] = min(t * 16 + lr, npv - 1);
                        kcl[t] = INVP(ccl[t], a);
                    }
#pragma unroll
                    for (int t = 0; t < TF; t++) lcl[t] = min(t * 16 + lr, ndf - 1);
                    f64x4 acc[NT];
#pragma unroll
                    for (int t = 0; t < NT; t++) acc[t] = (f64x4){0.0, 0.0, 0.0, 0.0};
                    for (int g0 = 0; g0 < ngf; g0 += 4) {
                        const int g = min(g0 + lk, ngf - 1), pp = g + ngf * a;
                        double wgt = SFG(kkb, g, 0) * jacf[pp];
                        wgt = g0 + lk < ngf ? wgt : 0.0;
                        double au[NT];
#pragma unroll
                        for (int t = 0; t < TD; t++) {
                            const double v = SFT(g, max(kcl[t], 0), 0);
                            au[t] = kcl[t] >= 0 ? v : 0.0;
                        }
#pragma unroll
                        for (int t = 0; t < TF; t++) {
                            const double v = SFT(g, lcl[t] % npf, 0);
                            au[TD + t] = lcl[t] / npf == a ? v : 0.0;
                        }
                        double aq[ND][NT];
#pragma unroll
                        for (int x = 0; x < ND; x++) {
#pragma unroll
                            for (int t = 0; t < TD; t++) aq[x][t] = cf[ccl[t] + npv * (g + ngf * x)];
#pragma unroll
                            for (int t = 0; t < TF; t++) aq[x][TD + t] = -ef[lcl[t] + ndf * (g + ngf * x)];
                        }
                        // u part: K with xu, H with xuh
                        const double bk = wgt * xa[g + ngf * (jb + nch * kcb)];
                        const double bh = wgt * xa[ngf * nch * nc + g + ngf * (jb + nch * kcb)];
#pragma unroll
                        for (int t = 0; t < TD; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(au[t], bk, acc[t], 0, 0, 0);
#pragma unroll
                        for (int t = TD; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(au[t], bh, acc[t], 0, 0, 0);
#pragma unroll
                        for (int x = 0; x < ND; x++) {
                            const double bq = wgt * xa[g + ngf * (jb + nch * (ncu * (x + 1) + kcb))];
#pragma unroll
                            for (int t = 0; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(aq[x][t], bq, acc[t], 0, 0, 0);
                        }
                    }
                    // result: lane holds C(row = l/16 + 4v, col = l%16) of each tile
                    const int colc = ct * 16 + lr;
                    if (colc < ncol) {
                        const int kcc = colc / (npf * nch), rc = colc % (npf * nch);
                        const long rowK = (long)(a * npf) * nch + rc;     // (a, kk, j) row of K and H
#pragma unroll
                        for (int t = 0; t < NT; t++)
#pragma unroll
                            for (int v = 0; v < 4; v++) {
                                const int rr = t < TD ? t * 16 + lk + 4 * v : (t - TD) * 16 + lk + 4 * v;
                                if (t < TD) {
                                    if (rr < npv) Km[e * (long)nb * na + (long)nb * (kcc * npv + rr) + rowK] = acc[t][v];
                                } else {
                                    if (rr < ndf) Hx[e * (nH + nb) + (long)nb * (rr * nch + kcc) + rowK] = acc[t][v];
                                }
                            }
                    }
                }
            }
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




// Far-field boundary flux at one face point (application/FM/fbou.cpp:126-217, ib 1):
// fb = (A+|A|)(u-uh) - (A-|A|)(ui-uh) = A(uh)(u-ui) + |A|(uh)(u+ui-2uh), with A the
// normal flux Jacobian at uh (getan).  Column k of d fb / d uh goes to fbuh; the
// direction-0 call also writes fb and d fb / d u (zero in the q columns).  Point
// arrays are strided by np: uh[np*c], u[np*c], nl[np*x]; outputs fb[np*j],
// fbu[np*(col + NC*j)], fbuh[np*(col + NCU*j)].
template <int ND>
__device__ __forceinline__ void boundary_point(const Phys &ph, const double *uh, const double *u, const double *nl,
                                            int np, int k, double *fb, double *fbu, double *fbuh)
{
    constexpr int NCU = ND + 2, NC = NCU * (ND + 1);
    dual uhd[NCU], Am[NCU * NCU], acc[NCU];
    double nrm[ND];
    for (int c = 0; c < NCU; c++) uhd[c] = mk(uh[np * c], c == k ? 1.0 : 0.0);
    for (int x = 0; x < ND; x++) nrm[x] = nl[np * x];
    for (int j = 0; j < NCU; j++) acc[j] = mk(0.0);
    for (int absflag = 0; absflag < 2; absflag++) {
        ns_normal_jac<ND>(ph.param[0], ph.param[1], uhd, nrm, absflag, Am);
        for (int q = 0; q < NCU; q++) {
            const dual uq = mk(u[np * q]);
            const dual w = absflag ? uq + mk(ph.ui[q]) - 2.0 * uhd[q] : uq - mk(ph.ui[q]);
            for (int j = 0; j < NCU; j++) acc[j] = acc[j] + Am[q * NCU + j] * w;
        }
        if (k == 0)
            for (int q = 0; q < NCU; q++)
                for (int j = 0; j < NCU; j++) {
                    double *o = &fbu[np * (q + NC * j)];
                    *o = absflag ? *o + Am[q * NCU + j].v : Am[q * NCU + j].v;
                }
    }
    for (int j = 0; j < NCU; j++) fbuh[np * (k + NCU * j)] = acc[j].d;
    if (k == 0)
        for (int j = 0; j < NCU; j++) {
            fb[np * j] = acc[j].v;
            for (int q = NCU; q < NC; q++) fbu[np * (q + NC * j)] = 0.0;
        }
}

// Pointwise fluxes and Jacobian columns (elementflux, :239-305; faceflux/fhat,
// :575-677 and application/fhatDriver.cpp:54-179; far-field fbou,
// application/FM/fbou.cpp:126-217).  Split from the assembly kernel so that the
// dual-number register pressure does not limit the contraction phases.
template <int ND, int NPV, int NPF, class Src>
__device__ __forceinline__ void
flux_body(const Dims &d, const Master &m, const Phys &ph, const double *__restrict__ dgnodes,
          const double *__restrict__ UDG, const double *__restrict__ UH, const int *__restrict__ elcon,
          const int *__restrict__ bf, int e0, int n, double *__restrict__ fluxbuf, long fstride, int res, Src &src)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x, nth = blockDim.x;
    constexpr int NCU = ND + 2, NC = NCU * (ND + 1);
    const int npv = NPV, ngv = d.ngv, npf = NPF, ngf = d.ngf, nfe = d.nfe, ndf = d.ndf, nch = NCU, nc = NC;
    const int np = d.np;
    double *p = lds;
    double *pn = p; p += npv * ND;
    double *udg = p; p += npv * nc;
    double *uh = p; p += ndf * nch;
    double *nl = p; p += np * ND;
    double *udgg = p; p += ngv * nc;
    double *ufg = p; p += np * nc;
    double *uhgl = p; p += np * nch;
    double *Xx = p; p += ngv * ND * ND;    // Xx(g,x,r) at [g + ngv*(x + ND*r)]
    int *bnd = (int *)p;
    const FluxLayout FL = flux_layout(d);
    const double tau = ph.param[5];
#define SVT(g, nn, dd) m.shapvt[(g) + ngv * ((nn) + npv * (dd))]
#define SFT(g, k, dd) m.shapft[(g) + ngf * ((k) + npf * (dd))]
#define PERM(k, a) m.perm[(k) + npf * (a)]
    for (int t = src.first(); t < n; t = src.next(t)) {
        const long e = e0 + t;
        double *F = fluxbuf + (long)t * fstride;
        double *fX = F + FL.fX, *FJ = F + FL.J, *fh = F + FL.fh;
        double *fb = F + FL.fb, *fbu = F + FL.fbu, *fbuh = F + FL.fbuh;
        const long sj = FL.sj, ohu = (long)ngv * ND * NC, ohh = ohu + (long)np * NC;
        for (int i = tid; i < npv * ND; i += nth) pn[i] = dgnodes[e * npv * ND + i];
        for (int i = tid; i < npv * nc; i += nth) udg[i] = UDG[e * npv * nc + i];
        for (int i = tid; i < ndf * nch; i += nth) {
            int j = i / nch, k = i % nch;
            uh[k * ndf + j] = UH[(long)elcon[e * ndf + j] * nch + k];
        }
        for (int a = tid; a < nfe; a += nth) bnd[a] = bf[e * nfe + a] < 0;
        __syncthreads();
        // unit normals at face points, solution at volume and face points
        for (int i = tid; i < np + ngv * nc + np * nc + np * nch + ngv; i += nth) {
            if (i >= np + ngv * nc + np * nc + np * nch) {
                const int g = i - (np + ngv * nc + np * nc + np * nch);
                double adj[ND][ND];
                volume_adjugate<ND>([&](int k, int r) { return SVT(g, k, 1 + r); }, pn, npv, adj);
                for (int x = 0; x < ND; x++)
                    for (int r = 0; r < ND; r++) Xx[g + ngv * (x + ND * r)] = adj[r][x];
            } else if (i < np) {
                const int pp = i, g = pp % ngf, a = pp / ngf;
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
                for (int x = 0; x < ND; x++) nl[pp + np * x] = nv[x] / len;
            } else if (i < np + ngv * nc) {
                int r0 = i - np, g = r0 % ngv, c = r0 / ngv;
                double s = 0;
                for (int q = 0; q < npv; q++) s += SVT(g, q, 0) * udg[c * npv + q];
                udgg[r0] = s;
            } else if (i < np + ngv * nc + np * nc) {
                int r0 = i - np - ngv * nc, pp = r0 % np, c = r0 / np, g = pp % ngf, a = pp / ngf;
                double s = 0;
                for (int q = 0; q < npf; q++) s += SFT(g, q, 0) * udg[c * npv + PERM(q, a)];
                ufg[r0] = s;
            } else {
                int r0 = i - np - ngv * nc - np * nc, pp = r0 % np, c = r0 / np, g = pp % ngf, a = pp / ngf;
                double s = 0;
                for (int q = 0; q < npf; q++) s += SFT(g, q, 0) * uh[c * ndf + a * npf + q];
                uhgl[r0] = s;
            }
        }
        __syncthreads();
        {
            // res: fluxes only (derivative direction 0 carries the values), for the
            // residual-only assembly
            const int nvol = ngv * (res ? 1 : NC), nfac = np * (res ? 1 : NC), nbnd = np * (res ? 1 : NCU);
            for (int i = tid; i < nvol; i += nth) {
                const int g = i % ngv, k = i / ngv;
                dual w[NC], fl[NCU * ND];
                for (int c = 0; c < NC; c++) w[c] = mk(udgg[g + ngv * c], c == k ? 1.0 : 0.0);
                ns_flux<ND>(ph.param, w, fl);
                // folded with Xx: row block j, column k of d f / d udg, and f itself
                for (int j = 0; j < NCU; j++)
                    for (int r = 0; r < ND; r++) {
                        double sd = 0.0, sv = 0.0;
                        for (int x = 0; x < ND; x++) {
                            const double xx = Xx[g + ngv * (x + ND * r)];
                            sd += fl[j + NCU * x].d * xx;
                            sv += fl[j + NCU * x].v * xx;
                        }
                        FJ[j * sj + g + ngv * (r + ND * k)] = sd;
                        if (k == 0) fX[g + ngv * (j + NCU * r)] = sv;
                    }
            }
            for (int r = tid; r < nfac; r += nth) {
                const int pp = r % np, k = r / np;
                dual w[NC], fl[NCU * ND];
                for (int c = 0; c < NC; c++)
                    w[c] = mk(c < NCU ? uhgl[pp + np * c] : ufg[pp + np * c], c == k ? 1.0 : 0.0);
                ns_flux<ND>(ph.param, w, fl);
                for (int j = 0; j < NCU; j++) {
                    dual fn = mk(0.0);
                    for (int x = 0; x < ND; x++) fn = fn + nl[pp + np * x] * fl[j + NCU * x];
                    FJ[j * sj + ohu + pp + np * k] = (k < NCU) ? (k == j ? tau : 0.0) : fn.d;
                    if (k < NCU) FJ[j * sj + ohh + pp + np * k] = fn.d - (k == j ? tau : 0.0);
                    if (k == 0) fh[pp + np * j] = fn.v + tau * (ufg[pp + np * j] - w[j].v);
                }
            }
            for (int r = tid; r < nbnd; r += nth) {
                const int pp = r % np, k = r / np, a = pp / ngf;
                if (bnd[a]) boundary_point<ND>(ph, uhgl + pp, ufg + pp, nl + pp, np, k, fb + pp, fbu + pp, fbuh + pp);
            }
        }

        __syncthreads();
    }
#undef SVT
#undef SFT
#undef PERM
}

}  // namespace hfem
