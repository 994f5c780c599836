// Element assembly for the compressible-NS HDG element path (gfx950).
//
// One workgroup assembles one element at a time (grid-stride over elements) and
// writes that element's D, F, K, H, Ru, Rh blocks (assembly/elementassembly.cpp:1481-1499
// and the faceint/elementint outputs).  D and K use the reference's column-major
// layouts; F is written row-major into Fx (na x ldx) with Ru as column nb, and H
// column-major into Hx (nb x (nb+1)) with Rh as column nb -- the layouts the
// condensation kernel streams through its MFMA operands.
//
// Phases (barriers between them; inside a phase threads own disjoint outputs):
//   P0 element data and the master tables in LDS  P5 traces of M^{-1}C, M^{-1}E;
//   P1 volume/face geometry                          solution at volume points
//   P2 q operators M, C, E (qint)                  P7 wrl = -[s_u jac ; f_u . Xx]
//   P3 M^{-1} (Gauss-Jordan)
//   P4 Q = [shapvt; face traces] M^{-1}, so that P5 forms the traces of M^{-1}C
//      and M^{-1}E as Q C and Q E without solving per column
//   (the pointwise fluxes and their Jacobians come from flux_kernel, launched
//   before this kernel on the same element chunk)
//   P8 element rows: one thread per column of D or F, all NPV rows in registers;
//      the q-folded integrand (elementint :510-549, faceint :962-1072) is formed
//      on the fly and contracted with the test functions, volume then faces
//      (faces containing the row node: the gather form of faceint's scatter)
//   P9 trace rows: one thread per (face, column) of K or H, npf rows in registers;
//      Ru and Rh entries.
#include "hfem_device.h"

#include <algorithm>

namespace hfem {

constexpr int ASM_THREADS = 512;

typedef double f64x4 __attribute__((ext_vector_type(4)));

// LDS copies of the master tables, transposed so that the thread-varying index is
// fastest: svt[n + npv*(g + ngv*dd)], svg[i + npv*(g + ngv*dd)] (as shapvg),
// sft[k + npf*(g + ngf*dd)], sfg[k + npf*(g + ngf*dd)] (as shapfg).
struct LdsTables {
    double *svt, *svg, *sft, *sfg;
    int *perm, *invperm;
};

template <int ND, int NPV, int NPF>
__global__ void __launch_bounds__(ASM_THREADS)
assemble_kernel(Dims d, Master m, Phys ph, const double *__restrict__ dgnodes, const double *__restrict__ UDG,
                const double *__restrict__ UH, const double *__restrict__ SH, const int *__restrict__ elcon,
                const int *__restrict__ bf, int e0, int n, double *__restrict__ scratch, long sstride, int ldx,
                const double *__restrict__ fluxbuf, long fstride, double *__restrict__ Ru,
                double *__restrict__ Dm, double *__restrict__ Fx, double *__restrict__ Km, double *__restrict__ Hx,
                double *__restrict__ Rh, double *__restrict__ ru2)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x, nth = blockDim.x;
    const FluxLayout FL = flux_layout(d);
    constexpr int NCU = ND + 2, NC = NCU * (ND + 1), ND1 = ND + 1;
    const int npv = NPV, ngv = d.ngv, npf = NPF, ngf = d.ngf, nfe = d.nfe, ndf = d.ndf;
    const int ncu = NCU, nch = NCU, nc = NC, na = d.na, nb = d.nb, nd1 = ND1, np = d.np;

    // ---- LDS layout ----
    double *p = lds;
    // shapvt; its derivative part (dead after P1) later holds fu of one row block
    double *svt = p; p += npv * ngv + (npv * ND > ND * NC ? npv : NC) * ngv * ND;
    double *svg = p; p += npv * ngv * nd1;
    double *sft = p; p += npf * ngf * ND;
    double *sfg = p; p += npf * ngf * ND;
    double *pn = p; p += npv * ND;
    double *udg = p; p += npv * nc;
    double *uh = p; p += ndf * nch;            // uh[k*ndf + j]
    double *sh = p; p += npv * ncu;
    double *jac = p; p += ngv;
    double *Xx = p; p += ngv * ND * ND;        // Xx(g,x,r) = jac d xi_r / d x_x
    double *jacf = p; p += np;
    double *nl = p; p += np * ND;
    // mass matrix / its inverse, Q and the Gauss-Jordan rows; later G (P8), then the
    // flux Jacobians of one face (P9k)
    double *M = p;
    double *Qm = M + npv * npv;                // [k + npv*r]: rows of shapvt / face traces times M^{-1}
    double *rowk = Qm + (ngv + np) * npv;
    double *colk = rowk + npv;
    double *xa = M, *xha = M + ngf * nch * nc;
    {
        int mr = (npv + ngv + np + 2) * npv;
        if (npv * ngv * ND > mr) mr = npv * ngv * ND;
        if (ngf * nch * (nc + nch) > mr) mr = ngf * nch * (nc + nch);
        p += mr;
    }
    double *shapMiC = p; p += npv * ngv * ND;  // [n + npv*(g + ngv*x)]
    double *shapMiE = p; p += ndf * ngv * ND;  // [l + ndf*(g + ngv*x)], includes the minus sign
    double *udgg = p; p += ngv * nc;           // [g + ngv*c]
    double *fhuj = p; p += np * nc;            // fhu of one row block [pp + np*c]
    double *fhuhj = p; p += np * nch;          // fhuh of one row block [pp + np*kc]
    double *red = p; p += ASM_THREADS;
    int *perm = (int *)p;
    int *invperm = perm + npf * nfe;
    int *bnd = invperm + npv * nfe;

    // ---- per-workgroup scratch (global) ----
    double *S = scratch + (long)blockIdx.x * sstride;
    const Scratch L = scratch_layout(d);
    double *C = S + L.C, *E = S + L.E, *shapMiCf = S + L.shapMiCf, *shapMiEf = S + L.shapMiEf;
    const double tau = ph.param[5];

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

    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const long e = e0 + t;
        const double *FB = fluxbuf + (long)t * fstride;
        const double *f = FB + FL.f, *fu = FB + FL.fu, *fh = FB + FL.fh, *fhu = FB + FL.fhu, *fhuh = FB + FL.fhuh;
        const double *fb = FB + FL.fb, *fbu = FB + FL.fbu, *fbuh = FB + FL.fbuh;
        // ---------------- P0: gather ----------------
        // the derivative part of shapvt held the previous element's fu in P8
        if (t != (int)blockIdx.x)
            for (int i = npv * ngv + tid; i < npv * ngv * nd1; i += nth) {
                const int nn = i % npv, r = i / npv, g = r % ngv, dd = r / ngv;
                svt[i] = m.shapvt[g + ngv * (nn + npv * dd)];
            }
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
                double J[ND][ND];
                for (int x = 0; x < ND; x++)
                    for (int r = 0; r < ND; r++) {
                        double s = 0;
                        for (int k = 0; k < npv; k++) s += SVT(g, k, 1 + r) * pn[x * npv + k];
                        J[x][r] = s;
                    }
                double adj[ND][ND], det;
                if (ND == 2) {
                    det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
                    adj[0][0] = J[1][1]; adj[0][1] = -J[0][1]; adj[1][0] = -J[1][0]; adj[1][1] = J[0][0];
                } else {
                    for (int r = 0; r < ND; r++)
                        for (int x = 0; x < ND; x++) {
                            int x1 = (x + 1) % 3, x2 = (x + 2) % 3, r1 = (r + 1) % 3, r2 = (r + 2) % 3;
                            adj[r][x] = J[x1][r1] * J[x2][r2] - J[x1][r2] * J[x2][r1];
                        }
                    det = J[0][0] * adj[0][0] + J[1][0] * adj[0][1] + J[2][0] * adj[0][2];
                }
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

        // ---------------- P2: q operators (qint) ----------------
        {
            const int nM = npv * npv, nC = ND * npv * npv, nE = ND * npv * ndf;
            for (int i = tid; i < nM + nC + nE; i += nth) {
                if (i < nM) {
                    int jj = i % npv, ii = i / npv;
                    double s = 0;
                    for (int g = 0; g < ngv; g++) s += (SVG(jj, g, 0) * SVT(g, ii, 0)) * jac[g];
                    M[i] = ph.fc_q * s;
                } else if (i < nM + nC) {
                    int r = i - nM, x = r / (npv * npv), rr = r % (npv * npv), jj = rr % npv, ii = rr / npv;
                    double s = 0;
                    for (int q = 0; q < ND; q++)
                        for (int g = 0; g < ngv; g++) s += (SVG(jj, g, q + 1) * SVT(g, ii, 0)) * Xx[g + ngv * (x + ND * q)];
                    C[r] = s;
                } else {
                    int r = i - nM - nC, x = r / (npv * ndf), rr = r % (npv * ndf), nn = rr % npv, l = rr / npv;
                    int a = l / npf, j = l % npf, k = INVP(nn, a);
                    double s = 0;
                    if (k >= 0)
                        for (int g = 0; g < ngf; g++) {
                            int pp = g + ngf * a;
                            s += (SFG(k, g, 0) * SFT(g, j, 0)) * nl[pp + np * x] * jacf[pp];
                        }
                    E[r] = s;
                }
            }
        }
        __syncthreads();

        // ---------------- P3: M^{-1} by Gauss-Jordan in LDS (M is SPD: no pivoting) ----------------
        for (int k = 0; k < npv; k++) {
            for (int i = tid; i < 2 * npv; i += nth) {
                if (i < npv) rowk[i] = M[k + npv * i];      // M(k, i)
                else colk[i - npv] = M[(i - npv) + npv * k];  // M(i, k)
            }
            __syncthreads();
            const double pinv = 1.0 / rowk[k];
            for (int idx = tid; idx < npv * npv; idx += nth) {
                const int i = idx % npv, j = idx / npv;       // M(i, j) at i + npv*j
                double v;
                if (i == k) v = (j == k) ? pinv : rowk[j] * pinv;
                else if (j == k) v = -colk[i] * pinv;
                else v = M[idx] - colk[i] * rowk[j] * pinv;
                M[idx] = v;
            }
            __syncthreads();
        }

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
            Qm[k + npv * r] = s;
        }
        __syncthreads();

        // ---------------- P5: traces of M^{-1}C, M^{-1}E; solution at points ----------------
        {
            const int n1 = ngv * npv * ND, n2 = ngv * ndf * ND, n3 = np * npv * ND, n4 = np * ndf * ND;
            const int n5 = ngv * nc;
            const int tot = n1 + n2 + n3 + n4 + n5;
            for (int i = tid; i < tot; i += nth) {
                if (i < n1) {
                    int nn = i % npv, r = i / npv, g = r % ngv, x = r / ngv;
                    const double *q = Qm + npv * g, *c = C + x * npv * npv + nn * npv;
                    double s = 0;
                    for (int l = 0; l < npv; l++) s += q[l] * c[l];
                    shapMiC[i] = s;
                } else if (i < n1 + n2) {
                    int r0 = i - n1, l = r0 % ndf, r = r0 / ndf, g = r % ngv, x = r / ngv;
                    const double *q = Qm + npv * g, *c = E + (long)x * npv * ndf + l * npv;
                    double s = 0;
                    for (int k = 0; k < npv; k++) s += q[k] * c[k];
                    shapMiE[r0] = -s;
                } else if (i < n1 + n2 + n3) {
                    int r0 = i - n1 - n2, nn = r0 % npv, r = r0 / npv, pp = r % np, x = r / np;
                    const double *q = Qm + npv * (ngv + pp), *c = C + x * npv * npv + nn * npv;
                    double s = 0;
                    for (int k = 0; k < npv; k++) s += q[k] * c[k];
                    shapMiCf[r0] = s;
                } else if (i < n1 + n2 + n3 + n4) {
                    int r0 = i - n1 - n2 - n3, l = r0 % ndf, r = r0 / ndf, pp = r % np, x = r / np;
                    const double *q = Qm + npv * (ngv + pp), *c = E + (long)x * npv * ndf + l * npv;
                    double s = 0;
                    for (int k = 0; k < npv; k++) s += q[k] * c[k];
                    shapMiEf[r0] = s;
                } else {
                    int r0 = i - n1 - n2 - n3 - n4, g = r0 % ngv, c = r0 / ngv;
                    double s = 0;
                    for (int q = 0; q < npv; q++) s += SVT(g, q, 0) * udg[c * npv + q];
                    udgg[r0] = s;
                }
            }
        }
        __syncthreads();

        // ---------------- P8: element rows of D and F on the FP64 MFMA ----------------
        // Row block j of D and F as one product over volume and face points.  With
        // G(i,g,x) = sum_r shapvg(i,g,r+1) Xx(g,x,r) (jac-weighted test-function
        // gradients), the q-folded volume integrand of elementint (:418-549) is
        //   D_j[(kc,i), cc] += sum_{x',g} P[(kc,i),(x',g)] Bd[(x',g), cc]
        //   F_j[(kc,i), l]  += sum_{x>=1,g} P[(kc,i),(x,g)] Be[(x,g), l]
        //   P = -sum_x fu(g,j,x, c) G(i,g,x) (+ fc_u jac shapvg(i,g,0) for c = j, tdep),
        //   c = x' ncu + kc;  Bd = [shapvt(g,cc); M^{-1}C traces], Be = M^{-1}E traces;
        // the face integrand of faceint (:862-1072), the gather over the faces that
        // contain node i, is the same product over face points pp = (a, g):
        //   A = [i on a] shapfg(k(i,a), g) jacf(pp) fhu(pp, j, c),
        //   D: B = [cc on a] shapft(g, k(cc,a)) (c = kc), M^{-1}C face traces (c = q);
        //   F: B = [l on a] shapft(g, l) with fhuh in A (u part), -M^{-1}E traces (q).
        // One wave per (j, 16-row tile of (kc,i)), all 16-column tiles of D (cc) and
        // F (l) accumulated together; A formed on the fly from LDS operands.
        {
            double *G = M;                                   // M, Qm dead after P5
            double *fuj = svt + npv * ngv;                   // shapvt derivatives dead after P1
            for (int i = tid; i < npv * ngv * ND; i += nth) {
                const int nn = i % npv, r = i / npv, g = r % ngv, x = r / ngv;
                double sg = 0;
#pragma unroll
                for (int q = 0; q < ND; q++) sg += SVG(nn, g, q + 1) * Xx[g + ngv * (x + ND * q)];
                G[i] = sg;
            }
            const int lane = tid & 63, wave = tid >> 6, nwave = nth >> 6, lr = lane & 15, lk = lane >> 4;
            const int nrt = (ncu * npv + 15) / 16;
            // column tiles: TD of D (cc < npv), TF of F (l < ndf); faces per element from
            // the face size (tets: npf triangular, hexes: npf square)
            constexpr int NFE = ND == 2 ? 3 : ((NPF == 4 || NPF == 9 || NPF == 16) ? 6 : 4);
            constexpr int TD = (NPV + 15) / 16, TF = (NPF * NFE + 15) / 16, NT = TD + TF;
            for (int j = 0; j < ncu; j++) {
                __syncthreads();
                for (int i = tid; i < ngv * ND * NC; i += nth) {
                    const int g = i % ngv, r = i / ngv, x = r % ND, c = r / ND;
                    fuj[i] = fu[g + ngv * (j + ncu * (x + ND * c))];
                }
                for (int i = tid; i < np * NC; i += nth) {
                    const int pp = i % np, c = i / np;
                    fhuj[i] = fhu[pp + np * (j + nch * c)];
                    if (c < nch) fhuhj[i] = fhuh[pp + np * (j + nch * c)];
                }
                __syncthreads();
                for (int rt = wave; rt < nrt; rt += nwave) {
                    // this lane's A row and its accumulator rows.  Operands are read at
                    // clamped indices and zeroed by selects, never under a branch, so that
                    // all loads of a k-step issue together; only zeros along the k index
                    // matter (rows and columns past the end are never stored).
                    const int ra = rt * 16 + lr;
                    const bool va = ra < ncu * npv;
                    const int kca = va ? ra / npv : 0, ia = va ? ra % npv : 0;
                    const double tdj = (ph.tdep && kca == j) ? ph.fc_u : 0.0;
                    int ccl[TD], lcl[TF];
#pragma unroll
                    for (int t = 0; t < TD; t++) ccl[t] = min(t * 16 + lr, npv - 1);
#pragma unroll
                    for (int t = 0; t < TF; t++) lcl[t] = min(t * 16 + lr, ndf - 1);
                    f64x4 acc[NT];
#pragma unroll
                    for (int t = 0; t < NT; t++) acc[t] = (f64x4){0.0, 0.0, 0.0, 0.0};
                    // volume points, u columns: B = shapvt
                    for (int g0 = 0; g0 < ngv; g0 += 4) {
                        const int g = min(g0 + lk, ngv - 1);
                        double sa = 0.0;
#pragma unroll
                        for (int x = 0; x < ND; x++) sa += fuj[g + ngv * (x + ND * kca)] * G[ia + npv * (g + ngv * x)];
                        double a = SVG(ia, g, 0) * tdj * jac[g] - sa;
                        a = g0 + lk < ngv ? a : 0.0;
                        double b[TD];
#pragma unroll
                        for (int t = 0; t < TD; t++) b[t] = SVT(g, ccl[t], 0);
#pragma unroll
                        for (int t = 0; t < TD; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[t], acc[t], 0, 0, 0);
                    }
                    // volume points, q columns: B = M^{-1}C and M^{-1}E traces
                    for (int xq = 1; xq < nd1; xq++) {
                        const int c = xq * ncu + kca;
                        for (int g0 = 0; g0 < ngv; g0 += 4) {
                            const int g = min(g0 + lk, ngv - 1);
                            double sa = 0.0;
#pragma unroll
                            for (int x = 0; x < ND; x++) sa += fuj[g + ngv * (x + ND * c)] * G[ia + npv * (g + ngv * x)];
                            const double a = g0 + lk < ngv ? -sa : 0.0;
                            double b[NT];
#pragma unroll
                            for (int t = 0; t < TD; t++) b[t] = shapMiC[ccl[t] + npv * (g + ngv * (xq - 1))];
#pragma unroll
                            for (int t = 0; t < TF; t++) b[TD + t] = shapMiE[lcl[t] + ndf * (g + ngv * (xq - 1))];
#pragma unroll
                            for (int t = 0; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[t], acc[t], 0, 0, 0);
                        }
                    }
                    // face points
                    for (int p0 = 0; p0 < np; p0 += 4) {
                        const int pp = min(p0 + lk, np - 1);
                        const int a = pp / ngf, gf = pp % ngf;
                        const int ki = INVP(ia, a);
                        double af = SFG(max(ki, 0), gf, 0) * jacf[pp];
                        af = (p0 + lk < np && ki >= 0) ? af : 0.0;
                        double bu[NT];
#pragma unroll
                        for (int t = 0; t < TD; t++) {
                            const int k = INVP(ccl[t], a);
                            const double v = SFT(gf, max(k, 0), 0);
                            bu[t] = k >= 0 ? v : 0.0;
                        }
#pragma unroll
                        for (int t = 0; t < TF; t++) {
                            const double v = SFT(gf, lcl[t] % npf, 0);
                            bu[TD + t] = lcl[t] / npf == a ? v : 0.0;
                        }
                        double bq[ND][NT];
#pragma unroll
                        for (int x = 0; x < ND; x++) {
#pragma unroll
                            for (int t = 0; t < TD; t++) bq[x][t] = shapMiCf[ccl[t] + npv * (pp + np * x)];
#pragma unroll
                            for (int t = 0; t < TF; t++) bq[x][TD + t] = -shapMiEf[lcl[t] + ndf * (pp + np * x)];
                        }
                        // u part: D with fhu, F with fhuh
                        const double ad = af * fhuj[pp + np * kca];
                        const double ah = af * fhuhj[pp + np * kca];
#pragma unroll
                        for (int t = 0; t < TD; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ad, bu[t], acc[t], 0, 0, 0);
#pragma unroll
                        for (int t = TD; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ah, bu[t], acc[t], 0, 0, 0);
                        // q parts
#pragma unroll
                        for (int x = 0; x < ND; x++) {
                            const double aq = af * fhuj[pp + np * (ncu * (x + 1) + kca)];
#pragma unroll
                            for (int t = 0; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(aq, bq[x][t], acc[t], 0, 0, 0);
                        }
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

        // ---------------- P9k: trace rows of K and H on the FP64 MFMA ----------------
        // faceint's trace rows (:1254-1318; boundary faces with the far-field flux,
        // :1074-1246): for face a, rows (a, kk, j) and columns (kc, cc) of K, (kc, l) of H,
        //   K^T[cc, (kc,kk,j)] = sum_{x',g} Bk[cc, (x',g)] Pk[(x',g), (kc,kk,j)],
        //   Pk = shapfg(kk,g) jacf(pp) xu(pp, j, x' ncu + kc)  (xuh for the u part of H),
        // with Bk / Bh the same face-point operands as the face part of D / F and xu the
        // interior (fhu) or boundary (fbu) flux Jacobian of the face.  Computed
        // transposed so that a tile column runs along contiguous (kk, j) rows of K, H.
        {
            const int lane = tid & 63, wave = tid >> 6, nwave = nth >> 6, lr = lane & 15, lk = lane >> 4;
            constexpr int NFE = ND == 2 ? 3 : ((NPF == 4 || NPF == 9 || NPF == 16) ? 6 : 4);
            constexpr int TD = (NPV + 15) / 16, TF = (NPF * NFE + 15) / 16, NT = TD + TF;
            const int ncol = ncu * npf * nch, nct = (ncol + 15) / 16;
            const long nH = (long)nb * nb;
            for (int a = 0; a < nfe; a++) {
                const bool bd = bnd[a];
                const double *xu = bd ? fbu : fhu, *xuh = bd ? fbuh : fhuh;
                __syncthreads();
                // this face's flux Jacobians: xa[g + ngf*(j + nch*c)], xha[g + ngf*(j + nch*kc)]
                for (int i = tid; i < ngf * nch * (nc + nch); i += nth) {
                    const int g = i % ngf, r = i / ngf, jj = r % nch, c = r / nch, pp = g + ngf * a;
                    if (c < nc) xa[i] = xu[pp + np * (jj + nch * c)];
                    else xha[g + ngf * (jj + nch * (c - nc))] = xuh[pp + np * (jj + nch * (c - nc))];
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
                            for (int t = 0; t < TD; t++) aq[x][t] = shapMiCf[ccl[t] + npv * (pp + np * x)];
#pragma unroll
                            for (int t = 0; t < TF; t++) aq[x][TD + t] = -shapMiEf[lcl[t] + ndf * (pp + np * x)];
                        }
                        // u part: K with xu, H with xuh
                        const double bk = wgt * xa[g + ngf * (jb + nch * kcb)];
                        const double bh = wgt * xha[g + ngf * (jb + nch * kcb)];
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

        // ---------------- P9: Ru, Rh ----------------
        {
            const long nH = (long)nb * nb;
            double ru2loc = 0.0;
            for (int col = tid; col < na + nb; col += nth) {
                if (col < na) {
                    int row = col, j = row / npv, ii = row % npv;
                    double s = 0;
                    for (int g = 0; g < ngv; g++) {
                        if (ph.tdep) {
                            double sv = 0;
                            for (int q = 0; q < npv; q++) sv += SVT(g, q, 0) * sh[j * npv + q];
                            s += SVG(ii, g, 0) * (sv - udgg[g + ngv * j] * ph.fc_u) * jac[g];
                        }
                        for (int mq = 0; mq < ND; mq++) {
                            double tx = 0;
                            for (int x = 0; x < ND; x++) tx += f[g + ngv * (j + ncu * x)] * Xx[g + ngv * (x + ND * mq)];
                            s += SVG(ii, g, mq + 1) * tx;
                        }
                    }
                    for (int a = 0; a < nfe; a++) {
                        int kk = INVP(ii, a);
                        if (kk < 0) continue;
                        for (int g = 0; g < ngf; g++) {
                            int pp = g + ngf * a;
                            s -= SFG(kk, g, 0) * fh[pp + np * j] * jacf[pp];
                        }
                    }
                    Ru[e * na + row] = s;
                    Fx[e * (long)na * ldx + (long)row * ldx + nb] = s;
                    ru2loc += s * s;
                } else {
                    int row = col - na, j = row % nch, tr = row / nch, a = tr / npf, kk = tr % npf;
                    const double *fx = bnd[a] ? fb : fh;
                    double s = 0;
                    for (int g = 0; g < ngf; g++) {
                        int pp = g + ngf * a;
                        s += SFG(kk, g, 0) * fx[pp + np * j] * jacf[pp];
                    }
                    Rh[e * nb + row] = -s;
                    Hx[e * (nH + nb) + nH + row] = -s;
                }
            }
            // |Ru_e|^2 for the residual norm (assembly/globalassembly.cpp:501-502)
            red[tid] = ru2loc;
            __syncthreads();
            for (int s = nth / 2; s > 0; s >>= 1) {
                if (tid < s) red[tid] += red[tid + s];
                __syncthreads();
            }
            if (tid == 0 && ru2) ru2[e] = red[0];
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
// arrays are strided by np: uh[np*c], u[np*c], nl[np*x], outputs [np*(j + NCU*col)].
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
                    double *o = &fbu[np * (j + NCU * q)];
                    *o = absflag ? *o + Am[q * NCU + j].v : Am[q * NCU + j].v;
                }
    }
    for (int j = 0; j < NCU; j++) fbuh[np * (j + NCU * k)] = acc[j].d;
    if (k == 0)
        for (int j = 0; j < NCU; j++) {
            fb[np * j] = acc[j].v;
            for (int q = NCU; q < NC; q++) fbu[np * (j + NCU * q)] = 0.0;
        }
}

// Pointwise fluxes and Jacobian columns (elementflux, :239-305; faceflux/fhat,
// :575-677 and application/fhatDriver.cpp:54-179; far-field fbou,
// application/FM/fbou.cpp:126-217).  Split from the assembly kernel so that the
// dual-number register pressure does not limit the contraction phases.
template <int ND, int NPV, int NPF>
__global__ void __launch_bounds__(256)
flux_kernel(Dims d, Master m, Phys ph, const double *__restrict__ dgnodes, const double *__restrict__ UDG,
            const double *__restrict__ UH, const int *__restrict__ elcon, const int *__restrict__ bf, int e0, int n,
            double *__restrict__ fluxbuf, long fstride)
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
    int *bnd = (int *)p;
    const FluxLayout FL = flux_layout(d);
    const double tau = ph.param[5];
#define SVT(g, nn, dd) m.shapvt[(g) + ngv * ((nn) + npv * (dd))]
#define SFT(g, k, dd) m.shapft[(g) + ngf * ((k) + npf * (dd))]
#define PERM(k, a) m.perm[(k) + npf * (a)]
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const long e = e0 + t;
        double *F = fluxbuf + (long)t * fstride;
        double *f = F + FL.f, *fu = F + FL.fu, *fh = F + FL.fh, *fhu = F + FL.fhu, *fhuh = F + FL.fhuh;
        double *fb = F + FL.fb, *fbu = F + FL.fbu, *fbuh = F + FL.fbuh;
        for (int i = tid; i < npv * ND; i += nth) pn[i] = dgnodes[e * npv * ND + i];
        for (int i = tid; i < npv * nc; i += nth) udg[i] = UDG[e * npv * nc + i];
        for (int i = tid; i < ndf * nch; i += nth) {
            int j = i / nch, k = i % nch;
            uh[k * ndf + j] = UH[(long)elcon[e * ndf + j] * nch + k];
        }
        for (int a = tid; a < nfe; a += nth) bnd[a] = bf[e * nfe + a] < 0;
        __syncthreads();
        // unit normals at face points, solution at volume and face points
        for (int i = tid; i < np + ngv * nc + np * nc + np * nch; i += nth) {
            if (i < np) {
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
            const int nvol = ngv * NC, nfac = np * NC, nbnd = np * NCU;
            for (int i = tid; i < nvol; i += nth) {
                const int g = i % ngv, k = i / ngv;
                dual w[NC], fl[NCU * ND];
                for (int c = 0; c < NC; c++) w[c] = mk(udgg[g + ngv * c], c == k ? 1.0 : 0.0);
                ns_flux<ND>(ph.param, w, fl);
                for (int jx = 0; jx < NCU * ND; jx++) {
                    fu[g + ngv * (jx + NCU * ND * k)] = fl[jx].d;
                    if (k == 0) f[g + ngv * jx] = fl[jx].v;
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
                    fhu[pp + np * (j + NCU * k)] = (k < NCU) ? (k == j ? tau : 0.0) : fn.d;
                    if (k < NCU) fhuh[pp + np * (j + NCU * k)] = fn.d - (k == j ? tau : 0.0);
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
                double J[ND][ND];
                for (int x = 0; x < ND; x++)
                    for (int r = 0; r < ND; r++) {
                        double s = 0;
                        for (int k = 0; k < npv; k++) s += SVT(g, k, 1 + r) * pn[x * npv + k];
                        J[x][r] = s;
                    }
                double adj[ND][ND], det;
                if (ND == 2) {
                    det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
                    adj[0][0] = J[1][1]; adj[0][1] = -J[0][1]; adj[1][0] = -J[1][0]; adj[1][1] = J[0][0];
                } else {
                    for (int r = 0; r < ND; r++)
                        for (int x = 0; x < ND; x++) {
                            int x1 = (x + 1) % 3, x2 = (x + 2) % 3, r1 = (r + 1) % 3, r2 = (r + 2) % 3;
                            adj[r][x] = J[x1][r1] * J[x2][r2] - J[x1][r2] * J[x2][r1];
                        }
                    det = J[0][0] * adj[0][0] + J[1][0] * adj[0][1] + J[2][0] * adj[0][2];
                }
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
               (size_t)d.ngv * d.nc + (size_t)d.np * d.nc + (size_t)d.np * d.nch;
    return n * sizeof(double) + 16 * sizeof(int);
}

size_t assemble_lds_bytes(const Dims &d)
{
    const size_t nc = d.nc, npv = d.npv, ngv = d.ngv, np = d.np, nd = d.nd;
    const size_t svt = npv * ngv + (npv > nc ? npv : nc) * ngv * nd;
    size_t mreg = (npv + ngv + np + 2) * npv > npv * ngv * nd ? (npv + ngv + np + 2) * npv : npv * ngv * nd;
    if ((size_t)d.ngf * d.nch * (nc + d.nch) > mreg) mreg = (size_t)d.ngf * d.nch * (nc + d.nch);
    size_t n = svt + npv * ngv * d.nd1 + 2 * (size_t)d.npf * d.ngf * nd + npv * nd + npv * nc + (size_t)d.ndf * d.nch +
               npv * d.ncu + ngv + ngv * nd * nd + np + np * nd + mreg + npv * ngv * nd + (size_t)d.ndf * ngv * nd +
               ngv * nc + np * nc + np * d.nch + ASM_THREADS;
    size_t ints = (size_t)d.npf * d.nfe + npv * d.nfe + d.nfe + 8;
    return n * sizeof(double) + ints * sizeof(int);
}

// element families with compiled kernels: (nd, npv, npf)
#define HFEM_FAMILIES(X)                                                            \
    X(2, 3, 2) X(2, 6, 3) X(2, 10, 4) X(2, 15, 5) X(2, 21, 6)                       \
    X(3, 4, 3) X(3, 10, 6) X(3, 20, 10) X(3, 35, 15) X(3, 8, 4) X(3, 27, 9)

bool assemble_supported(int nd, int npv, int npf)
{
#define CHECK(a, b, c) if (nd == a && npv == b && npf == c) return true;
    HFEM_FAMILIES(CHECK)
#undef CHECK
    return false;
}

hipError_t launch_assemble(const Dims &d, const Master &m, const Phys &ph, const double *dgnodes, const double *UDG,
                           const double *UH, const double *SH, const int *elcon, const int *bf, int e0, int n,
                           double *scratch, long sstride, int grid, int ldx, double *fluxbuf, long fstride, int chunk,
                           double *Ru, double *D, double *Fx, double *K, double *Hx, double *Rh, double *ru2,
                           hipStream_t stream)
{
    const size_t lds = assemble_lds_bytes(d), flds = flux_lds_bytes(d);
    for (int c0 = e0; c0 < e0 + n; c0 += chunk) {
        const int cn = std::min(chunk, e0 + n - c0);
        const int fgrid = std::min(cn, 2048), agrid = std::min(cn, grid);
        bool done = false;
#define LAUNCH(a, b, c)                                                                                             \
        if (!done && d.nd == a && d.npv == b && d.npf == c) {                                                       \
            hipFuncSetAttribute((const void *)flux_kernel<a, b, c>, hipFuncAttributeMaxDynamicSharedMemorySize,    \
                                (int)flds);                                                                         \
            hipLaunchKernelGGL((flux_kernel<a, b, c>), dim3(fgrid), dim3(256), flds, stream, d, m, ph, dgnodes,     \
                               UDG, UH, elcon, bf, c0, cn, fluxbuf, fstride);                                       \
            hipFuncSetAttribute((const void *)assemble_kernel<a, b, c>, hipFuncAttributeMaxDynamicSharedMemorySize,\
                                (int)lds);                                                                          \
            hipLaunchKernelGGL((assemble_kernel<a, b, c>), dim3(agrid), dim3(ASM_THREADS), lds, stream, d, m, ph,   \
                               dgnodes, UDG, UH, SH, elcon, bf, c0, cn, scratch, sstride, ldx, fluxbuf, fstride, Ru, \
                               D, Fx, K, Hx, Rh, ru2);                                                              \
            done = true;                                                                                            \
        }
        HFEM_FAMILIES(LAUNCH)
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
