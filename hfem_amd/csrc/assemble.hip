// Element assembly for the compressible-NS HDG element path (gfx950).
//
// One workgroup assembles one element at a time (grid-stride over elements) and
// writes that element's D, F, K, H, Ru, Rh blocks (assembly/elementassembly.cpp:1481-1499
// and the faceint/elementint outputs).  D and K use the reference's column-major
// layouts; F is written row-major into Fx (na x ldx) with Ru as column nb, and H
// column-major into Hx (nb x (nb+1)) with Rh as column nb -- the layouts the
// condensation kernel streams through its MFMA operands.
// The work is organised as a sequence of phases; inside a phase every thread
// computes independent entries, so no atomics are needed:
//   P0 gather element data into LDS          P5 traces of M^{-1}C, M^{-1}E
//   P1 volume/face geometry                   P6 flux Jacobians by forward-mode AD,
//   P2 q operators M, C, E                       one Jacobian column per thread
//   P3 Cholesky of M (LDS)                    P7 wrl = -[s_u jac; f_u . Xx]
//   P4 M^{-1}C, M^{-1}E (column per thread)   P8 q-folded integrands (Wd, Wf, Zd, Zf)
//                                             P9 contractions with the test functions
// Element rows sum the faces that contain the row's node ("gather" form of the
// reference's per-face scatter, assembly/elementassembly.cpp:862-924,1154-1318).
#include "hfem_device.h"

namespace hfem {

#define SVT(g, n, dd) m.shapvt[(g) + d.ngv * ((n) + d.npv * (dd))]
#define SVG(n, g, dd) m.shapvg[(n) + d.npv * ((g) + d.ngv * (dd))]
#define SFT(g, k, dd) m.shapft[(g) + d.ngf * ((k) + d.npf * (dd))]
#define SFG(k, g, dd) m.shapfg[(k) + d.npf * ((g) + d.ngf * (dd))]
#define PERM(k, a) m.perm[(k) + d.npf * (a)]
#define INVP(n, a) m.invperm[(n) + d.npv * (a)]

template <int ND>
__global__ void __launch_bounds__(256)
assemble_kernel(Dims d, Master m, Phys ph, const double *__restrict__ dgnodes, const double *__restrict__ UDG,
                const double *__restrict__ UH, const double *__restrict__ SH, const int *__restrict__ elcon,
                const int *__restrict__ bf, int e0, int n, double *__restrict__ scratch, long sstride, int ldx,
                double *__restrict__ Ru, double *__restrict__ Dm, double *__restrict__ Fx,
                double *__restrict__ Km, double *__restrict__ Hx, double *__restrict__ Rh,
                double *__restrict__ ru2)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x, nth = blockDim.x;
    const int npv = d.npv, ngv = d.ngv, npf = d.npf, ngf = d.ngf, nfe = d.nfe, ndf = d.ndf;
    const int ncu = d.ncu, nch = d.nch, nc = d.nc, na = d.na, nb = d.nb, nd1 = d.nd1, np = d.np;
    constexpr int NCU = ND + 2, NC = NCU * (ND + 1);

    double *pn = lds;                         // npv*nd
    double *udg = pn + npv * ND;              // npv*nc
    double *uh = udg + npv * nc;              // ndf*nch  (uh[k*ndf + j])
    double *sh = uh + ndf * nch;              // npv*ncu
    double *jac = sh + npv * ncu;             // ngv
    double *Xx = jac + ngv;                   // ngv*nd*nd   Xx(g,x,r) = jac d xi_r / d x_x
    double *jacf = Xx + ngv * ND * ND;        // np
    double *nl = jacf + np;                   // np*nd
    double *M = nl + np * ND;                 // npv*npv
    double *red = M + npv * npv;              // 256 (reduction)
    int *bnd = (int *)(red + 256);            // nfe

    const Scratch L = scratch_layout(d);
    double *S = scratch + (long)blockIdx.x * sstride;
    double *C = S + L.C, *E = S + L.E, *shapMiC = S + L.shapMiC, *shapMiE = S + L.shapMiE;
    double *shapMiCf = S + L.shapMiCf, *shapMiEf = S + L.shapMiEf, *f = S + L.f, *fu = S + L.fu;
    double *fh = S + L.fh, *fhu = S + L.fhu, *fhuh = S + L.fhuh, *fb = S + L.fb, *fbu = S + L.fbu;
    double *fbuh = S + L.fbuh, *wrl = S + L.wrl, *Wd = S + L.Wd, *Wf = S + L.Wf;
    double *Zd = S + L.Zd, *ZdT = S + L.ZdT, *Zf = S + L.Zf, *ZfT = S + L.ZfT;
    const double tau = ph.param[5];

    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const long e = e0 + t;
        // ---------------- P0: gather ----------------
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
                const int p = i - ngv, g = p % ngf, a = p / ngf;
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
                jacf[p] = len;
                for (int x = 0; x < ND; x++) nl[p + np * x] = nv[x] / len;
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
                            int p = g + ngf * a;
                            s += (SFG(k, g, 0) * SFT(g, j, 0)) * nl[p + np * x] * jacf[p];
                        }
                    E[r] = s;
                }
            }
        }
        __syncthreads();

        // ---------------- P3: Cholesky M = U^T U in LDS ----------------
        for (int j = 0; j < npv; j++) {
            if (tid == 0) M[j + npv * j] = sqrt(M[j + npv * j]);
            __syncthreads();
            const double ujj = M[j + npv * j];
            for (int i = j + 1 + tid; i < npv; i += nth) M[j + npv * i] /= ujj;
            __syncthreads();
            const int w = npv - j - 1;
            for (int idx = tid; idx < w * w; idx += nth) {
                int k = j + 1 + idx % w, i = j + 1 + idx / w;
                if (k <= i) M[k + npv * i] -= M[j + npv * k] * M[j + npv * i];
            }
            __syncthreads();
        }

        // ---------------- P4: M^{-1} C, M^{-1} E (a column per thread) ----------------
        for (int col = tid; col < ND * (npv + ndf); col += nth) {
            double *b = col < ND * npv ? C + (long)col * npv : E + (long)(col - ND * npv) * npv;
            for (int i = 0; i < npv; i++) {
                double s = b[i];
                for (int k = 0; k < i; k++) s -= M[k + npv * i] * b[k];
                b[i] = s / M[i + npv * i];
            }
            for (int i = npv - 1; i >= 0; i--) {
                double s = b[i];
                for (int k = i + 1; k < npv; k++) s -= M[i + npv * k] * b[k];
                b[i] = s / M[i + npv * i];
            }
        }
        __syncthreads();

        // ---------------- P5: traces of M^{-1}C, M^{-1}E ----------------
        {
            const int n1 = ngv * npv * ND, n2 = ngv * ndf * ND, n3 = np * npv * ND, n4 = np * ndf * ND;
            for (int i = tid; i < n1 + n2 + n3 + n4; i += nth) {
                if (i < n1) {
                    int g = i % ngv, r = i / ngv, nn = r % npv, x = r / npv;
                    double s = 0;
                    for (int l = 0; l < npv; l++) s += SVT(g, l, 0) * C[x * npv * npv + nn * npv + l];
                    shapMiC[i] = s;
                } else if (i < n1 + n2) {
                    int r0 = i - n1, g = r0 % ngv, r = r0 / ngv, l = r % ndf, x = r / ndf;
                    double s = 0;
                    for (int k = 0; k < npv; k++) s += SVT(g, k, 0) * E[(long)x * npv * ndf + l * npv + k];
                    shapMiE[r0] = -s;
                } else if (i < n1 + n2 + n3) {
                    int r0 = i - n1 - n2, p = r0 % np, r = r0 / np, nn = r % npv, x = r / npv;
                    int g = p % ngf, a = p / ngf;
                    double s = 0;
                    for (int k = 0; k < npf; k++) s += SFT(g, k, 0) * C[x * npv * npv + nn * npv + PERM(k, a)];
                    shapMiCf[r0] = s;
                } else {
                    int r0 = i - n1 - n2 - n3, p = r0 % np, r = r0 / np, l = r % ndf, x = r / ndf;
                    int g = p % ngf, a = p / ngf;
                    double s = 0;
                    for (int k = 0; k < npf; k++) s += SFT(g, k, 0) * E[(long)x * npv * ndf + l * npv + PERM(k, a)];
                    shapMiEf[r0] = s;
                }
            }
        }

        // ---------------- P6: fluxes and Jacobian columns ----------------
        {
            const int nvol = ngv * NC, nfac = np * NC, nbnd = np * NCU;
            for (int i = tid; i < nvol + nfac + nbnd; i += nth) {
                if (i < nvol) {
                    const int g = i % ngv, k = i / ngv;
                    dual w[NC], fl[NCU * ND];
                    for (int c = 0; c < NC; c++) {
                        double s = 0;
                        for (int q = 0; q < npv; q++) s += SVT(g, q, 0) * udg[c * npv + q];
                        w[c] = mk(s, c == k ? 1.0 : 0.0);
                    }
                    ns_flux<ND>(ph.param, w, fl);
                    for (int jx = 0; jx < NCU * ND; jx++) {
                        fu[g + ngv * (jx + NCU * ND * k)] = fl[jx].d;
                        if (k == 0) f[g + ngv * jx] = fl[jx].v;
                    }
                } else if (i < nvol + nfac) {
                    const int r = i - nvol, p = r % np, k = r / np, g = p % ngf, a = p / ngf;
                    dual w[NC], fl[NCU * ND];
                    double ug[NCU];
                    for (int c = 0; c < NC; c++) {
                        double s = 0;
                        if (c < NCU) {
                            double su = 0;
                            for (int q = 0; q < npf; q++) {
                                s += SFT(g, q, 0) * uh[c * ndf + a * npf + q];
                                su += SFT(g, q, 0) * udg[c * npv + PERM(q, a)];
                            }
                            ug[c] = su;
                        } else {
                            for (int q = 0; q < npf; q++) s += SFT(g, q, 0) * udg[c * npv + PERM(q, a)];
                        }
                        w[c] = mk(s, c == k ? 1.0 : 0.0);
                    }
                    ns_flux<ND>(ph.param, w, fl);
                    for (int j = 0; j < NCU; j++) {
                        dual fn = mk(0.0);
                        for (int x = 0; x < ND; x++) fn = fn + nl[p + np * x] * fl[j + NCU * x];
                        fhu[p + np * (j + NCU * k)] = (k < NCU) ? (k == j ? tau : 0.0) : fn.d;
                        if (k < NCU) fhuh[p + np * (j + NCU * k)] = fn.d - (k == j ? tau : 0.0);
                        if (k == 0) fh[p + np * j] = fn.v + tau * (ug[j] - w[j].v);
                    }
                } else {
                    const int r = i - nvol - nfac, p = r % np, k = r / np, g = p % ngf, a = p / ngf;
                    if (!bnd[a]) continue;
                    dual uhd[NCU], A0[NCU * NCU], A1[NCU * NCU];
                    double ug[NCU], nrm[ND];
                    for (int c = 0; c < NCU; c++) {
                        double s = 0, su = 0;
                        for (int q = 0; q < npf; q++) {
                            s += SFT(g, q, 0) * uh[c * ndf + a * npf + q];
                            su += SFT(g, q, 0) * udg[c * npv + PERM(q, a)];
                        }
                        uhd[c] = mk(s, c == k ? 1.0 : 0.0);
                        ug[c] = su;
                    }
                    for (int x = 0; x < ND; x++) nrm[x] = nl[p + np * x];
                    ns_normal_jac<ND>(ph.param[0], ph.param[1], uhd, nrm, 0, A0);
                    ns_normal_jac<ND>(ph.param[0], ph.param[1], uhd, nrm, 1, A1);
                    for (int j = 0; j < NCU; j++) {
                        dual acc = mk(0.0);
                        for (int q = 0; q < NCU; q++) {
                            dual ap = A0[q * NCU + j] + A1[q * NCU + j];
                            dual am = A0[q * NCU + j] - A1[q * NCU + j];
                            acc = acc + ap * (mk(ug[q]) - uhd[q]);
                            acc = acc - am * (mk(ph.ui[q]) - uhd[q]);
                        }
                        fbuh[p + np * (j + NCU * k)] = acc.d;
                        if (k == 0) {
                            fb[p + np * j] = acc.v;
                            for (int q = 0; q < NC; q++)
                                fbu[p + np * (j + NCU * q)] = q < NCU ? A0[q * NCU + j].v + A1[q * NCU + j].v : 0.0;
                        }
                    }
                }
            }
        }
        __syncthreads();

        // ---------------- P7: wrl ----------------
        for (int i = tid; i < ngv * nd1 * ncu * nc; i += nth) {
            int g = i % ngv, r = i / ngv, mq = r % nd1, r2 = r / nd1, j = r2 % ncu, c = r2 / ncu;
            double v;
            if (mq == 0) {
                double su = (ph.tdep && c == j) ? -ph.fc_u : 0.0;
                v = -su * jac[g];
            } else {
                double s = 0;
                for (int x = 0; x < ND; x++) s += fu[g + ngv * (j + ncu * (x + ND * c))] * Xx[g + ngv * (x + ND * (mq - 1))];
                v = -s;
            }
            wrl[i] = v;
        }
        __syncthreads();

        // ---------------- P8: q-folded integrands ----------------
        {
            const long nWd = (long)ngv * nd1 * ncu * ncu * npv, nWf = (long)ngv * nd1 * ncu * ncu * ndf;
            const long nZd = (long)np * nch * ncu * npv, nZf = (long)np * nch * nch * ndf;
            const long tot = nWd + nWf + 2 * nZd + 2 * nZf;
            for (long i = tid; i < tot; i += nth) {
                if (i < nWd) {
                    int g = i % ngv; long r = i / ngv;
                    int mq = r % nd1; r /= nd1; int j = r % ncu; r /= ncu; int kc = r % ncu; int nn = r / ncu;
                    const double *wr = wrl + g + ngv * (mq + nd1 * j);
                    const long cs = (long)ngv * nd1 * ncu;
                    double v = wr[cs * kc] * SVT(g, nn, 0);
                    for (int x = 0; x < ND; x++) v += wr[cs * (ncu * (x + 1) + kc)] * shapMiC[g + ngv * (nn + npv * x)];
                    Wd[i] = v;
                } else if (i < nWd + nWf) {
                    long i2 = i - nWd;
                    int g = i2 % ngv; long r = i2 / ngv;
                    int mq = r % nd1; r /= nd1; int j = r % ncu; r /= ncu; int kc = r % ncu; int l = r / ncu;
                    const double *wr = wrl + g + ngv * (mq + nd1 * j);
                    const long cs = (long)ngv * nd1 * ncu;
                    double v = 0;
                    for (int x = 0; x < ND; x++) v += wr[cs * (ncu * (x + 1) + kc)] * shapMiE[g + ngv * (l + ndf * x)];
                    Wf[i2] = v;
                } else if (i < nWd + nWf + 2 * nZd) {
                    long i2 = i - nWd - nWf;
                    const int trace = i2 >= nZd;
                    if (trace) i2 -= nZd;
                    int p = i2 % np; long r = i2 / np;
                    int j = r % nch; r /= nch; int kc = r % ncu; int nn = r / ncu;
                    int a = p / ngf, g = p % ngf;
                    if (trace && !bnd[a]) continue;
                    const double *xu = trace ? fbu : fhu;
                    int k = INVP(nn, a);
                    double v = k >= 0 ? xu[p + np * (j + nch * kc)] * SFT(g, k, 0) : 0.0;
                    for (int x = 0; x < ND; x++)
                        v += xu[p + np * (j + nch * (ncu * (x + 1) + kc))] * shapMiCf[p + np * (nn + npv * x)];
                    (trace ? ZdT : Zd)[i2] = v;
                } else {
                    long i2 = i - nWd - nWf - 2 * nZd;
                    const int trace = i2 >= nZf;
                    if (trace) i2 -= nZf;
                    int p = i2 % np; long r = i2 / np;
                    int j = r % nch; r /= nch; int kc = r % nch; int l = r / nch;
                    int a = p / ngf, g = p % ngf;
                    if (trace && !bnd[a]) continue;
                    const double *xu = trace ? fbu : fhu, *xuh = trace ? fbuh : fhuh;
                    double v = (l / npf == a) ? xuh[p + np * (j + nch * kc)] * SFT(g, l % npf, 0) : 0.0;
                    for (int x = 0; x < ND; x++)
                        v -= xu[p + np * (j + nch * (ncu * (x + 1) + kc))] * shapMiEf[p + np * (l + ndf * x)];
                    (trace ? ZfT : Zf)[i2] = v;
                }
            }
        }
        __syncthreads();

        // ---------------- P9: contractions -> D, F, K, H, Ru, Rh ----------------
        {
            const long nD = (long)na * na, nF = (long)na * nb, nK = (long)nb * na, nH = (long)nb * nb;
            const long tot = nD + nF + nK + nH + na + nb;
            double ru2loc = 0.0;
            for (long i = tid; i < tot; i += nth) {
                if (i < nD + nF) {
                    const bool isD = i < nD;
                    long i2 = isD ? i : i - nD;
                    // D: column-major (row fastest); F: row-major in Fx (column fastest)
                    int row = isD ? (int)(i2 % na) : (int)(i2 / nb);
                    long col = isD ? i2 / na : i2 % nb;
                    int j = row / npv, ii = row % npv;
                    int kc, cc;   // column component and column node/trace index
                    if (isD) { kc = col / npv; cc = col % npv; }
                    else { kc = col % nch; cc = col / nch; }
                    const double *W = isD ? Wd + (long)ngv * nd1 * (j + ncu * (kc + ncu * (long)cc))
                                          : Wf + (long)ngv * nd1 * (j + ncu * (kc + ncu * (long)cc));
                    double s = 0;
                    for (int mq = 0; mq < nd1; mq++)
                        for (int g = 0; g < ngv; g++) s += SVG(ii, g, mq) * W[g + ngv * mq];
                    for (int a = 0; a < nfe; a++) {
                        int kk = INVP(ii, a);
                        if (kk < 0) continue;
                        const double *Z = isD ? Zd + np * (j + nch * (kc + ncu * (long)cc))
                                              : Zf + np * (j + nch * (kc + nch * (long)cc));
                        for (int g = 0; g < ngf; g++) {
                            int p = g + ngf * a;
                            s += SFG(kk, g, 0) * jacf[p] * Z[p];
                        }
                    }
                    if (isD) Dm[e * nD + i2] = s;
                    else Fx[e * (long)na * ldx + (long)row * ldx + col] = s;
                } else if (i < nD + nF + nK + nH) {
                    const bool isK = i < nD + nF + nK;
                    long i2 = isK ? i - nD - nF : i - nD - nF - nK;
                    int row = i2 % nb; long col = i2 / nb;
                    int j = row % nch, tr = row / nch, a = tr / npf, kk = tr % npf;
                    int kc, cc;
                    if (isK) { kc = col / npv; cc = col % npv; }
                    else { kc = col % nch; cc = col / nch; }
                    const double *Z = bnd[a] ? (isK ? ZdT : ZfT) : (isK ? Zd : Zf);
                    Z += isK ? np * (j + nch * (kc + ncu * (long)cc)) : np * (j + nch * (kc + nch * (long)cc));
                    double s = 0;
                    for (int g = 0; g < ngf; g++) {
                        int p = g + ngf * a;
                        s += SFG(kk, g, 0) * jacf[p] * Z[p];
                    }
                    if (isK) Km[e * nK + i2] = s;
                    else Hx[e * (nH + nb) + i2] = s;
                } else if (i < nD + nF + nK + nH + na) {
                    int row = i - (nD + nF + nK + nH), j = row / npv, ii = row % npv;
                    double s = 0;
                    for (int g = 0; g < ngv; g++) {
                        if (ph.tdep) {
                            double sv = 0, ug = 0;
                            for (int q = 0; q < npv; q++) {
                                sv += SVT(g, q, 0) * sh[j * npv + q];
                                ug += SVT(g, q, 0) * udg[j * npv + q];
                            }
                            s += SVG(ii, g, 0) * (sv - ug * ph.fc_u) * jac[g];
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
                            int p = g + ngf * a;
                            s -= SFG(kk, g, 0) * fh[p + np * j] * jacf[p];
                        }
                    }
                    Ru[e * na + row] = s;
                    Fx[e * (long)na * ldx + (long)row * ldx + nb] = s;
                    ru2loc += s * s;
                } else {
                    int row = i - (nD + nF + nK + nH + na), j = row % nch, tr = row / nch, a = tr / npf, kk = tr % npf;
                    const double *fx = bnd[a] ? fb : fh;
                    double s = 0;
                    for (int g = 0; g < ngf; g++) {
                        int p = g + ngf * a;
                        s += SFG(kk, g, 0) * fx[p + np * j] * jacf[p];
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
}

size_t assemble_lds_bytes(const Dims &d)
{
    size_t n = (size_t)d.npv * d.nd + (size_t)d.npv * d.nc + (size_t)d.ndf * d.nch + (size_t)d.npv * d.ncu + d.ngv +
               (size_t)d.ngv * d.nd * d.nd + d.np + (size_t)d.np * d.nd + (size_t)d.npv * d.npv + 256;
    return n * sizeof(double) + 64 * sizeof(int);
}

hipError_t launch_assemble(const Dims &d, const Master &m, const Phys &ph, const double *dgnodes, const double *UDG,
                           const double *UH, const double *SH, const int *elcon, const int *bf, int e0, int n,
                           double *scratch, long sstride, int grid, int ldx, double *Ru, double *D, double *Fx,
                           double *K, double *Hx, double *Rh, double *ru2, hipStream_t stream)
{
    size_t lds = assemble_lds_bytes(d);
    if (d.nd == 2)
        hipLaunchKernelGGL(assemble_kernel<2>, dim3(grid), dim3(256), lds, stream, d, m, ph, dgnodes, UDG, UH, SH,
                           elcon, bf, e0, n, scratch, sstride, ldx, Ru, D, Fx, K, Hx, Rh, ru2);
    else
        hipLaunchKernelGGL(assemble_kernel<3>, dim3(grid), dim3(256), lds, stream, d, m, ph, dgnodes, UDG, UH, SH,
                           elcon, bf, e0, n, scratch, sstride, ldx, Ru, D, Fx, K, Hx, Rh, ru2);
    return hipGetLastError();
}

}  // namespace hfem
