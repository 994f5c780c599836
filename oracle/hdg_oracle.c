/*
 * CPU FP64 oracle for the HDG element path -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates, in plain C loops, the reference's per-element algorithm:
 *   geometry      assembly/elementassembly.cpp:11-127 (elementgeom), :129-237 (facegeom)
 *   q operators   assembly/elementassembly.cpp:1323-1412 (qint)
 *   volume terms  assembly/elementassembly.cpp:239-305 (elementflux), :307-572 (elementint)
 *   face terms    assembly/elementassembly.cpp:575-677 (faceflux), :717-1321 (faceint)
 *   physics       application/FM/nsNew/flux_ns{2,3}dNEW.c (NS flux, mu = 1 from
 *                 application/FM/getViscosity.c:5-21), application/fhatDriver.cpp:54-179
 *                 (fh2fhn, fhatExpression 0), constant tau
 *                 (application/FM/stabilizationTensor.cpp:9-28), far-field boundary
 *                 application/FM/fbou.cpp:126-217 with getan (application/FM/getan.c)
 *   condensation  assembly/elementschur.cpp:6-329 (impl 0 and 1)
 *   state         assembly/elementassembly.cpp:1414-1479 (getQ)
 *   trace system  assembly/globalassembly.cpp:8-86 (scatter), kernel/matvec.m:1-12
 * Flux Jacobians are obtained by forward-mode automatic differentiation of the
 * flux (dual numbers), not from the reference's symbolic expressions, so that
 * the oracle is an independent check of the Jacobian the HIP kernels form.
 * It is linked only by the tests; the shipped library never calls it.
 */
#include "hdg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NDMAX 20

typedef struct { double v; double d[NDMAX]; } dual;

static dual dc(double v) { dual r; memset(&r, 0, sizeof r); r.v = v; return r; }
static dual dadd(int n, dual a, dual b) { dual r; r.v = a.v + b.v; for (int i = 0; i < n; i++) r.d[i] = a.d[i] + b.d[i]; return r; }
static dual dsub(int n, dual a, dual b) { dual r; r.v = a.v - b.v; for (int i = 0; i < n; i++) r.d[i] = a.d[i] - b.d[i]; return r; }
static dual dmul(int n, dual a, dual b) { dual r; r.v = a.v * b.v; for (int i = 0; i < n; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
static dual dscale(int n, double s, dual a) { dual r; r.v = s * a.v; for (int i = 0; i < n; i++) r.d[i] = s * a.d[i]; return r; }
static dual dinv(int n, dual a) { dual r; r.v = 1.0 / a.v; double g = -r.v * r.v; for (int i = 0; i < n; i++) r.d[i] = g * a.d[i]; return r; }
static dual ddiv(int n, dual a, dual b) { return dmul(n, a, dinv(n, b)); }
static dual dsqrt(int n, dual a) { dual r; r.v = sqrt(a.v); double g = 0.5 / r.v; for (int i = 0; i < n; i++) r.d[i] = g * a.d[i]; return r; }
static dual dabs(int n, dual a) { return a.v < 0 ? dscale(n, -1.0, a) : a; }

struct hfo_ctx {
    int nd, ncu, nch, nc, npv, ngv, npf, ngf, nfe, ndf, na, nb, nd1;
    double *shapvt, *shapvg, *sgl, *shapft, *shapfg, *sfgl;
    int *perm;
    double param[6], ui[8], fc_u, fc_q;
    int tdep;
    /* scratch */
    double *pg, *Jg, *jac, *Xx, *pgf, *jacf, *nl;
    double *M, *Mc, *C, *E, *MiC, *MiE, *shapMiC, *shapMiE, *shapMiCf, *shapMiEf;
    double *udgg, *f, *fu, *s, *su, *wrl;
    double *uh, *ugf, *uhg, *fh, *fhu, *fhuh, *fb, *fbu, *fbuh, *T;
};

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

hfo_ctx *hfo_create(const int *dims, const double *shapvt, const double *shapvg,
                    const double *shapvgdotshapvl, const double *shapft, const double *shapfg,
                    const double *shapfgdotshapfc, const int *perm, const double *param,
                    const double *ui, double fc_u, double fc_q, int tdep)
{
    hfo_ctx *c = (hfo_ctx *)calloc(1, sizeof(hfo_ctx));
    c->nd = dims[0]; c->ncu = dims[1]; c->npv = dims[2]; c->ngv = dims[3];
    c->npf = dims[4]; c->ngf = dims[5]; c->nfe = dims[6];
    c->nch = c->ncu; c->nd1 = c->nd + 1; c->nc = c->ncu * c->nd1;
    c->ndf = c->npf * c->nfe; c->na = c->npv * c->ncu; c->nb = c->ndf * c->nch;
    int nd = c->nd, nd1 = c->nd1, npv = c->npv, ngv = c->ngv, npf = c->npf, ngf = c->ngf;
    int nfe = c->nfe, ndf = c->ndf, ncu = c->ncu, nc = c->nc, nch = c->nch;
    size_t n;
    n = (size_t)ngv * npv * nd1; c->shapvt = dalloc(n); memcpy(c->shapvt, shapvt, n * 8);
    c->shapvg = dalloc(n); memcpy(c->shapvg, shapvg, n * 8);
    n = (size_t)npv * npv * ngv * nd1; c->sgl = dalloc(n); memcpy(c->sgl, shapvgdotshapvl, n * 8);
    n = (size_t)ngf * npf * nd; c->shapft = dalloc(n); memcpy(c->shapft, shapft, n * 8);
    c->shapfg = dalloc(n); memcpy(c->shapfg, shapfg, n * 8);
    n = (size_t)npf * npf * ngf * nd; c->sfgl = dalloc(n); memcpy(c->sfgl, shapfgdotshapfc, n * 8);
    c->perm = (int *)calloc((size_t)npf * nfe, sizeof(int)); memcpy(c->perm, perm, sizeof(int) * npf * nfe);
    memcpy(c->param, param, 6 * 8);
    memcpy(c->ui, ui, (size_t)nch * 8);
    c->fc_u = fc_u; c->fc_q = fc_q; c->tdep = tdep;
    int np = ngf * nfe;
    c->pg = dalloc((size_t)ngv * nd); c->Jg = dalloc((size_t)ngv * nd * nd); c->jac = dalloc(ngv);
    c->Xx = dalloc((size_t)ngv * nd * nd); c->pgf = dalloc((size_t)np * nd); c->jacf = dalloc(np);
    c->nl = dalloc((size_t)np * nd);
    c->M = dalloc((size_t)npv * npv); c->Mc = dalloc((size_t)npv * npv);
    c->C = dalloc((size_t)nd * npv * npv); c->E = dalloc((size_t)nd * npv * ndf);
    c->MiC = dalloc((size_t)nd * npv * npv); c->MiE = dalloc((size_t)nd * npv * ndf);
    c->shapMiC = dalloc((size_t)ngv * npv * nd); c->shapMiE = dalloc((size_t)ngv * ndf * nd);
    c->shapMiCf = dalloc((size_t)np * npv * nd); c->shapMiEf = dalloc((size_t)np * ndf * nd);
    c->udgg = dalloc((size_t)ngv * nc); c->f = dalloc((size_t)ngv * ncu * nd);
    c->fu = dalloc((size_t)ngv * ncu * nd * nc); c->s = dalloc((size_t)ngv * ncu);
    c->su = dalloc((size_t)ngv * ncu * nc); c->wrl = dalloc((size_t)ngv * nd1 * ncu * nc);
    c->uh = dalloc((size_t)ndf * nch); c->ugf = dalloc((size_t)np * nc); c->uhg = dalloc((size_t)np * nch);
    c->fh = dalloc((size_t)np * nch); c->fhu = dalloc((size_t)np * nch * nc); c->fhuh = dalloc((size_t)np * nch * nch);
    c->fb = dalloc((size_t)np * nch); c->fbu = dalloc((size_t)np * nch * nc); c->fbuh = dalloc((size_t)np * nch * nch);
    c->T = dalloc((size_t)ngv * nd1 * (npv > ndf ? npv : ndf));
    return c;
}

void hfo_destroy(hfo_ctx *c)
{
    if (!c) return;
    double **p[] = {&c->shapvt, &c->shapvg, &c->sgl, &c->shapft, &c->shapfg, &c->sfgl, &c->pg, &c->Jg,
                    &c->jac, &c->Xx, &c->pgf, &c->jacf, &c->nl, &c->M, &c->Mc, &c->C, &c->E, &c->MiC,
                    &c->MiE, &c->shapMiC, &c->shapMiE, &c->shapMiCf, &c->shapMiEf, &c->udgg, &c->f,
                    &c->fu, &c->s, &c->su, &c->wrl, &c->uh, &c->ugf, &c->uhg, &c->fh, &c->fhu, &c->fhuh,
                    &c->fb, &c->fbu, &c->fbuh, &c->T};
    for (size_t i = 0; i < sizeof p / sizeof p[0]; i++) free(*p[i]);
    free(c->perm);
    free(c);
}

/* ------------------------------------------------------------------------- */
/* physics                                                                    */
/* ------------------------------------------------------------------------- */

/* Compressible NS flux f[j + ncu*x] of state w[nc] (u then q = gradient columns),
 * as in application/FM/nsNew/flux_ns2dNEW.c:78-85 and flux_ns3dNEW.c (mu = 1):
 * inviscid part plus tau(q)/Re plus gamma/(Re Pr (gamma-1)) d(p/rho). */
static void ns_flux(int nd, int n, const double *param, const dual *w, dual *f)
{
    int ncu = nd + 2;
    double gam = param[0], Re = param[2], Pr = param[3];
    double gm1 = gam - 1.0, mu = 1.0;
    dual r = w[0], ri = dinv(n, r), rE = w[nd + 1];
    dual v[3], g[3][3], tau[3][3], th[3];
    dual ke = dc(0.0);
    for (int i = 0; i < nd; i++) { v[i] = dmul(n, w[1 + i], ri); ke = dadd(n, ke, dmul(n, v[i], v[i])); }
    ke = dscale(n, 0.5, ke);
    dual p = dscale(n, gm1, dsub(n, rE, dmul(n, r, ke)));
    for (int d = 0; d < nd; d++) {
        const dual *q = &w[ncu + d * ncu];
        for (int i = 0; i < nd; i++)
            g[i][d] = dmul(n, dsub(n, q[1 + i], dmul(n, v[i], q[0])), ri);       /* d_d v_i */
        dual t = dmul(n, dsub(n, q[nd + 1], dmul(n, dmul(n, rE, ri), q[0])), ri);
        for (int i = 0; i < nd; i++) t = dsub(n, t, dmul(n, v[i], g[i][d]));
        th[d] = dscale(n, gm1, t);                                            /* d_d (p/rho) */
    }
    dual div = dc(0.0);
    for (int i = 0; i < nd; i++) div = dadd(n, div, g[i][i]);
    for (int i = 0; i < nd; i++)
        for (int d = 0; d < nd; d++) {
            dual t = dadd(n, g[i][d], g[d][i]);
            if (i == d) t = dsub(n, t, dscale(n, 2.0 / 3.0, div));
            tau[i][d] = dscale(n, mu / Re, t);
        }
    double kappa = mu * gam / (Re * Pr * gm1);
    dual H = dmul(n, dadd(n, rE, p), ri);
    for (int d = 0; d < nd; d++) {
        f[0 + ncu * d] = w[1 + d];
        for (int i = 0; i < nd; i++) {
            dual t = dadd(n, dmul(n, w[1 + i], v[d]), tau[i][d]);
            if (i == d) t = dadd(n, t, p);
            f[1 + i + ncu * d] = t;
        }
        dual e = dmul(n, w[1 + d], H);
        for (int i = 0; i < nd; i++) e = dadd(n, e, dmul(n, v[i], tau[i][d]));
        f[nd + 1 + ncu * d] = dadd(n, e, dscale(n, kappa, th[d]));
    }
}

/* Normal flux-Jacobian A_n(uh) (absflag 0) or |A_n|(uh) (absflag 1) in the
 * eigen-split form of application/FM/getan.c (getanonly2d/3d): column k is
 * lam3 e_k + cc1_k [1, v, h] + cc2_k [0, n, un]; M[k*nch + row]. */
static void ns_normal_jac(int nd, int n, double gam, double eps, const dual *uh, const double *nrm,
                          int absflag, dual *M)
{
    int nch = nd + 2;
    double gm1 = gam - 1.0;
    dual r = uh[0], ri = dinv(n, r), v[3], af = dc(0.0), un = dc(0.0);
    for (int i = 0; i < nd; i++) {
        v[i] = dmul(n, uh[1 + i], ri);
        af = dadd(n, af, dmul(n, v[i], v[i]));
        un = dadd(n, un, dscale(n, nrm[i], v[i]));
    }
    af = dscale(n, 0.5, af);
    dual rE = uh[nd + 1], E = dmul(n, rE, ri);
    dual p = dscale(n, gm1, dsub(n, rE, dmul(n, r, af)));
    dual h = dadd(n, E, dmul(n, p, ri));
    dual c2 = dscale(n, gam, dmul(n, p, ri)), cs = dsqrt(n, c2);
    dual l1 = dadd(n, un, cs), l2 = dsub(n, un, cs), l3 = un;
    if (absflag) {
        dual *ls[3] = {&l1, &l2, &l3};
        for (int k = 0; k < 3; k++) {
            dual a = dabs(n, *ls[k]);
            if (eps > 0) {
                dual ec = dscale(n, eps, cs);
                if (a.v < ec.v) a = dscale(n, 0.5, dadd(n, ddiv(n, dmul(n, a, a), ec), ec));
            }
            *ls[k] = a;
        }
    }
    dual s1 = dscale(n, 0.5, dadd(n, l1, l2)), s2 = dscale(n, 0.5, dsub(n, l1, l2));
    dual s13 = dsub(n, s1, l3);
    for (int k = 0; k < nch; k++) {
        dual cc1, cc2;
        if (k == 0) {
            cc1 = dsub(n, ddiv(n, dscale(n, gm1, dmul(n, s13, af)), c2), ddiv(n, dmul(n, s2, un), cs));
            cc2 = dsub(n, ddiv(n, dscale(n, gm1, dmul(n, s2, af)), cs), dmul(n, s13, un));
        } else if (k <= nd) {
            dual vi = v[k - 1];
            double ni = nrm[k - 1];
            cc1 = dadd(n, dscale(n, -gm1, ddiv(n, dmul(n, s13, vi), c2)), dscale(n, ni, ddiv(n, s2, cs)));
            cc2 = dadd(n, dscale(n, -gm1, ddiv(n, dmul(n, s2, vi), cs)), dscale(n, ni, s13));
        } else {
            cc1 = dscale(n, gm1, ddiv(n, s13, c2));
            cc2 = dscale(n, gm1, ddiv(n, s2, cs));
        }
        M[k * nch + 0] = cc1;
        for (int m = 0; m < nd; m++) M[k * nch + 1 + m] = dadd(n, dmul(n, cc1, v[m]), dscale(n, nrm[m], cc2));
        M[k * nch + nd + 1] = dadd(n, dmul(n, cc1, h), dmul(n, cc2, un));
        M[k * nch + k] = dadd(n, M[k * nch + k], l3);
    }
}

/* far-field flux (application/FM/fbou.cpp:126-192) at one point:
 * fb = (A+|A|)(u-uh) - (A-|A|)(ui-uh), with A, |A| evaluated at uh. */
static void fbou_point(const hfo_ctx *c, const double *u, const double *uhv, const double *nrm,
                       double *fb, double *fbu, double *fbuh, int np, int ip)
{
    int nd = c->nd, nch = c->nch, nc = c->nc, n = nch;
    dual uh[8], A0[64], A1[64];
    for (int k = 0; k < nch; k++) { uh[k] = dc(uhv[k]); uh[k].d[k] = 1.0; }
    ns_normal_jac(nd, n, c->param[0], c->param[1], uh, nrm, 0, A0);
    ns_normal_jac(nd, n, c->param[0], c->param[1], uh, nrm, 1, A1);
    for (int j = 0; j < nch; j++) {
        dual acc = dc(0.0);
        for (int k = 0; k < nch; k++) {
            dual ap = dadd(n, A0[k * nch + j], A1[k * nch + j]);
            dual am = dsub(n, A0[k * nch + j], A1[k * nch + j]);
            acc = dadd(n, acc, dmul(n, ap, dsub(n, dc(u[k]), uh[k])));
            acc = dsub(n, acc, dmul(n, am, dsub(n, dc(c->ui[k]), uh[k])));
        }
        fb[ip + np * j] = acc.v;
        for (int k = 0; k < nc; k++)
            fbu[ip + np * (j + nch * k)] = (k < nch) ? A0[k * nch + j].v + A1[k * nch + j].v : 0.0;
        for (int k = 0; k < nch; k++) fbuh[ip + np * (j + nch * k)] = acc.d[k];
    }
}

void hfo_flux(hfo_ctx *c, int np, const double *udg, double *f, double *f_udg)
{
    int nd = c->nd, ncu = c->ncu, nc = c->nc;
    dual w[NDMAX], fl[15];
    for (int i = 0; i < np; i++) {
        for (int k = 0; k < nc; k++) { w[k] = dc(udg[k * np + i]); w[k].d[k] = 1.0; }
        ns_flux(nd, nc, c->param, w, fl);
        for (int jx = 0; jx < ncu * nd; jx++) {
            f[i + np * jx] = fl[jx].v;
            for (int k = 0; k < nc; k++) f_udg[i + np * (jx + ncu * nd * k)] = fl[jx].d[k];
        }
    }
}

void hfo_fbou(hfo_ctx *c, int np, const double *udg, const double *uh, const double *nl,
              double *fb, double *fb_u, double *fb_uh)
{
    int nd = c->nd, nch = c->nch;
    double u[8], uhv[8], nrm[3];
    for (int i = 0; i < np; i++) {
        for (int k = 0; k < nch; k++) { u[k] = udg[k * np + i]; uhv[k] = uh[k * np + i]; }
        for (int d = 0; d < nd; d++) nrm[d] = nl[d * np + i];
        fbou_point(c, u, uhv, nrm, fb, fb_u, fb_uh, np, i);
    }
}

/* ------------------------------------------------------------------------- */
/* small dense helpers                                                        */
/* ------------------------------------------------------------------------- */

/* upper Cholesky M = U^T U (DPOTRF 'U') in place; returns 0 on success */
static int chol_u(int n, double *A)
{
    for (int j = 0; j < n; j++) {
        double s = A[j + n * j];
        for (int k = 0; k < j; k++) s -= A[k + n * j] * A[k + n * j];
        if (s <= 0) return j + 1;
        double ujj = sqrt(s);
        A[j + n * j] = ujj;
        for (int i = j + 1; i < n; i++) {
            double t = A[j + n * i];
            for (int k = 0; k < j; k++) t -= A[k + n * j] * A[k + n * i];
            A[j + n * i] = t / ujj;
        }
    }
    return 0;
}

/* solve U^T U X = B for nrhs columns (DPOTRS 'U') */
static void chol_solve(int n, const double *U, int nrhs, double *B, int ldb)
{
    for (int r = 0; r < nrhs; r++) {
        double *b = B + (size_t)ldb * r;
        for (int i = 0; i < n; i++) {
            double t = b[i];
            for (int k = 0; k < i; k++) t -= U[k + n * i] * b[k];
            b[i] = t / U[i + n * i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double t = b[i];
            for (int k = i + 1; k < n; k++) t -= U[i + n * k] * b[k];
            b[i] = t / U[i + n * i];
        }
    }
}

/* LU with partial pivoting (DGETRF semantics), A n x n with leading dim lda */
static int lu_pp(int n, double *A, int lda, int *piv)
{
    int info = 0;
    for (int j = 0; j < n; j++) {
        int p = j;
        double mx = fabs(A[j + (size_t)lda * j]);
        for (int i = j + 1; i < n; i++)
            if (fabs(A[i + (size_t)lda * j]) > mx) { mx = fabs(A[i + (size_t)lda * j]); p = i; }
        piv[j] = p;
        if (A[p + (size_t)lda * j] == 0.0) { if (!info) info = j + 1; continue; }
        if (p != j)
            for (int k = 0; k < n; k++) {
                double t = A[j + (size_t)lda * k]; A[j + (size_t)lda * k] = A[p + (size_t)lda * k]; A[p + (size_t)lda * k] = t;
            }
        double inv = 1.0 / A[j + (size_t)lda * j];
        for (int i = j + 1; i < n; i++) A[i + (size_t)lda * j] *= inv;
        for (int k = j + 1; k < n; k++) {
            double akj = A[j + (size_t)lda * k];
            if (akj != 0.0)
                for (int i = j + 1; i < n; i++) A[i + (size_t)lda * k] -= A[i + (size_t)lda * j] * akj;
        }
    }
    return info;
}

/* solve with LU factors (DGETRS 'N') */
static void lu_solve(int n, const double *A, int lda, const int *piv, int nrhs, double *B, int ldb)
{
    for (int r = 0; r < nrhs; r++) {
        double *b = B + (size_t)ldb * r;
        for (int j = 0; j < n; j++) if (piv[j] != j) { double t = b[j]; b[j] = b[piv[j]]; b[piv[j]] = t; }
        for (int j = 0; j < n; j++) { double bj = b[j]; if (bj != 0.0) for (int i = j + 1; i < n; i++) b[i] -= A[i + (size_t)lda * j] * bj; }
        for (int j = n - 1; j >= 0; j--) {
            b[j] /= A[j + (size_t)lda * j];
            double bj = b[j];
            if (bj != 0.0) for (int i = 0; i < j; i++) b[i] -= A[i + (size_t)lda * j] * bj;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* element assembly                                                           */
/* ------------------------------------------------------------------------- */

#define SVT(g, n, d) c->shapvt[(g) + ngv * ((n) + npv * (d))]
#define SVG(n, g, d) c->shapvg[(n) + npv * ((g) + ngv * (d))]
#define SFT(g, k, d) c->shapft[(g) + ngf * ((k) + npf * (d))]
#define SFG(k, g, d) c->shapfg[(k) + npf * ((g) + ngf * (d))]
#define PERM(k, a) c->perm[(k) + npf * (a)]

/* elementgeom + facegeom (assembly/elementassembly.cpp:11-237) */
static void geometry(hfo_ctx *c, const double *pn)
{
    int nd = c->nd, npv = c->npv, ngv = c->ngv, npf = c->npf, ngf = c->ngf, nfe = c->nfe;
    int np = ngf * nfe;
    for (int g = 0; g < ngv; g++) {
        double J[3][3] = {{0}};   /* J[x][r] = d x_x / d xi_r */
        for (int x = 0; x < nd; x++) {
            double s = 0;
            for (int n = 0; n < npv; n++) s += SVT(g, n, 0) * pn[x * npv + n];
            c->pg[g + ngv * x] = s;
            for (int r = 0; r < nd; r++) {
                double t = 0;
                for (int n = 0; n < npv; n++) t += SVT(g, n, 1 + r) * pn[x * npv + n];
                J[x][r] = t;
            }
        }
        double det, adj[3][3];   /* adj[r][x] = det * d xi_r / d x_x */
        if (nd == 2) {
            det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            adj[0][0] = J[1][1]; adj[0][1] = -J[0][1]; adj[1][0] = -J[1][0]; adj[1][1] = J[0][0];
        } else {
            for (int r = 0; r < 3; r++)
                for (int x = 0; x < 3; x++) {
                    int x1 = (x + 1) % 3, x2 = (x + 2) % 3, r1 = (r + 1) % 3, r2 = (r + 2) % 3;
                    adj[r][x] = J[x1][r1] * J[x2][r2] - J[x1][r2] * J[x2][r1];
                }
            det = J[0][0] * adj[0][0] + J[1][0] * adj[0][1] + J[2][0] * adj[0][2];
        }
        c->jac[g] = det;
        for (int x = 0; x < nd; x++)
            for (int r = 0; r < nd; r++) c->Xx[g + ngv * (x + nd * r)] = adj[r][x];
    }
    for (int a = 0; a < nfe; a++)
        for (int g = 0; g < ngf; g++) {
            int p = g + ngf * a;
            double t[2][3] = {{0}};
            for (int x = 0; x < nd; x++) {
                double s = 0;
                for (int k = 0; k < npf; k++) s += SFT(g, k, 0) * pn[x * npv + PERM(k, a)];
                c->pgf[p + np * x] = s;
                for (int r = 0; r < nd - 1; r++) {
                    double u = 0;
                    for (int k = 0; k < npf; k++) u += SFT(g, k, 1 + r) * pn[x * npv + PERM(k, a)];
                    t[r][x] = u;
                }
            }
            double nv[3];
            if (nd == 2) { nv[0] = t[0][1]; nv[1] = -t[0][0]; }
            else {
                nv[0] = t[0][1] * t[1][2] - t[0][2] * t[1][1];
                nv[1] = t[0][2] * t[1][0] - t[0][0] * t[1][2];
                nv[2] = t[0][0] * t[1][1] - t[0][1] * t[1][0];
            }
            double len = 0;
            for (int x = 0; x < nd; x++) len += nv[x] * nv[x];
            len = sqrt(len);
            c->jacf[p] = len;
            for (int x = 0; x < nd; x++) c->nl[p + np * x] = nv[x] / len;
        }
}

/* qint (assembly/elementassembly.cpp:1323-1412): M, C_i, E_i */
static void qoperators(hfo_ctx *c)
{
    int nd = c->nd, npv = c->npv, ngv = c->ngv, npf = c->npf, ngf = c->ngf, nfe = c->nfe, ndf = c->ndf;
    int n2 = npv * npv;
    for (int r = 0; r < n2; r++) {
        double s = 0;
        for (int g = 0; g < ngv; g++) s += c->sgl[r + n2 * g] * c->jac[g];
        c->M[r] = c->fc_q * s;
    }
    for (int i = 0; i < nd; i++)
        for (int r = 0; r < n2; r++) {
            double s = 0;
            for (int rr = 0; rr < nd; rr++)
                for (int g = 0; g < ngv; g++)
                    s += c->sgl[r + n2 * (g + ngv * (rr + 1))] * c->Xx[g + ngv * (i + nd * rr)];
            c->C[i * n2 + r] = s;
        }
    memset(c->E, 0, sizeof(double) * nd * npv * ndf);
    int f2 = npf * npf, np = ngf * nfe;
    for (int i = 0; i < nd; i++)
        for (int a = 0; a < nfe; a++)
            for (int j = 0; j < npf; j++)
                for (int k = 0; k < npf; k++) {
                    double s = 0;
                    for (int g = 0; g < ngf; g++) {
                        int p = g + ngf * a;
                        s += c->sfgl[(j * npf + k) + f2 * g] * c->nl[p + np * i] * c->jacf[p];
                    }
                    c->E[(size_t)i * npv * ndf + (a * npf + j) * npv + PERM(k, a)] = s;
                }
}

static void gather_uh(hfo_ctx *c, const double *UH, const int *elc)
{
    for (int j = 0; j < c->ndf; j++)
        for (int k = 0; k < c->nch; k++) c->uh[k * c->ndf + j] = UH[(size_t)elc[j] * c->nch + k];
}

static void assemble_one(hfo_ctx *c, const double *pn, const double *udg, const double *UH,
                         const double *sh, const int *elc, const int *bfe,
                         double *Ru, double *D, double *F, double *K, double *H, double *Rh)
{
    int nd = c->nd, nd1 = c->nd1, npv = c->npv, ngv = c->ngv, npf = c->npf, ngf = c->ngf;
    int nfe = c->nfe, ndf = c->ndf, ncu = c->ncu, nch = c->nch, nc = c->nc, na = c->na, nb = c->nb;
    int np = ngf * nfe, f2 = npf * npf;
    double tau = c->param[5];

    geometry(c, pn);
    qoperators(c);

    /* Cholesky of M and M^{-1} C, M^{-1} E (elementint, :463-478) */
    memcpy(c->Mc, c->M, sizeof(double) * npv * npv);
    chol_u(npv, c->Mc);
    memcpy(c->MiC, c->C, sizeof(double) * nd * npv * npv);
    chol_solve(npv, c->Mc, npv * nd, c->MiC, npv);
    memcpy(c->MiE, c->E, sizeof(double) * nd * npv * ndf);
    chol_solve(npv, c->Mc, ndf * nd, c->MiE, npv);
    for (int d = 0; d < nd; d++)
        for (int g = 0; g < ngv; g++) {
            for (int n = 0; n < npv; n++) {
                double s = 0;
                for (int l = 0; l < npv; l++) s += SVT(g, l, 0) * c->MiC[d * npv * npv + n * npv + l];
                c->shapMiC[g + ngv * (n + npv * d)] = s;
            }
            for (int l = 0; l < ndf; l++) {
                double s = 0;
                for (int n = 0; n < npv; n++) s += SVT(g, n, 0) * c->MiE[(size_t)d * npv * ndf + l * npv + n];
                c->shapMiE[g + ngv * (l + ndf * d)] = -s;
            }
        }

    /* volume flux and its Jacobian at Gauss points (elementflux, :239-305) */
    {
        dual w[NDMAX], fl[15];
        for (int g = 0; g < ngv; g++) {
            for (int k = 0; k < nc; k++) {
                double s = 0;
                for (int n = 0; n < npv; n++) s += SVT(g, n, 0) * udg[k * npv + n];
                c->udgg[g + ngv * k] = s;
                w[k] = dc(s); w[k].d[k] = 1.0;
            }
            ns_flux(nd, nc, c->param, w, fl);
            for (int jx = 0; jx < ncu * nd; jx++) {
                c->f[g + ngv * jx] = fl[jx].v;
                for (int k = 0; k < nc; k++) c->fu[g + ngv * (jx + ncu * nd * k)] = fl[jx].d[k];
            }
            for (int j = 0; j < ncu; j++) {
                c->s[g + ngv * j] = 0.0;                      /* NS source is zero */
                for (int k = 0; k < nc; k++) c->su[g + ngv * (j + ncu * k)] = 0.0;
            }
            if (c->tdep) {
                for (int j = 0; j < ncu; j++) {
                    double s = 0;
                    for (int n = 0; n < npv; n++) s += SVT(g, n, 0) * sh[j * npv + n];
                    c->s[g + ngv * j] += s - c->udgg[g + ngv * j] * c->fc_u;
                    c->su[g + ngv * (j + ncu * j)] -= c->fc_u;
                }
            }
        }
    }

    /* Ru = int v s + grad v . f (elementint, :392-410) */
    for (int j = 0; j < ncu; j++)
        for (int i = 0; i < npv; i++) {
            double s = 0;
            for (int g = 0; g < ngv; g++) {
                s += SVG(i, g, 0) * c->s[g + ngv * j] * c->jac[g];
                for (int m = 0; m < nd; m++) {
                    double t = 0;
                    for (int x = 0; x < nd; x++) t += c->f[g + ngv * (j + ncu * x)] * c->Xx[g + ngv * (x + nd * m)];
                    s += SVG(i, g, m + 1) * t;
                }
            }
            Ru[j * npv + i] = s;
        }

    /* wrl(g, m', j, k) = -[s_udg jac ; f_udg . Xx] (elementint, :418-436) */
#define WRL(g, m, j, k) c->wrl[(g) + ngv * ((m) + nd1 * ((j) + ncu * (k)))]
    for (int k = 0; k < nc; k++)
        for (int j = 0; j < ncu; j++)
            for (int g = 0; g < ngv; g++) {
                WRL(g, 0, j, k) = -c->su[g + ngv * (j + ncu * k)] * c->jac[g];
                for (int m = 0; m < nd; m++) {
                    double t = 0;
                    for (int x = 0; x < nd; x++)
                        t += c->fu[g + ngv * (j + ncu * (x + nd * k))] * c->Xx[g + ngv * (x + nd * m)];
                    WRL(g, m + 1, j, k) = -t;
                }
            }

    /* D and F from the volume, with q folded in through M^{-1}C, M^{-1}E
     * (elementint, :510-569) */
    double *T = c->T;
    for (int j = 0; j < ncu; j++)
        for (int k = 0; k < ncu; k++) {
            for (int g = 0; g < ngv; g++)
                for (int m = 0; m < nd1; m++)
                    for (int n = 0; n < npv; n++) {
                        double t = WRL(g, m, j, k) * SVT(g, n, 0);
                        for (int d = 0; d < nd; d++) t += WRL(g, m, j, ncu * (d + 1) + k) * c->shapMiC[g + ngv * (n + npv * d)];
                        T[(g + ngv * m) + ngv * nd1 * n] = t;
                    }
            for (int n = 0; n < npv; n++)
                for (int i = 0; i < npv; i++) {
                    double s = 0;
                    for (int m = 0; m < nd1; m++)
                        for (int g = 0; g < ngv; g++) s += SVG(i, g, m) * T[(g + ngv * m) + ngv * nd1 * n];
                    D[(size_t)(j * npv + i) + (size_t)na * (k * npv + n)] = s;
                }
            for (int g = 0; g < ngv; g++)
                for (int m = 0; m < nd1; m++)
                    for (int l = 0; l < ndf; l++) {
                        double t = 0;
                        for (int d = 0; d < nd; d++) t += WRL(g, m, j, ncu * (d + 1) + k) * c->shapMiE[g + ngv * (l + ndf * d)];
                        T[(g + ngv * m) + ngv * nd1 * l] = t;
                    }
            for (int l = 0; l < ndf; l++)
                for (int i = 0; i < npv; i++) {
                    double s = 0;
                    for (int m = 0; m < nd1; m++)
                        for (int g = 0; g < ngv; g++) s += SVG(i, g, m) * T[(g + ngv * m) + ngv * nd1 * l];
                    F[(size_t)(j * npv + i) + (size_t)na * (l * nch + k)] = s;
                }
        }

    /* face states (faceflux, :575-677) */
    gather_uh(c, UH, elc);
    for (int a = 0; a < nfe; a++)
        for (int g = 0; g < ngf; g++) {
            int p = g + ngf * a;
            for (int k = 0; k < nc; k++) {
                double s = 0;
                for (int kk = 0; kk < npf; kk++) s += SFT(g, kk, 0) * udg[k * npv + PERM(kk, a)];
                c->ugf[p + np * k] = s;
            }
            for (int k = 0; k < nch; k++) {
                double s = 0;
                for (int kk = 0; kk < npf; kk++) s += SFT(g, kk, 0) * c->uh[k * ndf + a * npf + kk];
                c->uhg[p + np * k] = s;
            }
        }
    /* interior trace flux fhat = F(uh, q).n + tau (u - uh) (fhatDriver/fh2fhn) */
    {
        dual w[NDMAX], fl[15];
        for (int p = 0; p < np; p++) {
            for (int k = 0; k < nc; k++) {
                w[k] = dc(k < ncu ? c->uhg[p + np * k] : c->ugf[p + np * k]);
                w[k].d[k] = 1.0;
            }
            ns_flux(nd, nc, c->param, w, fl);
            for (int j = 0; j < nch; j++) {
                dual fn = dc(0.0);
                for (int x = 0; x < nd; x++) fn = dadd(nc, fn, dscale(nc, c->nl[p + np * x], fl[j + ncu * x]));
                c->fh[p + np * j] = fn.v + tau * (c->ugf[p + np * j] - c->uhg[p + np * j]);
                for (int k = 0; k < nc; k++)
                    c->fhu[p + np * (j + nch * k)] = (k < ncu) ? (k == j ? tau : 0.0) : fn.d[k];
                for (int k = 0; k < nch; k++)
                    c->fhuh[p + np * (j + nch * k)] = fn.d[k] - (k == j ? tau : 0.0);
            }
        }
    }
    /* boundary flux on boundary faces (faceint, :1074-1246) */
    for (int a = 0; a < nfe; a++) {
        if (bfe[a] >= 0) continue;
        for (int g = 0; g < ngf; g++) {
            int p = g + ngf * a;
            double u[8], uhv[8], nrm[3];
            for (int k = 0; k < nch; k++) { u[k] = c->ugf[p + np * k]; uhv[k] = c->uhg[p + np * k]; }
            for (int x = 0; x < nd; x++) nrm[x] = c->nl[p + np * x];
            fbou_point(c, u, uhv, nrm, c->fb, c->fbu, c->fbuh, np, p);
        }
    }
    /* traces of M^{-1}C, M^{-1}E on the faces (faceint, :927-960) */
    for (int d = 0; d < nd; d++)
        for (int a = 0; a < nfe; a++)
            for (int g = 0; g < ngf; g++) {
                int p = g + ngf * a;
                for (int n = 0; n < npv; n++) {
                    double s = 0;
                    for (int k = 0; k < npf; k++) s += SFT(g, k, 0) * c->MiC[d * npv * npv + n * npv + PERM(k, a)];
                    c->shapMiCf[p + np * (n + npv * d)] = s;
                }
                for (int l = 0; l < ndf; l++) {
                    double s = 0;
                    for (int k = 0; k < npf; k++) s += SFT(g, k, 0) * c->MiE[(size_t)d * npv * ndf + l * npv + PERM(k, a)];
                    c->shapMiEf[p + np * (l + ndf * d)] = s;
                }
            }

    memset(K, 0, sizeof(double) * nb * na);
    memset(H, 0, sizeof(double) * nb * nb);
    for (int a = 0; a < nfe; a++) {
        int bnd = bfe[a] < 0;
        for (int pass = 0; pass < 2; pass++) {
            /* pass 0: element rows with the interior fhat; pass 1: trace rows with
             * fhat (interior face) or fb (boundary face) */
            const double *fx = (pass == 1 && bnd) ? c->fb : c->fh;
            const double *fxu = (pass == 1 && bnd) ? c->fbu : c->fhu;
            const double *fxuh = (pass == 1 && bnd) ? c->fbuh : c->fhuh;
            for (int j = 0; j < nch; j++)
                for (int k = 0; k < npf; k++) {
                    double s = 0;
                    for (int g = 0; g < ngf; g++) { int p = g + ngf * a; s += SFG(k, g, 0) * fx[p + np * j] * c->jacf[p]; }
                    if (pass == 0) Ru[j * npv + PERM(k, a)] -= s;
                    else Rh[(a * npf + k) * nch + j] = -s;
                }
            for (int j = 0; j < nch; j++)
                for (int m = 0; m < nch; m++)
                    for (int k = 0; k < npf; k++)
                        for (int ks = 0; ks < npf; ks++) {
                            double su = 0, suh = 0;
                            for (int g = 0; g < ngf; g++) {
                                int p = g + ngf * a;
                                double w = c->sfgl[(k * npf + ks) + f2 * g] * c->jacf[p];
                                su += w * fxu[p + np * (j + nch * m)];
                                suh += w * fxuh[p + np * (j + nch * m)];
                            }
                            if (pass == 0) {
                                D[(size_t)(j * npv + PERM(ks, a)) + (size_t)na * (m * npv + PERM(k, a))] += su;
                                F[(size_t)(j * npv + PERM(ks, a)) + (size_t)na * ((a * npf + k) * nch + m)] += suh;
                            } else {
                                K[(size_t)((a * npf + ks) * nch + j) + (size_t)nb * (m * npv + PERM(k, a))] = su;
                                H[(size_t)((a * npf + ks) * nch + j) + (size_t)nb * ((a * npf + k) * nch + m)] = suh;
                            }
                        }
            /* q folded into the face terms: Df, Ff (faceint, :962-1072, :1298-1318) */
            for (int j = 0; j < nch; j++)
                for (int kc = 0; kc < ncu; kc++)
                    for (int kk = 0; kk < npf; kk++) {
                        for (int n = 0; n < npv; n++) {
                            double s = 0;
                            for (int g = 0; g < ngf; g++) {
                                int p = g + ngf * a;
                                double t = 0;
                                for (int d = 0; d < nd; d++)
                                    t += fxu[p + np * (j + nch * (ncu * (d + 1) + kc))] * c->shapMiCf[p + np * (n + npv * d)];
                                s += SFG(kk, g, 0) * t * c->jacf[p];
                            }
                            if (pass == 0) D[(size_t)(j * npv + PERM(kk, a)) + (size_t)na * (kc * npv + n)] += s;
                            else K[(size_t)((a * npf + kk) * nch + j) + (size_t)nb * (kc * npv + n)] += s;
                        }
                        for (int l = 0; l < ndf; l++) {
                            double s = 0;
                            for (int g = 0; g < ngf; g++) {
                                int p = g + ngf * a;
                                double t = 0;
                                for (int d = 0; d < nd; d++)
                                    t += fxu[p + np * (j + nch * (ncu * (d + 1) + kc))] * c->shapMiEf[p + np * (l + ndf * d)];
                                s += SFG(kk, g, 0) * t * c->jacf[p];
                            }
                            if (pass == 0) F[(size_t)(j * npv + PERM(kk, a)) + (size_t)na * (l * nch + kc)] -= s;
                            else H[(size_t)((a * npf + kk) * nch + j) + (size_t)nb * (l * nch + kc)] -= s;
                        }
                    }
        }
    }
#undef WRL
}

int hfo_assemble(hfo_ctx *c, int e0, int n, const double *dgnodes, const double *UDG,
                 const double *UH, const double *SH, const int *elcon, const int *bf,
                 double *Ru, double *D, double *F, double *K, double *H, double *Rh)
{
    int nd = c->nd, npv = c->npv, nc = c->nc, ndf = c->ndf, nfe = c->nfe, na = c->na, nb = c->nb;
    for (int t = 0; t < n; t++) {
        int ie = e0 + t;
        assemble_one(c, dgnodes + (size_t)ie * npv * nd, UDG + (size_t)ie * npv * nc, UH,
                     SH + (size_t)ie * npv * nc, elcon + (size_t)ie * ndf, bf + (size_t)ie * nfe,
                     Ru + (size_t)t * na, D + (size_t)t * na * na, F + (size_t)t * na * nb,
                     K + (size_t)t * nb * na, H + (size_t)t * nb * nb, Rh + (size_t)t * nb);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* condensation (assembly/elementschur.cpp)                                   */
/* ------------------------------------------------------------------------- */

/* inverse of an m x m block in place (DGETRF + DGETRI semantics) */
static void block_inverse(int m, double *A, int lda, double *work, int *piv)
{
    for (int j = 0; j < m; j++) for (int i = 0; i < m; i++) work[i + m * j] = A[i + (size_t)lda * j];
    lu_pp(m, work, m, piv);
    for (int j = 0; j < m; j++) {
        double *col = A + (size_t)lda * j;
        for (int i = 0; i < m; i++) col[i] = (i == j);
    }
    /* solve in place column by column into A */
    double *tmp = work + m * m;
    for (int j = 0; j < m; j++) {
        for (int i = 0; i < m; i++) tmp[i] = (i == j);
        lu_solve(m, work, m, piv, 1, tmp, m);
        for (int i = 0; i < m; i++) A[i + (size_t)lda * j] = tmp[i];
    }
}

/* block LU over nbk x nbk blocks of size m (elementschur.cpp:106-186): diagonal
 * blocks replaced by their inverses, L_ji = A_ji inv(A_ii), A_jk -= L_ji A_ik */
static void block_lu(int nbk, int m, double *A, int lda, double *work, int *piv)
{
#define BLK(i, j) (A + (size_t)(i) * m + (size_t)lda * (j) * m)
    double *tmp = work + 2 * m * m + m;
    for (int i = 0; i < nbk; i++) {
        block_inverse(m, BLK(i, i), lda, work, piv);
        for (int j = i + 1; j < nbk; j++) {
            double *L = BLK(j, i), *Di = BLK(i, i);
            for (int cc = 0; cc < m; cc++)
                for (int r = 0; r < m; r++) {
                    double s = 0;
                    for (int t = 0; t < m; t++) s += L[r + (size_t)lda * t] * Di[t + (size_t)lda * cc];
                    tmp[r + m * cc] = s;
                }
            for (int cc = 0; cc < m; cc++) for (int r = 0; r < m; r++) L[r + (size_t)lda * cc] = tmp[r + m * cc];
            for (int k = i + 1; k < nbk; k++) {
                double *U = BLK(i, k), *Ajk = BLK(j, k);
                for (int cc = 0; cc < m; cc++)
                    for (int r = 0; r < m; r++) {
                        double s = 0;
                        for (int t = 0; t < m; t++) s += L[r + (size_t)lda * t] * U[t + (size_t)lda * cc];
                        Ajk[r + (size_t)lda * cc] -= s;
                    }
            }
        }
    }
#undef BLK
}

/* forward/backward solve with the block factors (elementschur.cpp:188-215, :289-323) */
static void block_solve(int nbk, int m, const double *A, int lda, int nrhs, double *B, int ldb, double *tmp)
{
#define BLK(i, j) (A + (size_t)(i) * m + (size_t)lda * (j) * m)
    for (int r = 0; r < nrhs; r++) {
        double *b = B + (size_t)ldb * r;
        for (int i = 0; i < nbk; i++)
            for (int k = 0; k < i; k++) {
                const double *L = BLK(i, k);
                for (int rr = 0; rr < m; rr++) {
                    double s = 0;
                    for (int t = 0; t < m; t++) s += L[rr + (size_t)lda * t] * b[k * m + t];
                    b[i * m + rr] -= s;
                }
            }
        for (int i = nbk - 1; i >= 0; i--) {
            for (int k = i + 1; k < nbk; k++) {
                const double *U = BLK(i, k);
                for (int rr = 0; rr < m; rr++) {
                    double s = 0;
                    for (int t = 0; t < m; t++) s += U[rr + (size_t)lda * t] * b[k * m + t];
                    b[i * m + rr] -= s;
                }
            }
            const double *Di = BLK(i, i);
            for (int rr = 0; rr < m; rr++) {
                double s = 0;
                for (int t = 0; t < m; t++) s += Di[rr + (size_t)lda * t] * b[i * m + t];
                tmp[rr] = s;
            }
            for (int rr = 0; rr < m; rr++) b[i * m + rr] = tmp[rr];
        }
    }
#undef BLK
}

int hfo_schur(int impl, int n, int npv, int ncu, int nb, double *D, double *F, double *K,
              double *H, double *Ru, double *Rh)
{
    int na = npv * ncu;
    int *piv = (int *)calloc((size_t)na, sizeof(int));
    double *work = dalloc((size_t)4 * npv * npv + 4 * na);
    int info = 0;
    for (int t = 0; t < n; t++) {
        double *De = D + (size_t)t * na * na, *Fe = F + (size_t)t * na * nb, *Ke = K + (size_t)t * nb * na;
        double *He = H + (size_t)t * nb * nb, *Rue = Ru + (size_t)t * na, *Rhe = Rh + (size_t)t * nb;
        if (impl == 0) {
            int e = lu_pp(na, De, na, piv);
            if (e && !info) info = e;
            lu_solve(na, De, na, piv, nb, Fe, na);
            lu_solve(na, De, na, piv, 1, Rue, na);
        } else {
            block_lu(ncu, npv, De, na, work, piv);
            block_solve(ncu, npv, De, na, nb, Fe, na, work);
            block_solve(ncu, npv, De, na, 1, Rue, na, work);
        }
        /* H <- H - K inv(D) F ; Rh <- Rh - K inv(D) Ru */
        for (int j = 0; j < nb; j++)
            for (int i = 0; i < nb; i++) {
                double s = 0;
                for (int k = 0; k < na; k++) s += Ke[i + (size_t)nb * k] * Fe[k + (size_t)na * j];
                He[i + (size_t)nb * j] -= s;
            }
        for (int i = 0; i < nb; i++) {
            double s = 0;
            for (int k = 0; k < na; k++) s += Ke[i + (size_t)nb * k] * Rue[k];
            Rhe[i] -= s;
        }
    }
    free(piv);
    free(work);
    return info;
}

/* getQ (assembly/elementassembly.cpp:1414-1479):
 * q_d = M^{-1} (M sh_q + C_d u - E_d uh) */
int hfo_getq(hfo_ctx *c, int e0, int n, const double *dgnodes, double *UDG, const double *UH,
             const double *SH, const int *elcon)
{
    int nd = c->nd, npv = c->npv, nc = c->nc, ncu = c->ncu, ndf = c->ndf;
    double *rhs = dalloc((size_t)npv * ncu * nd);
    for (int t = 0; t < n; t++) {
        int ie = e0 + t;
        const double *pn = dgnodes + (size_t)ie * npv * nd;
        double *udg = UDG + (size_t)ie * npv * nc;
        const double *sh = SH + (size_t)ie * npv * nc;
        geometry(c, pn);
        qoperators(c);
        gather_uh(c, UH, elcon + (size_t)ie * ndf);
        for (int d = 0; d < nd; d++)
            for (int cc = 0; cc < ncu; cc++)
                for (int i = 0; i < npv; i++) {
                    double s = 0;
                    for (int j = 0; j < npv; j++) {
                        s += c->M[i + npv * j] * sh[(ncu + d * ncu + cc) * npv + j];
                        s += c->C[d * npv * npv + i + npv * j] * udg[cc * npv + j];
                    }
                    for (int l = 0; l < ndf; l++) s -= c->E[(size_t)d * npv * ndf + i + npv * l] * c->uh[cc * ndf + l];
                    rhs[(d * ncu + cc) * npv + i] = s;
                }
        memcpy(c->Mc, c->M, sizeof(double) * npv * npv);
        chol_u(npv, c->Mc);
        chol_solve(npv, c->Mc, ncu * nd, rhs, npv);
        memcpy(udg + (size_t)npv * ncu, rhs, sizeof(double) * npv * ncu * nd);
    }
    free(rhs);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* trace system                                                               */
/* ------------------------------------------------------------------------- */

/* mapElement2LinearSystem, HDG branch (assembly/globalassembly.cpp:42-86), with the
 * face-node map applied through its true inverse (identical to the reference
 * whenever the map is an involution). Hg blocks are column-major bsz x bsz. */
void hfo_scatter(int ne, int nfe, int npf, int nch, const int *t2f, const int *elcon,
                 const int *rowstart, const int *cols, const double *AH, const double *FH,
                 double *Hg, double *Rg)
{
    int bsz = nch * npf, ndf = npf * nfe, nb = ndf * nch;
    int ind[64], jnd[64];
    for (int ie = 0; ie < ne; ie++) {
        const double *He = AH + (size_t)ie * nb * nb, *Re = FH + (size_t)ie * nb;
        for (int a = 0; a < nfe; a++) {
            int i = t2f[ie * nfe + a];
            for (int m = 0; m < npf; m++) ind[m] = elcon[ie * ndf + a * npf + m] - i * npf;
            for (int m = 0; m < npf; m++)
                for (int q = 0; q < nch; q++) Rg[(size_t)i * bsz + ind[m] * nch + q] += Re[a * bsz + m * nch + q];
            for (int b = 0; b < nfe; b++) {
                int j = t2f[ie * nfe + b], k;
                for (k = rowstart[i]; k < rowstart[i + 1]; k++) if (cols[k] == j) break;
                for (int m = 0; m < npf; m++) jnd[m] = elcon[ie * ndf + b * npf + m] - j * npf;
                double *blk = Hg + (size_t)k * bsz * bsz;
                for (int m2 = 0; m2 < npf; m2++)          /* element column node on face b */
                    for (int n2 = 0; n2 < nch; n2++)
                        for (int m1 = 0; m1 < npf; m1++)  /* element row node on face a */
                            for (int n1 = 0; n1 < nch; n1++)
                                blk[(jnd[m2] * nch + n2) * bsz + ind[m1] * nch + n1] +=
                                    He[(size_t)(b * bsz + m2 * nch + n2) * nb + a * bsz + m1 * nch + n1];
            }
        }
    }
}

/* y = Hg x, rows = faces (kernel/matvec.m:8-12) */
void hfo_bsr_matvec(int nrow, int bsz, const int *rowstart, const int *cols, const double *Hg,
                    const double *x, double *y)
{
    for (int i = 0; i < nrow; i++) {
        double *yi = y + (size_t)i * bsz;
        for (int r = 0; r < bsz; r++) yi[r] = 0;
        for (int k = rowstart[i]; k < rowstart[i + 1]; k++) {
            const double *blk = Hg + (size_t)k * bsz * bsz, *xj = x + (size_t)cols[k] * bsz;
            for (int cc = 0; cc < bsz; cc++)
                for (int r = 0; r < bsz; r++) yi[r] += blk[cc * bsz + r] * xj[cc];
        }
    }
}
