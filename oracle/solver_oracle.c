/*
 * CPU FP64 oracle for the trace-system solver (test infrastructure only).
 *
 * Plain-C restatements of the reference's MATLAB specifications of the solver
 * pieces that its C++ binds from closed libraries (solvers/dgsolvers.cpp:16-20):
 *   approximate and exact MDF ordering   kernel/computeMDFordering.m:1-150,
 *                                        kernel/computeExactMDFordering.m:1-87
 *   BILU0 factorisation and application  kernel/computeBILU0.m:1-46, kernel/applyBILU0.m:1-52
 *   block-Jacobi                          kernel/mkBlockJacobi.m:1-16
 *   restarted left-preconditioned GMRES   kernel/gmres_bilu0.m:1-105
 * and of the solution update of solvers/solutionupdate.cpp:6-67 (getU).
 * BSR blocks are bsz x bsz column-major, rows = faces, the diagonal block first
 * in every row (preprocessing/elcon2entcon.m:54).  Indices are 0-based.
 * Nothing in the shipped library links or calls this file.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "solver_oracle.h"

/* ---- dense helpers (n x n column-major) ---- */

/* In-place inverse by LU with partial pivoting (LAPACK DGETRF + DGETRI semantics). */
int hfo_dense_inverse(int n, double *A)
{
    int *piv = (int *)malloc(sizeof(int) * n);
    double *W = (double *)malloc(sizeof(double) * n * n);
    int info = 0;
    memcpy(W, A, sizeof(double) * n * n);
    for (int k = 0; k < n; k++) {
        int p = k;
        double mx = fabs(W[k + n * k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(W[i + n * k]) > mx) { mx = fabs(W[i + n * k]); p = i; }
        piv[k] = p;
        if (mx == 0.0) { if (!info) info = k + 1; continue; }
        if (p != k)
            for (int j = 0; j < n; j++) { double t = W[k + n * j]; W[k + n * j] = W[p + n * j]; W[p + n * j] = t; }
        for (int i = k + 1; i < n; i++) W[i + n * k] /= W[k + n * k];
        for (int j = k + 1; j < n; j++)
            for (int i = k + 1; i < n; i++) W[i + n * j] -= W[i + n * k] * W[k + n * j];
    }
    /* solve for the identity, column by column */
    for (int c = 0; c < n; c++) {
        double *x = A + (size_t)n * c;
        for (int i = 0; i < n; i++) x[i] = (i == c);
        for (int k = 0; k < n; k++)
            if (piv[k] != k) { double t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
        for (int k = 0; k < n; k++)
            for (int i = k + 1; i < n; i++) x[i] -= W[i + n * k] * x[k];
        for (int k = n - 1; k >= 0; k--) {
            x[k] /= W[k + n * k];
            for (int i = 0; i < k; i++) x[i] -= W[i + n * k] * x[k];
        }
    }
    free(piv);
    free(W);
    return info;
}

/* C = A B */
static void mm(int n, const double *A, const double *B, double *C)
{
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += A[i + n * k] * B[k + n * j];
            C[i + n * j] = s;
        }
}

/* C -= A B */
static void mm_sub(int n, const double *A, const double *B, double *C)
{
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += A[i + n * k] * B[k + n * j];
            C[i + n * j] -= s;
        }
}

/* y = A x (acc = 0) or y -= A x (acc = 1) */
static void mv(int n, const double *A, const double *x, double *y, int sub)
{
    for (int i = 0; i < n; i++) {
        double s = 0;
        for (int k = 0; k < n; k++) s += A[i + n * k] * x[k];
        y[i] = sub ? y[i] - s : s;
    }
}

static double frob(int n, const double *A)
{
    double s = 0;
    for (int i = 0; i < n * n; i++) s += A[i] * A[i];
    return sqrt(s);
}

static int find_in_row(const int *rowpts, const int *colind, int r, int c)
{
    for (int k = rowpts[r]; k < rowpts[r + 1]; k++)
        if (colind[k] == c) return k;
    return -1;
}

/* ---- MDF ordering (kernel/computeMDFordering.m, kernel/computeExactMDFordering.m) ---- */

/* Weight of row k: Frobenius norm of the discarded fill DeltaC(i,j) over pairs of
 * distinct unnumbered neighbours ri, rj of k that are not neighbours of each other
 * (:23-45 initially, :57-85 after each pivot).  approximate: C(ik) C(kj) with
 * C(n) = ||inv(H_ii) H_n||_F of the block's row; exact: ||H_ik inv(H_kk) H_kj||_F. */
static double mdf_weight(int k, int bsz, const int *rowpts, const int *colind, const double *Hg,
                         const double *Cw, const double *Dinv, const double *w, double big, int exact,
                         double *T1, double *T2)
{
    const size_t b2 = (size_t)bsz * bsz;
    double s = 0;
    for (int i = rowpts[k] + 1; i < rowpts[k + 1]; i++) {
        int ri = colind[i];
        if (w && !(w[ri] < big)) continue;
        int ik = find_in_row(rowpts, colind, ri, k);
        for (int j = rowpts[k] + 1; j < rowpts[k + 1]; j++) {
            if (j == i) continue;
            int rj = colind[j];
            if (w && !(w[rj] < big)) continue;
            if (find_in_row(rowpts, colind, ri, rj) >= 0) continue;
            double dc;
            if (!exact) dc = Cw[ik] * Cw[j];
            else {
                /* H_ik * (H_kk \ H_kj) */
                mm(bsz, Dinv + (size_t)k * b2, Hg + (size_t)j * b2, T1);
                mm(bsz, Hg + (size_t)ik * b2, T1, T2);
                dc = frob(bsz, T2);
            }
            s += dc * dc;
        }
    }
    return sqrt(s);
}

int hfo_mdf_order(int nrows, int bsz, const int *rowpts, const int *colind, const double *Hg, int exact,
                  int *o2u, int *u2o)
{
    const size_t b2 = (size_t)bsz * bsz;
    const int nblk = rowpts[nrows];
    double *Dinv = (double *)malloc(sizeof(double) * b2 * nrows);
    double *Cw = (double *)calloc(nblk, sizeof(double));
    double *w = (double *)malloc(sizeof(double) * nrows);
    double *T1 = (double *)malloc(sizeof(double) * b2), *T2 = (double *)malloc(sizeof(double) * b2);
    int info = 0;
    for (int i = 0; i < nrows; i++) {
        memcpy(Dinv + (size_t)i * b2, Hg + (size_t)rowpts[i] * b2, sizeof(double) * b2);
        if (hfo_dense_inverse(bsz, Dinv + (size_t)i * b2)) info = -1;
    }
    /* C(n) = ||H_ii \ H_n||_F (:14-20) */
    for (int i = 0; i < nrows; i++)
        for (int n = rowpts[i]; n < rowpts[i + 1]; n++) {
            mm(bsz, Dinv + (size_t)i * b2, Hg + (size_t)n * b2, T1);
            Cw[n] = frob(bsz, T1);
        }
    for (int k = 0; k < nrows; k++) w[k] = mdf_weight(k, bsz, rowpts, colind, Hg, Cw, Dinv, NULL, 0, exact, T1, T2);
    const double big = 1e100;
    for (int l = 0; l < nrows; l++) {
        /* [~, p(l)] = min(w): the first index of the minimum (:93) */
        int p = 0;
        for (int k = 1; k < nrows; k++)
            if (w[k] < w[p]) p = k;
        o2u[l] = p;
        w[p] = big;
        for (int e = rowpts[p] + 1; e < rowpts[p + 1]; e++) {
            int k = colind[e];
            if (w[k] < big) w[k] = mdf_weight(k, bsz, rowpts, colind, Hg, Cw, Dinv, w, big, exact, T1, T2);
        }
    }
    for (int l = 0; l < nrows; l++) u2o[o2u[l]] = l;
    free(Dinv); free(Cw); free(w); free(T1); free(T2);
    return info;
}

/* ---- BILU0 (kernel/computeBILU0.m, kernel/applyBILU0.m) ---- */

int hfo_bilu0_factor(int nrows, int bsz, const int *rowpts, const int *colind, const int *o2u, const int *u2o,
                     double *Mg)
{
    const size_t b2 = (size_t)bsz * bsz;
    int info = 0;
    double *T = (double *)malloc(sizeof(double) * b2);
    for (int j = 0; j < nrows; j++) {
        const int rj = o2u[j], jj = rowpts[rj];
        if (hfo_dense_inverse(bsz, Mg + (size_t)jj * b2) && !info) info = rj + 1;   /* inv(Ujj) (:12) */
        for (int i = rowpts[rj] + 1; i < rowpts[rj + 1]; i++) {
            const int ri = colind[i];
            if (u2o[ri] <= j) continue;
            const int ij = find_in_row(rowpts, colind, ri, rj), ii = rowpts[ri];
            /* Lij = Mij inv(Ujj) (:21) */
            mm(bsz, Mg + (size_t)ij * b2, Mg + (size_t)jj * b2, T);
            memcpy(Mg + (size_t)ij * b2, T, sizeof(double) * b2);
            /* Uii -= Lij Uji (:25) */
            mm_sub(bsz, Mg + (size_t)ij * b2, Mg + (size_t)i * b2, Mg + (size_t)ii * b2);
            /* U_il -= L_ij U_jm for later common neighbours (:26-40) */
            for (int il = rowpts[ri] + 1; il < rowpts[ri + 1]; il++) {
                const int rij = colind[il];
                if (u2o[rij] <= j) continue;
                for (int jm = rowpts[rj] + 1; jm < rowpts[rj + 1]; jm++)
                    if (colind[jm] == rij) mm_sub(bsz, Mg + (size_t)ij * b2, Mg + (size_t)jm * b2, Mg + (size_t)il * b2);
            }
        }
    }
    free(T);
    return info;
}

void hfo_bilu0_apply(int nrows, int bsz, const int *rowpts, const int *colind, const int *o2u, const int *u2o,
                     const double *Mg, const double *x, double *v)
{
    const size_t b2 = (size_t)bsz * bsz;
    double *y = (double *)malloc(sizeof(double) * (size_t)nrows * bsz);
    double *t = (double *)malloc(sizeof(double) * bsz);
    /* L y = x (:12-27) */
    for (int j = 0; j < nrows; j++) {
        const int rj = o2u[j];
        double *yj = y + (size_t)rj * bsz;
        memcpy(yj, x + (size_t)rj * bsz, sizeof(double) * bsz);
        for (int k = rowpts[rj] + 1; k < rowpts[rj + 1]; k++)
            if (u2o[colind[k]] < j) mv(bsz, Mg + (size_t)k * b2, y + (size_t)colind[k] * bsz, yj, 1);
    }
    /* U v = y, inverted diagonal (:29-48) */
    for (int j = nrows - 1; j >= 0; j--) {
        const int rj = o2u[j];
        memcpy(t, y + (size_t)rj * bsz, sizeof(double) * bsz);
        for (int k = rowpts[rj] + 1; k < rowpts[rj + 1]; k++)
            if (u2o[colind[k]] > j) mv(bsz, Mg + (size_t)k * b2, v + (size_t)colind[k] * bsz, t, 1);
        mv(bsz, Mg + (size_t)rowpts[rj] * b2, t, v + (size_t)rj * bsz, 0);
    }
    free(y);
    free(t);
}

/* Block-Jacobi preconditioner: the inverted diagonal blocks (kernel/mkBlockJacobi.m:9-11). */
int hfo_block_jacobi_factor(int nrows, int bsz, const int *rowpts, const double *Hg, double *Dinv)
{
    const size_t b2 = (size_t)bsz * bsz;
    int info = 0;
    for (int i = 0; i < nrows; i++) {
        memcpy(Dinv + (size_t)i * b2, Hg + (size_t)rowpts[i] * b2, sizeof(double) * b2);
        if (hfo_dense_inverse(bsz, Dinv + (size_t)i * b2) && !info) info = i + 1;
    }
    return info;
}

/* ---- GMRES (kernel/gmres_bilu0.m) ---- */

typedef struct {
    int nrows, bsz, prec;
    const int *rowpts, *colind, *o2u, *u2o;
    const double *Hg, *Mg;
} sysops;

static void matvec(const sysops *s, const double *x, double *y)
{
    const size_t b2 = (size_t)s->bsz * s->bsz;
    for (int i = 0; i < s->nrows; i++) {
        double *yi = y + (size_t)i * s->bsz;
        for (int r = 0; r < s->bsz; r++) yi[r] = 0;
        for (int k = s->rowpts[i]; k < s->rowpts[i + 1]; k++) {
            const double *blk = s->Hg + (size_t)k * b2, *xj = x + (size_t)s->colind[k] * s->bsz;
            for (int c = 0; c < s->bsz; c++)
                for (int r = 0; r < s->bsz; r++) yi[r] += blk[c * s->bsz + r] * xj[c];
        }
    }
}

static void precond(const sysops *s, const double *x, double *y)
{
    const size_t N = (size_t)s->nrows * s->bsz;
    if (s->prec == 0) hfo_bilu0_apply(s->nrows, s->bsz, s->rowpts, s->colind, s->o2u, s->u2o, s->Mg, x, y);
    else if (s->prec == 2)
        for (int i = 0; i < s->nrows; i++)
            mv(s->bsz, s->Mg + (size_t)i * s->bsz * s->bsz, x + (size_t)i * s->bsz, y + (size_t)i * s->bsz, 0);
    else memcpy(y, x, sizeof(double) * N);
}

static double dot(size_t n, const double *a, const double *b)
{
    double s = 0;
    for (size_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* Least squares min |beta e1 - Hbar y| by Givens rotations, Hbar (j+1) x j stored
 * column-major with leading dimension ldh.  Returns the residual norm. */
static double hessenberg_ls(int j, int ldh, const double *hh, double beta, double *y)
{
    double *R = (double *)malloc(sizeof(double) * ldh * j), *g = (double *)calloc(j + 1, sizeof(double));
    memcpy(R, hh, sizeof(double) * ldh * j);
    g[0] = beta;
    for (int k = 0; k < j; k++) {
        double a = R[k + ldh * k], b = R[k + 1 + ldh * k], r = hypot(a, b);
        double c = r == 0 ? 1 : a / r, s = r == 0 ? 0 : b / r;
        for (int col = k; col < j; col++) {
            double u = R[k + ldh * col], v = R[k + 1 + ldh * col];
            R[k + ldh * col] = c * u + s * v;
            R[k + 1 + ldh * col] = -s * u + c * v;
        }
        double u = g[k], v = g[k + 1];
        g[k] = c * u + s * v;
        g[k + 1] = -s * u + c * v;
    }
    for (int k = j - 1; k >= 0; k--) {
        double s = g[k];
        for (int c2 = k + 1; c2 < j; c2++) s -= R[k + ldh * c2] * y[c2];
        y[k] = s / R[k + ldh * k];
    }
    double res = fabs(g[j]);
    free(R);
    free(g);
    return res;
}

int hfo_gmres(int nrows, int bsz, const int *rowpts, const int *colind, const double *Hg, int prec,
              const double *Mg, const int *o2u, const int *u2o, const double *b_in, double *x, int restart,
              double tol, int maxit, int orthog, int *iters, double *rev)
{
    sysops s = {nrows, bsz, prec, rowpts, colind, o2u, u2o, Hg, Mg};
    const size_t N = (size_t)nrows * bsz;
    double *b = (double *)malloc(sizeof(double) * N), *r = (double *)malloc(sizeof(double) * N);
    double *t = (double *)malloc(sizeof(double) * N);
    double *V = (double *)malloc(sizeof(double) * N * (restart + 1));
    double *hh = (double *)calloc((size_t)(restart + 1) * restart, sizeof(double));
    double *y = (double *)malloc(sizeof(double) * restart), *hc = (double *)malloc(sizeof(double) * (restart + 1));
    const int ldh = restart + 1;
    precond(&s, b_in, b);                      /* b = M \ b (:17) */
    const double nrmb = sqrt(dot(N, b, b));
    int flag = -1, iter = 0;
    double res = 0;
    while (1) {
        matvec(&s, x, t);                      /* r = b - M \ (A x) (:33-39) */
        precond(&s, t, r);
        for (size_t i = 0; i < N; i++) r[i] = b[i] - r[i];
        const double beta = sqrt(dot(N, r, r));
        if (beta == 0 || nrmb == 0) { flag = 0; res = 0; break; }
        for (size_t i = 0; i < N; i++) V[i] = r[i] / beta;
        memset(hh, 0, sizeof(double) * ldh * restart);
        int jdone = 0;
        for (int j = 0; j < restart; j++) {
            double *vj = V + N * j, *w = V + N * (j + 1);
            matvec(&s, vj, t);
            precond(&s, t, w);
            if (orthog == 0) {                 /* modified Gram-Schmidt (:56-60) */
                for (int i = 0; i <= j; i++) {
                    double h = dot(N, V + N * i, w);
                    hh[i + ldh * j] = h;
                    for (size_t q = 0; q < N; q++) w[q] -= h * V[N * i + q];
                }
            } else {                           /* classical Gram-Schmidt, applied twice (ICGS) */
                for (int pass = 0; pass < 2; pass++) {
                    for (int i = 0; i <= j; i++) hc[i] = dot(N, V + N * i, w);
                    for (int i = 0; i <= j; i++) {
                        hh[i + ldh * j] += hc[i];
                        for (size_t q = 0; q < N; q++) w[q] -= hc[i] * V[N * i + q];
                    }
                }
            }
            const double hn = sqrt(dot(N, w, w));
            hh[j + 1 + ldh * j] = hn;
            jdone = j + 1;
            if (hn != 0.0)
                for (size_t q = 0; q < N; q++) w[q] /= hn;
            res = hessenberg_ls(j + 1, ldh, hh, beta, y);   /* (:70-73) */
            iter++;
            if (rev) rev[iter - 1] = res / nrmb;
            if (hn == 0.0) { flag = 0; break; }             /* breakdown: exact solution in the space */
            if (res / nrmb <= tol) { flag = 0; break; }
            if (iter >= maxit) { flag = 1; break; }
        }
        for (int k = 0; k < jdone; k++)                      /* x += V y (:93) */
            for (size_t q = 0; q < N; q++) x[q] += V[N * k + q] * y[k];
        if (flag >= 0) break;
    }
    *iters = iter;
    free(b); free(r); free(t); free(V); free(hh); free(y); free(hc);
    return flag;
}

/* ---- getU (solvers/solutionupdate.cpp:6-67) ---- */

/* UDG u-part: u -= alpha inv(D)F duh_e, u += alpha inv(D)Ru;  UH += alpha duh. */
void hfo_update_u(int ne, int npv, int nc, int ncu, int nch, int ndf, const int *elcon, const double *DinvF,
                  const double *DinvRu, const double *duh, double alpha, double *UDG)
{
    const int na = npv * ncu, nb = ndf * nch;
    double *de = (double *)malloc(sizeof(double) * nb);
    for (int e = 0; e < ne; e++) {
        for (int j = 0; j < ndf; j++)
            for (int k = 0; k < nch; k++) de[j * nch + k] = duh[(size_t)elcon[e * ndf + j] * nch + k];
        double *u = UDG + (size_t)e * npv * nc;
        const double *F = DinvF + (size_t)e * na * nb, *R = DinvRu + (size_t)e * na;
        for (int i = 0; i < na; i++) {
            double s = 0;
            for (int c = 0; c < nb; c++) s += F[i + (size_t)na * c] * de[c];
            u[i] = u[i] - alpha * s + alpha * R[i];
        }
    }
    free(de);
}
