/*
 * CPU FP64 oracle for the HDG element path (test infrastructure only).
 *
 * A plain-C restatement of the reference's element assembly, static
 * condensation, state recovery and trace-system kernels, used by the tests to
 * check the HIP library.  Nothing in the shipped library links or calls it.
 * Each function cites the reference code it restates.
 */
#ifndef HFEM_ORACLE_H
#define HFEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hfo_ctx hfo_ctx;

/* dims: nd, ncu, npv, ngv, npf, ngf, nfe (nch = ncu, nc = ncu*(nd+1)).
 * Master arrays in the reference's column-major layouts (masterstruct,
 * preprocessing/readbinaryfiles.cpp:95-223); perm is (npf, nfe), 0-based.
 * phys: param[6] = {gamma, eps_lm, Re, Pr, Minf, tau}, ui[nch] the far-field
 * state of boundary 0, then fc_u, fc_q, tdep. */
hfo_ctx *hfo_create(const int *dims, const double *shapvt, const double *shapvg,
                    const double *shapvgdotshapvl, const double *shapft, const double *shapfg,
                    const double *shapfgdotshapfc, const int *perm, const double *param,
                    const double *ui, double fc_u, double fc_q, int tdep);
void hfo_destroy(hfo_ctx *c);

/* assembleElementMatrixVector for elements [e0, e0+n): outputs per element
 * (contiguous, element-major): Ru[na], D[na*na], F[na*nb], K[nb*na], H[nb*nb], Rh[nb]. */
int hfo_assemble(hfo_ctx *c, int e0, int n, const double *dgnodes, const double *UDG,
                 const double *UH, const double *SH, const int *elcon, const int *bf,
                 double *Ru, double *D, double *F, double *K, double *H, double *Rh);

/* schurElementMatrixVector, in place: D -> factors, F -> inv(D)F, H -> H - K inv(D) F,
 * Ru -> inv(D)Ru, Rh -> Rh - K inv(D) Ru.  impl 0: LU with partial pivoting;
 * impl 1: block LU over ncu x ncu blocks of npv x npv with inverted diagonal blocks. */
int hfo_schur(int impl, int n, int npv, int ncu, int nb, double *D, double *F, double *K,
              double *H, double *Ru, double *Rh);

/* getQ for elements [e0,e0+n): recompute q in UDG from u and UH (in place). */
int hfo_getq(hfo_ctx *c, int e0, int n, const double *dgnodes, double *UDG, const double *UH,
             const double *SH, const int *elcon);

/* Flux and Jacobian at points (for tests): udg[np*nc] point-major columns like the
 * reference (udg[c*np + i]); f[np*ncu*nd], f_udg[np*ncu*nd*nc]. */
void hfo_flux(hfo_ctx *c, int np, const double *udg, double *f, double *f_udg);

/* Far-field boundary flux (application/FM/fbou.cpp ib==1) at np points. */
void hfo_fbou(hfo_ctx *c, int np, const double *udg, const double *uh, const double *nl,
              double *fb, double *fb_u, double *fb_uh);

/* Global trace system (serial restatements). */
void hfo_scatter(int ne, int nfe, int npf, int nch, const int *t2f, const int *elcon,
                 const int *rowstart, const int *cols, const double *AH, const double *FH,
                 double *Hg, double *Rg);
void hfo_bsr_matvec(int nrow, int bsz, const int *rowstart, const int *cols, const double *Hg,
                    const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
