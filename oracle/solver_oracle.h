/*
 * CPU FP64 oracle for the trace-system solver (test infrastructure only):
 * MDF ordering, BILU0, block-Jacobi, GMRES and the solution update, restated
 * from the reference's MATLAB specifications (the kernel directory) and
 * solvers/solutionupdate.cpp.  BSR blocks are column-major bsz x bsz with the
 * diagonal block first in each row; all indices 0-based.
 */
#ifndef HFEM_SOLVER_ORACLE_H
#define HFEM_SOLVER_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

int hfo_dense_inverse(int n, double *A);
/* o2u[l] = row pivoted at step l, u2o its inverse; exact = 0 approximate MDF. */
int hfo_mdf_order(int nrows, int bsz, const int *rowpts, const int *colind, const double *Hg, int exact,
                  int *o2u, int *u2o);
/* In place on a copy of Hg: diagonal blocks become inv(U_ii), off-diagonal blocks L or U. */
int hfo_bilu0_factor(int nrows, int bsz, const int *rowpts, const int *colind, const int *o2u, const int *u2o,
                     double *Mg);
void hfo_bilu0_apply(int nrows, int bsz, const int *rowpts, const int *colind, const int *o2u, const int *u2o,
                     const double *Mg, const double *x, double *v);
int hfo_block_jacobi_factor(int nrows, int bsz, const int *rowpts, const double *Hg, double *Dinv);
/* prec: -1 none, 0 BILU0 (Mg from hfo_bilu0_factor), 2 block-Jacobi (Mg = nrows inverted
 * diagonal blocks).  orthog: 0 MGS, 1 ICGS.  Returns 0 converged, 1 max iterations. */
int hfo_gmres(int nrows, int bsz, const int *rowpts, const int *colind, const double *Hg, int prec,
              const double *Mg, const int *o2u, const int *u2o, const double *b, double *x, int restart,
              double tol, int maxit, int orthog, int *iters, double *rev);
void hfo_update_u(int ne, int npv, int nc, int ncu, int nch, int ndf, const int *elcon, const double *DinvF,
                  const double *DinvRu, const double *duh, double alpha, double *UDG);

#ifdef __cplusplus
}
#endif
#endif
