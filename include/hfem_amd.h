/*
 * hfem_amd: MI355X (gfx950) HDG element path -- C ABI.
 *
 * Drop-in boundary for the reference's per-element hot path.  The reference
 * binds it at link time from C++ with struct arguments; each entry point below
 * replaces the reference function named beside it, batched over elements and
 * with the element data resident in device memory:
 *
 *   hfem_assemble        assembleElementMatrixVector  assembly/elementassembly.cpp:1481-1499
 *   hfem_assemble_vector assembleElementVector        assembly/elementassembly.cpp:1501-1512
 *   hfem_condense        schurElementMatrixVector     assembly/elementschur.cpp:331-343
 *   hfem_assemble_linear_system
 *                        assembleLinearSystem         assembly/globalassembly.cpp:423-556
 *   hfem_bsr_matvec      matrix-vector product (closed in C++; spec kernel/matvec.m:1-12)
 *
 * Conventions: every function returns 0 on success and a negative code on
 * error (hfem_last_error() gives the message) instead of the reference's
 * error()/exit(1) (preprocessing/errormsg.cpp:4-21).  Host arrays use the
 * reference's layouts (digasopre.h structs; flat, column-major, 0-based
 * int32 indices like its Int).  One context owns one device and one stream; a
 * context is not thread-safe.
 */
#ifndef HFEM_AMD_H
#define HFEM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hfem_ctx hfem_ctx;

/* Element family and problem sizes.  nch = ncu (compressible NS), nc = ncu*(nd+1).
 * Mirrors master.ndims / mesh sizes read by readmasterstruct and readmeshstruct
 * (preprocessing/readbinaryfiles.cpp:95-223, :311-460). */
typedef struct {
    int nd;   /* spatial dimension (2 or 3) */
    int ncu;  /* solution components (nd+2) */
    int npv;  /* nodes per element */
    int ngv;  /* volume quadrature points */
    int npf;  /* nodes per face */
    int ngf;  /* face quadrature points */
    int nfe;  /* faces per element */
    int ne;   /* elements */
    int nf;   /* faces (trace entities) */
} hfem_dims;

/* Condensation precision. */
enum {
    HFEM_FP64 = 0,        /* FP64 LU with partial pivoting (schurImplementation 0 semantics) */
    HFEM_MIXED = 1        /* FP32 factorisation with FP64 iterative refinement */
};

const char *hfem_last_error(void);

/* Create a context on HIP device `device`, allocating the element block storage
 * for dims->ne elements (D, F, K, H, Ru, Rh and the condensed inv(D)F, inv(D)Ru). */
int hfem_create(hfem_ctx **ctx, const hfem_dims *dims, int device);
int hfem_destroy(hfem_ctx *ctx);

/* Master tables in the masterstruct layouts: shapvt (ngv,npv,nd+1), shapvg
 * (npv,ngv,nd+1), shapft (ngf,npf,nd), shapfg (npf,ngf,nd), perm (npf,nfe). */
int hfem_set_master(hfem_ctx *ctx, const double *shapvt, const double *shapvg, const double *shapft,
                    const double *shapfg, const int *perm);

/* physicsparam = {gamma, eps_lm, Re, Pr, Minf, tau} (preprocessing/digasopre.m:513),
 * ui = far-field state of boundary 0 (app.bcs), time-discretisation factors. */
int hfem_set_physics(hfem_ctx *ctx, const double *physicsparam, const double *ui, double fc_u,
                     double fc_q, int tdep);

/* Mesh: dgnodes (ne,nd,npv), elcon (ne,nfe,npf), bf (ne,nfe; <0 boundary),
 * t2f (ne,nfe); BSR trace graph rowstart[nf+1], cols[nblocks] (self block first). */
int hfem_set_mesh(hfem_ctx *ctx, const double *dgnodes, const int *elcon, const int *bf, const int *t2f,
                  const int *rowstart, const int *cols, int nblocks);

/* State: UDG (ne,nc,npv), UH (nf*npf,nch), SH (ne,nc,npv); host <-> device. */
int hfem_set_solution(hfem_ctx *ctx, const double *UDG, const double *UH, const double *SH);
int hfem_get_solution(hfem_ctx *ctx, double *UDG, double *UH);

/* Element assembly for elements [e0, e0+n) into the device element blocks. */
int hfem_assemble(hfem_ctx *ctx, int e0, int n);
/* assembleElementVector (assembly/elementassembly.cpp:1501-1512): the residual rows
 * Ru, Rh of elements [e0, e0+n) only (fluxes without derivatives, no matrix
 * phases; the reference still forms the Jacobians there).  D, F, K, H are left as
 * they are.  Used by hfem_residual_norm and hfem_assemble_rhs. */
int hfem_assemble_vector(hfem_ctx *ctx, int e0, int n);
/* Static condensation of elements [e0, e0+n) on the device blocks: inv(D)F,
 * inv(D)Ru, AH (in place of H) and FH; D, Ru and Rh keep their assembled values. */
int hfem_condense(hfem_ctx *ctx, int e0, int n, int precision);
/* Mixed path (HFEM_MIXED): inv(D) by Gauss-Jordan on the FP32 MFMA, X0 = inv32 [F Ru]
 * on the FP32 MFMA, then sweeps R = [F Ru] - D X (FP64) and X += inv32 R until the
 * estimated remaining error of every element is below tol max|X| (at most
 * max_sweeps; defaults 1e-11 and 4), then the FP64 Schur product.  A range with
 * an element that does not converge is condensed again by the FP64 path.  No
 * counterpart in the reference (its FP32 copies are storage only,
 * assembly/globalassembly.cpp:353-371). */
int hfem_set_refinement(hfem_ctx *ctx, double tol, int max_sweeps);
/* Report of the last hfem_condense: refinement sweeps run, elements left above the
 * tolerance, 1 if the range was redone in FP64 (all 0 after an FP64 condensation). */
int hfem_condense_report(hfem_ctx *ctx, int *sweeps, int *unconverged, int *fallback);
/* No counterpart in the reference (it calls LAPACK/BLAS per element): times the
 * launch candidates of every stage of hfem_condense on a sample of nsample
 * elements (0: min(ne, 32768)) and keeps the fastest in the context -- the
 * inverse kernels' launch shapes (FP64, FP32) and, per product (0 X = inv(D)[F Ru],
 * 1 [H Rh] - K X, 2 mixed X0, 3 mixed residual [F Ru] - D X, 4 mixed correction
 * X += inv32 R), the hand-written MFMA variants that fit plus, for the FP64
 * products 0, 1 and 3, rocBLAS (its own choice and each solution it lists) as the
 * measured comparison.  Element blocks are not changed.  choice[7] (may be NULL)
 * receives {inverse FP64, inverse FP32, product 0..4} (product: variant >= 0, -1
 * rocBLAS), ms[7] (may be NULL) the best time of each in milliseconds. */
int hfem_tune_condensation(hfem_ctx *ctx, int nsample, int *choice, float *ms);
/* Set the seven launch choices in hfem_tune_condensation's layout (a pick kept from
 * an earlier tuning); an error if a choice cannot run this context's shapes. */
int hfem_set_condensation_choice(hfem_ctx *ctx, const int *choice);
/* Number of hand-written product variants (valid product choices 0 .. n-1). */
int hfem_condense_variants(void);

/* Ragged element batches.  No counterpart in the reference, which runs one
 * element family per process (main/digasoser.cpp:62-63): the elements of up to
 * six contexts of different 3D families (tets of orders 1-4, hexes of orders
 * 1-2, one context per family, one device, at most one flux chunk of 32768
 * elements each) go through single persistent launches with ragged per-bucket
 * offsets.  Each element gets the per-element work of hfem_assemble /
 * hfem_condense (FP64), and each context's buffers receive exactly what those
 * entry points write.  The contexts stay usable on their own; batch calls are
 * ordered after the work already queued on the contexts' streams and before
 * work queued on them later. */
typedef struct hfem_batch hfem_batch;
/* nelem[nctx] (may be NULL: every element): the batch takes elements
 * [0, nelem[k]) of ctxs[k].  Destroy the batch before its contexts. */
int hfem_batch_create(hfem_ctx *const *ctxs, int nctx, const int *nelem, hfem_batch **batch);
int hfem_batch_destroy(hfem_batch *batch);
/* assembleElementMatrixVector (assembly/elementassembly.cpp:1481-1499) over every
 * element of every context: one flux launch and one assembly launch. */
int hfem_batch_assemble(hfem_batch *batch);
/* schurElementMatrixVector (assembly/elementschur.cpp:331-343) over every element.
 * HFEM_FP64: one inverse launch, then each context's two products with its own
 * tuned shapes on its own stream.  HFEM_MIXED: each context's mixed condensation
 * (hfem_condense(.., HFEM_MIXED): FP32-MFMA inverse, FP64 refinement to its
 * tolerance, FP64 Schur product) on its own stream. */
int hfem_batch_condense(hfem_batch *batch, int precision);
/* Per bucket (context): runs hfem_tune_condensation on the context, times its
 * assembly and the inverse with 8, 4 or 2 waves per element (those its tile
 * count allows) in the batch launch, keeps the fastest and orders the buckets by
 * measured cost per element, most expensive first; then picks the launch forms
 * (hfem_batch_set_launch).  waves[nctx] (may be NULL)
 * receives the inverse's waves per element, ms[2*nctx] (may be NULL) the
 * assembly and inverse milliseconds of each bucket alone.  Assembles (the
 * element blocks are overwritten). */
int hfem_batch_tune(hfem_batch *batch, int *waves, float *ms);
/* Waves per element of the batch inverse for each context (8, 4 or 2). */
int hfem_batch_set_inverse_waves(hfem_batch *batch, const int *waves);
/* Launch form of the batch assembly (flux and assembly kernels) and of the batch
 * inverse: 0 one persistent launch over all buckets (the default), 1 one launch
 * per bucket with the per-family kernels, on the batch stream.  hfem_batch_tune
 * times both over the whole batch and keeps the faster. */
int hfem_batch_set_launch(hfem_batch *batch, int assemble_mode, int inverse_mode);
/* modes[2] (may be NULL): the current launch forms; tuned_ms[4] (may be NULL): the
 * last tuning's times of the assembly in one launch / per bucket and of the
 * inverse in one launch / per bucket. */
int hfem_batch_get_launch(hfem_batch *batch, int *modes, float *tuned_ms);
/* Milliseconds of the last hfem_batch_assemble and hfem_batch_condense (HIP events). */
int hfem_batch_kernel_ms(hfem_batch *batch, float *assemble_ms, float *condense_ms);
/* Synchronise the batch stream and the contexts' streams. */
int hfem_batch_sync(hfem_batch *batch);

/* No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * The FP64 condensation of hfem_condense on caller-owned device blocks of n
 * elements in the library's internal layouts -- D (na x na column-major), [F Ru]
 * (na x ldx row-major, Ru in column nb), K (nb x na column-major), [H Rh]
 * (nb x (nb+1) column-major, Rh in column nb) -- writing X = [inv(D)F inv(D)Ru]
 * (na x ldx row-major) and [AH FH] over [H Rh], with info[n] the singular-pivot
 * flags, on `stream` (a hipStream_t; NULL the default stream).  It allocates
 * nothing: inv(D) lives in a fixed 32 MiB device array per GPU, so n*na*na must
 * not exceed 4 Mi doubles (419 elements of na = 100; larger batches are refused
 * with an error -- call it per range).  Calls that overlap on one GPU share that
 * array and race. */
int hfem_condense_noalloc(int na, int nb, int ldx, int n, const double *D, const double *Fx, const double *K,
                          double *Hx, double *X, int *info, void *stream);

/* Copy the assembled element blocks of [e0, e0+n) to the host (any pointer may
 * be NULL), each element's block contiguous and column-major like elemstruct
 * (Ru na, D na*na, F na*nb, K nb*na, H nb*nb, Rh nb).  H is overwritten by
 * hfem_condense; read it before condensing. */
int hfem_read_elements(hfem_ctx *ctx, int e0, int n, double *Ru, double *D, double *F, double *K, double *H,
                       double *Rh);
/* Copy the condensation outputs of [e0, e0+n): inv(D)Ru (na), inv(D)F (na*nb),
 * AH (nb*nb), FH (nb) -- sol.DinvRu/sol.DinvF and elem.H/elem.Rh after
 * schurElementMatrixVector -- and a per-element singular-pivot flag (0 = ok,
 * else 1 + first zero pivot). */
int hfem_read_condensed(hfem_ctx *ctx, int e0, int n, double *DinvRu, double *DinvF, double *AH, double *FH,
                        int *info);

/* assembleLinearSystem: assemble + condense all elements, scatter AH/FH into the
 * BSR matrix Hg and right-hand side Rg (device), and return
 * rNorm = sqrt(sum_e |Ru_e|^2 + |r|^2) with r the assembled uncondensed Rh.
 * With quasi-Newton storage enabled it also keeps inv(D) and K of every element
 * (updateQuasiNewtonStructures, assembly/globalassembly.cpp:327-389). */
int hfem_assemble_linear_system(hfem_ctx *ctx, int precision, double *rNorm);
/* Quasi-Newton storage (app.quasiNewton, sys.quasiNewtonAccuracy,
 * preprocessing/initializestructs.cpp:193,223-230): accuracy -1 off, 0 inv(D) and K
 * kept in float (DinvFloat, Kfloat), 1 kept in double. */
int hfem_quasi_newton_storage(hfem_ctx *ctx, int accuracy);
/* assembleRHS (assembly/globalassembly.cpp:559-616): Rg, r and rNorm at the current
 * state with the stored inv(D) and K of the last hfem_assemble_linear_system
 * (schurElementVector, assembly/elementschur.cpp:345-458); Hg is left as it is.
 * inv(D) Ru is updated for hfem_update_solution; inv(D) F stays the stored one. */
int hfem_assemble_rhs(hfem_ctx *ctx, double *rNorm);
/* Copy Hg (nblocks*bsz*bsz, column-major blocks) and Rg (nf*bsz) to the host. */
int hfem_read_system(hfem_ctx *ctx, double *Hg, double *Rg);
/* y = Hg x with host vectors of length nf*bsz. */
int hfem_bsr_matvec(hfem_ctx *ctx, const double *x, double *y);

/* ---- trace-system solver --------------------------------------------------
 * The reference binds these from closed libraries (solvers/dgsolvers.cpp:16-20:
 * computeOrdering, computePreconditioner, gmres); their semantics follow its
 * MATLAB specifications.  They act on the Hg, Rg of the last
 * hfem_assemble_linear_system. */

/* computeOrdering: method 0 natural, 1 approximate minimum discarded fill
 * (kernel/computeMDFordering.m:1-150; reorderMethod 1). */
int hfem_compute_ordering(hfem_ctx *ctx, int method);
/* Use the caller's ordering: o2u[l] = face pivoted at step l. */
int hfem_set_ordering(hfem_ctx *ctx, const int *o2u);
/* Current ordering and the number of levels of the forward / backward sweeps
 * (nlevels[2]); either pointer may be NULL. */
int hfem_get_ordering(hfem_ctx *ctx, int *o2u, int *nlevels);
/* computePreconditioner: kind -1 none, 0 BILU0 in the current ordering
 * (kernel/computeBILU0.m:1-46), 2 block-Jacobi (kernel/mkBlockJacobi.m:1-16). */
int hfem_precond_factor(hfem_ctx *ctx, int kind);
/* sys.precPrecision (preprocessing/initializestructs.cpp:476-486): 0 keeps the
 * preconditioner factors in single precision (Mg_sp; BILU0 factored on the FP32
 * MFMA, sweeps reading FP32 factors with FP64 arithmetic), 1 in double (default).
 * Takes effect at the next hfem_precond_factor. */
int hfem_set_precond_precision(hfem_ctx *ctx, int prec_precision);
/* No counterpart in the reference: 1 replays the BILU0 application (the
 * level-by-level launches of both sweeps) as a HIP graph captured once per
 * factorisation; 0 (default) launches the levels one by one.  Measured equal on
 * MI355X: a sweep is bound by the latency of its chain of dependent levels, not
 * by launch overhead. */
int hfem_set_apply_graph(hfem_ctx *ctx, int use);
/* No counterpart in the reference (kernel/applyBILU0.m runs its sweeps row by
 * row): mode 1 (default) runs each BILU0 sweep as one persistent launch whose
 * workgroups take the rows in level order and wait on per-row ready flags, each
 * wait bounded by spin_limit polls (0: 2^20); mode 0 launches the sweeps level by
 * level.  Same results bit for bit.  A wait that gives up makes the next
 * hfem_gmres / hfem_precond_apply return an error. */
int hfem_set_bilu0_sweep(hfem_ctx *ctx, int mode, int spin_limit);
/* y = M^{-1} x with host vectors of length nf*bsz (kernel/applyBILU0.m:1-52). */
int hfem_precond_apply(hfem_ctx *ctx, const double *x, double *y);
/* gmres: left-preconditioned restarted GMRES on Hg x = Rg from x = 0
 * (kernel/gmres_bilu0.m:1-105); orthog 0 modified, 1 iterated classical
 * Gram-Schmidt (orthogMethod).  rev (maxit doubles, may be NULL) receives the
 * relative preconditioned residual of every iteration; flag 0 converged,
 * 1 iteration limit.  x stays on the device (hfem_read_trace_solution). */
int hfem_gmres(hfem_ctx *ctx, int restart, double tol, int maxit, int orthog, int *iters, double *rev, int *flag);
int hfem_read_trace_solution(hfem_ctx *ctx, double *x);

/* updateSol (solvers/solutionupdate.cpp:69-98): UH += alpha x, getU on every
 * element with the inv(D)F, inv(D)Ru of the last condensation, then getQ. */
int hfem_update_solution(hfem_ctx *ctx, double alpha);
/* getQ on every element (assembly/elementassembly.cpp:1414-1479). */
int hfem_get_q(hfem_ctx *ctx);
/* computeResidualNorm (assembly/globalassembly.cpp:618-678). */
int hfem_residual_norm(hfem_ctx *ctx, double *rNorm);

/* Newton / GMRES parameters (sysstruct flags, preprocessing/digasopre.h:749-778). */
typedef struct {
    double newton_tol;   /* NewtonTol */
    int newton_maxiter;  /* NewtonMaxiter */
    int precond;         /* -1 none, 0 BILU0, 2 block-Jacobi */
    int ordering;        /* 0 natural, 1 approximate MDF */
    int reuse_ordering;  /* 1: keep an existing ordering (reuseOrdering) */
    int restart;         /* GMRES restart */
    double gmres_tol;    /* GMREStol */
    int gmres_maxit;
    int orthog;          /* 0 MGS, 1 ICGS */
    int precision;       /* condensation precision, HFEM_FP64 or HFEM_MIXED */
    int quasi_newton;    /* 1: reuse the Jacobian and its preconditioner (app.quasiNewton) */
    int qn_accuracy;     /* stored inv(D) and K: 0 float, 1 double (quasiNewtonAccuracy) */
    int true_newton_maxiter; /* limit on iterations that assemble the Jacobian (trueNewtonMaxiter); 0 none */
    double qn_reuse_ratio;   /* reuse while the GMRES time since the last Jacobian is below this
                                multiple of that Jacobian's assembly + preconditioner time (3 in
                                solvers/dgsolvers.cpp:137-141) */
    int prec_precision;      /* preconditioner factors: 0 single (FP32, factored on the FP32 MFMA),
                                1 double (sys.precPrecision) */
} hfem_solver_params;

/* solveNonlinearProblem (solvers/dgsolvers.cpp:142-205): Newton with the
 * reference's halving line search (a failed line search returns to the previous
 * iterate the first time in a solve, or always with quasi-Newton: robustMode,
 * :114, :175-177) and, with quasi_newton, its Jacobian reuse (solveLinearProblem
 * :40-105; a damped step or a reuse test that fails recomputes the Jacobian).
 * history (HFEM_HIST doubles per iteration, may be NULL): residual before,
 * residual after, GMRES iterations (negative when the iteration reused the
 * Jacobian), accumulated step, number of line-search steps applied
 * (updateSol calls). */
#define HFEM_HIST 5
int hfem_solve_steady(hfem_ctx *ctx, const hfem_solver_params *p, int *converged, int *niter, double *history);
/* One time step of solveUnsteadyProblem (solvers/dgsolvers.cpp:221-444) with
 * DIRK(nstage, order) of solvers/DIRKcoeff.cpp; each stage is a steady solve
 * with fc_u = d_ii/dt and the stage source SH. */
int hfem_dirk_step(hfem_ctx *ctx, double dt, int nstage, int order, const hfem_solver_params *p, int *converged,
                   int *newton_iters);

/* Timings of the last solver calls, like the breakdown solveLinearProblem prints
 * (solvers/dgsolvers.cpp:107-113): out[7] = {ordering ms, factorisation ms,
 * GMRES ms, matvec ms (device), matvec count, preconditioner ms (device),
 * preconditioner count}. */
int hfem_solver_stats(hfem_ctx *ctx, double *out);

/* Milliseconds spent in the last hfem_assemble / hfem_condense launch (HIP events). */
int hfem_last_kernel_ms(hfem_ctx *ctx, float *assemble_ms, float *condense_ms);
/* Synchronise the context stream. */
int hfem_sync(hfem_ctx *ctx);

/* No counterpart in the reference: use = 1 makes hfem_assemble_linear_system (and
 * so every Newton iteration of hfem_solve_steady and hfem_dirk_step) replay its
 * FP64 device work -- element fluxes and assembly, the residual norm's sums, the
 * condensation, the BSR scatter -- as one HIP graph, captured on first use and
 * again whenever its launch arguments change (physics, condensation choices,
 * buffers), the default; 0 launches it kernel by kernel.  The mixed
 * condensation and quasi-Newton storage always run kernel by kernel. */
int hfem_set_linearization_graph(hfem_ctx *ctx, int use);

#ifdef __cplusplus
}
#endif
#endif
