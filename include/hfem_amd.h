/*
 * hfem_amd: MI355X (gfx950) HDG element path -- C ABI.
 *
 * Drop-in boundary for the reference's per-element hot path.  The reference
 * binds it at link time from C++ with struct arguments; each entry point below
 * replaces the reference function named beside it, batched over elements and
 * with the element data resident in device memory:
 *
 *   hfem_assemble        assembleElementMatrixVector  assembly/elementassembly.cpp:1481-1499
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
/* Static condensation of elements [e0, e0+n) on the device blocks: inv(D)F,
 * inv(D)Ru, AH (in place of H) and FH; D, Ru and Rh keep their assembled values. */
int hfem_condense(hfem_ctx *ctx, int e0, int n, int precision);

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
 * rNorm = sqrt(sum_e |Ru_e|^2 + |r|^2) with r the assembled uncondensed Rh. */
int hfem_assemble_linear_system(hfem_ctx *ctx, int precision, double *rNorm);
/* Copy Hg (nblocks*bsz*bsz, column-major blocks) and Rg (nf*bsz) to the host. */
int hfem_read_system(hfem_ctx *ctx, double *Hg, double *Rg);
/* y = Hg x with host vectors of length nf*bsz. */
int hfem_bsr_matvec(hfem_ctx *ctx, const double *x, double *y);

/* Milliseconds spent in the last hfem_assemble / hfem_condense launch (HIP events). */
int hfem_last_kernel_ms(hfem_ctx *ctx, float *assemble_ms, float *condense_ms);
/* Synchronise the context stream. */
int hfem_sync(hfem_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
