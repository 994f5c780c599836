// Private definitions shared by the hfem_amd host sources (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

// rocblas_gemm_strided_batched_ex_get_solutions (solution selection for the
// condensation GEMMs) is declared under the beta API
#define ROCBLAS_BETA_FEATURES_API
#include <rocblas/rocblas.h>

#include "../../include/hfem_amd.h"
#include "hfem_device.h"

namespace hfem {
// after a BILU0 factorisation: upper blocks U_rl -> inv(U_rr) U_rl, the form the
// backward sweeps read
hipError_t launch_bilu0_scale_upper(int bsz, int n, const int *rowpts, const int *colind, const int *u2o,
                                    double *Mg, hipStream_t s);
hipError_t launch_bilu0_scale_upper(int bsz, int n, const int *rowpts, const int *colind, const int *u2o,
                                    float *Mg, hipStream_t s);
hipError_t launch_assemble(const Dims &, const Master &, const Phys &, const double *, const double *, const double *,
                           const double *, const int *, const int *, int, int, int, int, double *, long,
                           int, double *, double *, double *, double *, double *, double *, double *, double *,
                           hipStream_t, int res = 0);
size_t flux_lds_bytes(const Dims &d);
hipError_t launch_dinv(int na, int e0, int n, const double *D, double *Dinv, int *info, hipStream_t s,
                       int variant = 0);
hipError_t launch_dinv32(int na, int e0, int n, const double *D, float *Dinv, int *info, hipStream_t s,
                         int variant = 0);
constexpr int kDinvVariants = 3;
hipError_t launch_refine_reset(int n, unsigned long long *dmax, unsigned long long *cmax, int *count, hipStream_t s);
hipError_t launch_refine_count(int n, const unsigned long long *dmax, const unsigned long long *cmax,
                               unsigned long long *dprev, int first, double tol, int *count, hipStream_t s);
// bgemm.hip: batched per-element products (see the file's header for the layout)
struct BgemmArgs {
    int M, N, K;
    const void *A;
    long sa;
    int lda;
    const void *B;
    long sb;
    int ldb;
    double alpha, beta;
    const double *C0;
    long sc0;
    int ldc0;
    double *C;
    long sc;
    int ldc;
    int e0, n;                        // element range
    int nstrip;                       // strips per element (set by the launcher)
    unsigned long long *dmax, *cmax;  // optional per-element maxima (index e - e0)
};
int bgemm_num_variants();
// xsolve.hip: fused per-element solve X = inv(D)[F Ru] (mode 0 FP64, mode 1 mixed)
bool xsolve_fits(int mode, int na, size_t maxlds);
hipError_t launch_xsolve(int mode, int na, int N, int ldx, int e0, int n, const double *D, const void *inv,
                         const double *Fx, double *X, int nsweep, int from_x, unsigned long long *dmax,
                         unsigned long long *cmax, hipStream_t s);
bool bgemm_fits(int mode, int variant, int M, int N, size_t maxlds);
hipError_t launch_bgemm(int mode, int variant, BgemmArgs g, size_t maxlds, hipStream_t s);
hipError_t launch_schur_vector(int na, int nb, int ldx, int ne, int fp32, const void *Dinv, const void *K,
                               const double *Ru, const double *Rh, double *X, double *Hx, hipStream_t s);
hipError_t launch_to_f32(long n, const double *x, float *y, hipStream_t s);
size_t assemble_lds_bytes(const Dims &d);
bool assemble_supported(const Dims &d);
long assemble_spill_doubles(const Dims &d);
hipError_t launch_bsr_gather(int, int, int, int, int, long, const int *, const int *, const int *, const double *,
                             double *, hipStream_t);
hipError_t launch_face_gather(int, int, int, int, int, long, const int *, const int *, const double *, double *,
                              const double *, double *, hipStream_t);
hipError_t launch_sumsq(long, const double *, int, double *, int, hipStream_t);
hipError_t launch_bsr_matvec(int, int, const int *, const int *, const double *, const double *, double *,
                             hipStream_t);
hipError_t launch_getq(const Dims &, const Master &, const Phys &, const double *, double *, const double *,
                       const double *, const int *, const int *, int, int, hipStream_t);
size_t getq_lds_bytes(const Dims &d);
// solver.hip
size_t block_inverse_lds(int bsz);
size_t bilu0_factor_lds(int bsz);
hipError_t launch_block_inverse(int nrows, int bsz, const int *rowpts, const double *Hg, double *Dinv, int *info,
                                hipStream_t s);
hipError_t launch_mdf_weights(int nrows, int bsz, const int *rowpts, const double *Hg, const double *Dinv,
                              double *Cw, hipStream_t s);
hipError_t launch_bilu0_factor_level(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *rows,
                                     int nlev, double *Mg, int *info, hipStream_t s);
hipError_t launch_bilu0_lsolve_level(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *rows,
                                     int nlev, const double *Mg, const double *x, double *y, hipStream_t s);
hipError_t launch_bilu0_usolve_level(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *rows,
                                     int nlev, const double *Mg, const double *y, double *v, hipStream_t s);
hipError_t launch_block_jacobi(int nrows, int bsz, const double *Dinv, const double *x, double *y, hipStream_t s);
// single-precision preconditioner factors (FP32 storage, FP32 MFMA factorisation)
hipError_t launch_bilu0_factor_level(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *rows,
                                     int nlev, float *Mg, int *info, hipStream_t s);
hipError_t launch_bilu0_lsolve_level(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *rows,
                                     int nlev, const float *Mg, const double *x, double *y, hipStream_t s);
hipError_t launch_bilu0_usolve_level(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *rows,
                                     int nlev, const float *Mg, const double *y, double *v, hipStream_t s);
hipError_t launch_block_jacobi(int nrows, int bsz, const float *Dinv, const double *x, double *y, hipStream_t s);
hipError_t launch_to_float(long n, const double *x, float *y, hipStream_t s);
hipError_t launch_bilu0_factor_syncfree(int bsz, const int *rowpts, const int *colind, const int *u2o,
                                        const int *order, int n, double *Mg, int *info, unsigned *flags,
                                        unsigned *sync, int nwg, int spin_limit, int *err, hipStream_t s);
hipError_t launch_bilu0_factor_syncfree(int bsz, const int *rowpts, const int *colind, const int *u2o,
                                        const int *order, int n, float *Mg, int *info, unsigned *flags,
                                        unsigned *sync, int nwg, int spin_limit, int *err, hipStream_t s);
hipError_t launch_bilu0_syncfree(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *lorder,
                                 const int *uorder, int n, const double *Mg, const double *x, double *t, double *y,
                                 unsigned *flags, unsigned *sync, int nwg, int spin_limit, int *err, hipStream_t s);
hipError_t launch_bilu0_syncfree(int bsz, const int *rowpts, const int *colind, const int *u2o, const int *lorder,
                                 const int *uorder, int n, const float *Mg, const double *x, double *t, double *y,
                                 unsigned *flags, unsigned *sync, int nwg, int spin_limit, int *err, hipStream_t s);
int multidot_blocks(long n);
hipError_t launch_multidot(long n, int k, const double *V, long ldv, const double *w, double *partial, double *out,
                           hipStream_t s);
hipError_t launch_multi_axpy(long n, int k, const double *V, long ldv, const double *coef, double sign, double *w,
                             hipStream_t s);
hipError_t launch_axpby(long n, double a, const double *x, double b, double *y, hipStream_t s);
hipError_t launch_fill(long n, double v, double *y, hipStream_t s);
hipError_t launch_update_u(int ne, int npv, int nc, int nch, int ndf, int na, int nb, int ldx, const int *elcon,
                           const double *X, const double *duh, double alpha, double *UDG, hipStream_t s);
hipError_t launch_stage_source(long ne, int npv, int nc, int ncu, double a, const double *UDGi, double b,
                               const double *UDG, double c, double *SH, hipStream_t s);
}  // namespace hfem


using namespace hfem;

int hfem_fail(int code, const char *fmt, ...);

#define HIPCHK(x)                                                                                     \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) return hfem_fail(-2, "%s failed: %s", #x, hipGetErrorString(e_));       \
    } while (0)

// Device state of the trace-system solver: preconditioner blocks, ordering and
// level schedules, Krylov basis.  Allocated on first use.
struct SolverState {
    int prec = -2;                      // -2 nothing factored, -1 none, 0 BILU0, 2 block-Jacobi
    int order_method = -1;              // ordering behind o2u/u2o: 0 natural, 1 approximate MDF, 3 caller's
    std::vector<int> rowpts, colind;    // host copy of the trace graph
    std::vector<int> o2u, u2o;          // host copies
    int *d_o2u = nullptr, *d_u2o = nullptr;
    std::vector<int> lptr, uptr;        // level pointers of the L (forward) and U (backward) sweeps
    int *d_lrows = nullptr, *d_urows = nullptr;  // rows of each level, concatenated
    double *Mg = nullptr;               // BILU0 factors (nblocks blocks) or block-Jacobi inverses (nf blocks)
    float *Mg32 = nullptr;              // the same in single precision (precPrecision 0)
    int prec_fp32 = 0;
    // optionally, the BILU0 application (every level launch of both sweeps, yv -> t
    // -> gy) captured once per factorisation as a HIP graph and replayed
    hipGraphExec_t apply_graph = nullptr;
    bool graph_dirty = true;
    int use_graph = 0;
    double *gy = nullptr;
    // the BILU0 application without level launches (sweep_mode 1, default): 2 nf ready
    // flags, then the epoch and two ticket counters; a flag set when a bounded wait
    // gave up
    int sweep_mode = 1, spin_limit = 1 << 20, sweep_wgs = 1024;
    unsigned *sf_flags = nullptr;
    int *sf_err = nullptr;
    double *Dinv = nullptr;             // inverted diagonal blocks of Hg (MDF weights)
    double *Cw = nullptr;               // MDF block weights ||inv(H_ii) H_ij||_F
    int *d_info = nullptr;
    double *x = nullptr;                // GMRES solution (trace update duh)
    double *V = nullptr;                // Krylov basis, (restart+1) columns of length N
    int vcols = 0;
    double *b = nullptr, *t = nullptr;  // preconditioned rhs, work vector
    double *dots = nullptr;             // partial sums of the multi-dot kernel
    double *coef = nullptr;             // device copy of Gram-Schmidt coefficients
    double *UDGi = nullptr;             // DIRK: solution at the start of the time step
    // timings of the last calls (milliseconds; device events inside GMRES)
    double ms_order = 0, ms_factor = 0, ms_gmres = 0, ms_matvec = 0, ms_prec = 0;
    int n_matvec = 0, n_prec = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

struct hfem_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hfem_dims hd{};
    Dims d{};
    Phys ph{};
    bool has_master = false, has_phys = false, has_mesh = false, has_sol = false;
    // master
    double *shapvt = nullptr, *shapvg = nullptr, *shapft = nullptr, *shapfg = nullptr;
    int *perm = nullptr, *invperm = nullptr;
    // mesh
    double *dgnodes = nullptr;
    int *elcon = nullptr, *bf = nullptr, *rowstart = nullptr, *cols = nullptr, *einv = nullptr, *f2e = nullptr;
    int *blkptr = nullptr, *blklist = nullptr;
    int nblocks = 0;
    // state
    double *UDG = nullptr, *UH = nullptr, *SH = nullptr;
    // element blocks: D (na x na), Fx and X (na x ldx, row-major, column nb = Ru / inv(D)Ru),
    // K (nb x na), Hx (nb x (nb+1), column nb = Rh / FH), Ru, Rh (uncondensed)
    double *D = nullptr, *Fx = nullptr, *X = nullptr, *K = nullptr, *Hx = nullptr, *Ru = nullptr, *Rh = nullptr,
           *ru2 = nullptr;
    int *info = nullptr;
    int ldx = 0;
    // work
    double *dinv = nullptr, *fluxbuf = nullptr;   // dinv: ne x na x na
    // mixed condensation: the FP64 residual [F Ru] - D X (ne x na x ldx), allocated
    // by the first mixed condensation
    double *rwork = nullptr;
    double *spill = nullptr;   // per-workgroup element tables of families beyond the LDS
    rocblas_handle blas = nullptr;
    // rocBLAS solution indices of the condensation GEMMs (0: rocBLAS's own choice):
    // X = inv(D)[F Ru], the Schur product, the mixed residual and update
    int gemm_sol[4] = {0, 0, 0, 0};
    // condensation launch choices (hfem_tune_condensation): inverse variants (FP64,
    // FP32) and, per product (0 X = inv(D)[F Ru], 1 Schur, 2 mixed X0, 3 mixed
    // residual, 4 mixed correction), the hand-written variant (>= 0) or rocBLAS (-1)
    int inv_var[2] = {0, 0};
    int prod_var[5] = {0, 0, 0, 0, 0};
    // mixed refinement: per-element maxima, previous correction, unconverged count
    unsigned long long *rdmax = nullptr, *rcmax = nullptr, *rdprev = nullptr;
    int *rcount = nullptr;
    int mix_max_sweeps = 4;
    double mix_tol = 1e-11;
    int last_sweeps = 0, last_unconverged = 0, last_fallback = 0;
    long fstride = 0;
    int chunk = 0;
    int grid_asm = 0, grid_cond = 0;
    int maxlds = 0;
    // trace system
    double *Hg = nullptr, *Rg = nullptr, *r = nullptr, *partial = nullptr, *xv = nullptr, *yv = nullptr;
    float ms_asm = 0, ms_cond = 0;
    // quasi-Newton: inv(D) and K of the last Jacobian (qn_acc -1 off, 0 float, 1 double)
    int qn_acc = -1;
    bool qn_valid = false;
    void *qn_dinv = nullptr, *qn_K = nullptr;
    // trace-system solver (solve.hip)
    SolverState sv;
    // linearisation graph (hfem_set_linearization_graph): the FP64 element path and
    // BSR scatter of hfem_assemble_linear_system captured once and replayed, with
    // the launch arguments it was captured with (recaptured when they change)
    bool lin_graph = true, capturing = false;
    hipGraphExec_t lin_exec = nullptr;
    int lin_mixed_sweeps = 0;   // sweeps the last kernel-by-kernel mixed condensation needed (0: none yet)
    struct LinKey {
        Phys ph;
        int precision, sweeps, inv_var1, prod_var2, prod_var3, prod_var4, gemm_sol2;
        double mix_tol;
        int inv_var0, prod_var0, prod_var1, gemm_sol0, gemm_sol1;
        const void *ptr[20];   // buffers the setters may (re)allocate
    } lin_key;
};


int hfem_reduce_sum(hfem_ctx *c, long n, const double *x, int square, double *out);
// X = inv(D)[F Ru] and the Schur product of [e0, e0+n) from the inverse in c->dinv
int condense_products_fp64(hfem_ctx *c, int e0, int n);


template <class T>
static inline hipError_t hfem_dalloc(T **p, size_t n)
{
    return hipMalloc((void **)p, (n > 0 ? n : 1) * sizeof(T));
}
