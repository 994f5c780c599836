// C ABI of the hfem_amd library (declared in include/hfem_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"



static thread_local std::string g_err;

int hfem_fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}


extern "C" const char *hfem_last_error(void) { return g_err.c_str(); }

template <class T>
static hipError_t dalloc(T **p, size_t n)
{
    return hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
}

static void default_choices(hfem_ctx *c);

extern "C" int hfem_create(hfem_ctx **out, const hfem_dims *hd, int device)
{
    if (!out || !hd) return hfem_fail(-1, "hfem_create: null argument");
    *out = nullptr;
    if (hd->nd != 2 && hd->nd != 3) return hfem_fail(-1, "hfem_create: nd must be 2 or 3 (got %d)", hd->nd);
    if (hd->ncu != hd->nd + 2) return hfem_fail(-1, "hfem_create: compressible NS needs ncu = nd+2");
    if (hd->npv <= 0 || hd->ngv <= 0 || hd->npf <= 0 || hd->ngf <= 0 || hd->nfe <= 0 || hd->ne <= 0 || hd->nf <= 0)
        return hfem_fail(-1, "hfem_create: sizes must be positive");
    if (hd->npv > 64 || hd->npf * hd->nfe > 128) return hfem_fail(-1, "hfem_create: element too large (npv<=64)");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return hfem_fail(-1, "hfem_create: device %d out of range (%d devices)", device, ndev);
    HIPCHK(hipSetDevice(device));
    hfem_ctx *c = new hfem_ctx();
    c->device = device;
    c->hd = *hd;
    Dims &d = c->d;
    d.nd = hd->nd; d.ncu = hd->ncu; d.nch = hd->ncu; d.nd1 = hd->nd + 1; d.nc = d.ncu * d.nd1;
    d.npv = hd->npv; d.ngv = hd->ngv; d.npf = hd->npf; d.ngf = hd->ngf; d.nfe = hd->nfe;
    d.ndf = d.npf * d.nfe; d.na = d.npv * d.ncu; d.nb = d.ndf * d.nch; d.np = d.ngf * d.nfe;
    if (d.na > 176) { delete c; return hfem_fail(-1, "hfem_create: na=%d exceeds 176 (npv*ncu)", d.na); }
    const long ne = hd->ne;
    hipError_t e = hipSuccess;
#define ALLOC(p, n) if (e == hipSuccess) e = dalloc(&c->p, (size_t)(n))
    c->ldx = (d.nb + 1 + 15) / 16 * 16;
    ALLOC(D, ne * d.na * d.na);
    ALLOC(Fx, ne * d.na * c->ldx);
    ALLOC(X, ne * d.na * c->ldx);
    ALLOC(K, ne * d.nb * d.na);
    ALLOC(Hx, ne * d.nb * (d.nb + 1));
    ALLOC(Ru, ne * d.na);
    ALLOC(Rh, ne * d.nb);
    ALLOC(ru2, ne);
    ALLOC(info, ne);
    ALLOC(partial, 1024);
    c->grid_asm = (int)std::min<long>(ne, 512);
    c->grid_cond = (int)std::min<long>(ne, 1024);
    c->fstride = flux_layout(d).total;
    // one chunk for meshes up to 32768 elements (the flux buffer holds about 160 KB per
    // element for p=3 tets): no chunk boundaries to drain
    c->chunk = (int)std::min<long>(ne, 32768);
    ALLOC(fluxbuf, (long)c->chunk * c->fstride);
    {
        int maxlds = 65536;
        if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
            maxlds = 65536;
        c->maxlds = maxlds;
        if (assemble_lds_bytes(d) > (size_t)maxlds) {
            delete c;
            return hfem_fail(-1, "hfem_create: element family needs %zu bytes of LDS, device offers %d",
                        assemble_lds_bytes(d), maxlds);
        }
        if (!assemble_supported(d)) {
            delete c;
            return hfem_fail(-1, "hfem_create: no assembly kernel compiled for nd=%d npv=%d npf=%d", d.nd, d.npv, d.npf);
        }
    }
    ALLOC(dinv, ne * d.na * d.na);
    ALLOC(rdmax, ne);
    ALLOC(rcmax, ne);
    ALLOC(rdprev, ne);
    ALLOC(rcount, 17);   // [0]: the adaptive loop's count; [1 + s]: sweep s of a fixed-sweep graph
    if (assemble_spill_doubles(d) > 0) ALLOC(spill, (long)c->grid_asm * assemble_spill_doubles(d));
#undef ALLOC
    if (e == hipSuccess) e = hipStreamCreate(&c->stream);
    if (e == hipSuccess && (rocblas_create_handle(&c->blas) != rocblas_status_success ||
                            rocblas_set_stream(c->blas, c->stream) != rocblas_status_success))
        e = hipErrorInitializationError;
    for (int i = 0; i < 4 && e == hipSuccess; i++) e = hipEventCreate(&c->ev[i]);
    if (e != hipSuccess) {
        hfem_destroy(c);
        return hfem_fail(-2, "hfem_create: allocation failed: %s", hipGetErrorString(e));
    }
    default_choices(c);
    *out = c;
    return 0;
}

extern "C" int hfem_destroy(hfem_ctx *c)
{
    if (!c) return 0;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    void *ptrs[] = {c->shapvt, c->shapvg, c->shapft, c->shapfg, c->perm, c->invperm, c->dgnodes, c->elcon, c->bf,
                    c->rowstart, c->cols, c->einv, c->f2e, c->blkptr, c->blklist, c->UDG, c->UH, c->SH, c->D, c->Fx,
                    c->X, c->K, c->Hx, c->Ru, c->Rh, c->ru2, c->info, c->dinv, c->spill, c->rwork, c->fluxbuf, c->Hg, c->Rg,
                    c->r, c->partial, c->xv, c->yv, c->sv.d_o2u, c->sv.d_u2o, c->sv.d_lrows, c->sv.d_urows,
                    c->sv.Mg, c->sv.Mg32, c->sv.gy, c->sv.Dinv, c->sv.Cw, c->sv.d_info, c->sv.x, c->sv.V, c->sv.b, c->sv.t, c->sv.dots,
                    c->sv.coef, c->sv.UDGi, c->sv.sf_flags, c->sv.sf_err, c->qn_dinv, c->qn_K, c->rdmax, c->rcmax, c->rdprev, c->rcount};
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (auto &ev : c->ev)
        if (ev) hipEventDestroy(ev);
    for (auto &ev : c->sv.ev)
        if (ev) hipEventDestroy(ev);
    if (c->sv.apply_graph) hipGraphExecDestroy(c->sv.apply_graph);
    if (c->lin_exec) hipGraphExecDestroy(c->lin_exec);
    if (c->blas) rocblas_destroy_handle(c->blas);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

template <class T>
static int upload(hfem_ctx *c, T **dst, const T *src, size_t n)
{
    if (!*dst) HIPCHK(dalloc(dst, n));
    HIPCHK(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hfem_set_master(hfem_ctx *c, const double *shapvt, const double *shapvg, const double *shapft,
                               const double *shapfg, const int *perm)
{
    if (!c || !shapvt || !shapvg || !shapft || !shapfg || !perm) return hfem_fail(-1, "hfem_set_master: null argument");
    HIPCHK(hipSetDevice(c->device));
    const Dims &d = c->d;
    std::vector<int> inv((size_t)d.npv * d.nfe, -1);
    for (int a = 0; a < d.nfe; a++)
        for (int k = 0; k < d.npf; k++) {
            int nn = perm[k + d.npf * a];
            if (nn < 0 || nn >= d.npv) return hfem_fail(-1, "hfem_set_master: perm(%d,%d)=%d out of range", k, a, nn);
            inv[nn + (size_t)d.npv * a] = k;
        }
    int rc;
    if ((rc = upload(c, &c->shapvt, shapvt, (size_t)d.ngv * d.npv * d.nd1))) return rc;
    if ((rc = upload(c, &c->shapvg, shapvg, (size_t)d.ngv * d.npv * d.nd1))) return rc;
    if ((rc = upload(c, &c->shapft, shapft, (size_t)d.ngf * d.npf * d.nd))) return rc;
    if ((rc = upload(c, &c->shapfg, shapfg, (size_t)d.ngf * d.npf * d.nd))) return rc;
    if ((rc = upload(c, &c->perm, perm, (size_t)d.npf * d.nfe))) return rc;
    if ((rc = upload(c, &c->invperm, inv.data(), inv.size()))) return rc;
    c->has_master = true;
    return 0;
}

extern "C" int hfem_set_physics(hfem_ctx *c, const double *param, const double *ui, double fc_u, double fc_q, int tdep)
{
    if (!c || !param || !ui) return hfem_fail(-1, "hfem_set_physics: null argument");
    if (param[0] <= 1.0 || param[2] <= 0 || param[3] <= 0) return hfem_fail(-1, "hfem_set_physics: invalid gamma/Re/Pr");
    for (int i = 0; i < 6; i++) c->ph.param[i] = param[i];
    for (int i = 0; i < c->d.nch; i++) c->ph.ui[i] = ui[i];
    c->ph.fc_u = fc_u;
    c->ph.fc_q = fc_q;
    c->ph.tdep = tdep;
    c->has_phys = true;
    return 0;
}

extern "C" int hfem_set_mesh(hfem_ctx *c, const double *dgnodes, const int *elcon, const int *bf, const int *t2f,
                             const int *rowstart, const int *cols, int nblocks)
{
    if (!c || !dgnodes || !elcon || !bf || !t2f || !rowstart || !cols) return hfem_fail(-1, "hfem_set_mesh: null argument");
    HIPCHK(hipSetDevice(c->device));
    const Dims &d = c->d;
    const int ne = c->hd.ne, nf = c->hd.nf;
    if (rowstart[0] != 0 || rowstart[nf] != nblocks) return hfem_fail(-1, "hfem_set_mesh: rowstart inconsistent with nblocks");
    // the trace-system buffers are sized by the first mesh's block count
    if (c->has_mesh && nblocks != c->nblocks)
        return hfem_fail(-1, "hfem_set_mesh: the trace graph has %d blocks, this context's has %d (create a new "
                             "context for a mesh of another size)", nblocks, c->nblocks);
    // einv, face-to-element lists and the gather lists of every BSR block
    std::vector<int> einv((size_t)ne * d.ndf, -1), f2e((size_t)nf * 4, -1);
    std::vector<int> cnt(nblocks + 1, 0);
    std::vector<int> kidx((size_t)ne * d.nfe * d.nfe);
    for (int e = 0; e < ne; e++)
        for (int a = 0; a < d.nfe; a++) {
            const int i = t2f[e * d.nfe + a];
            if (i < 0 || i >= nf) return hfem_fail(-1, "hfem_set_mesh: t2f[%d,%d]=%d out of range", e, a, i);
            for (int m = 0; m < d.npf; m++) {
                int g = elcon[(long)e * d.ndf + a * d.npf + m] - i * d.npf;
                if (g < 0 || g >= d.npf) return hfem_fail(-1, "hfem_set_mesh: elcon of element %d face %d not on face %d", e, a, i);
                einv[(long)e * d.ndf + a * d.npf + g] = m;
            }
            int side = f2e[4 * i] < 0 ? 0 : 1;
            if (side == 1 && f2e[4 * i + 2] >= 0) return hfem_fail(-1, "hfem_set_mesh: face %d has more than two elements", i);
            f2e[4 * i + 2 * side] = e;
            f2e[4 * i + 2 * side + 1] = a;
            for (int b = 0; b < d.nfe; b++) {
                const int j = t2f[e * d.nfe + b];
                int k = rowstart[i];
                while (k < rowstart[i + 1] && cols[k] != j) k++;
                if (k == rowstart[i + 1]) return hfem_fail(-1, "hfem_set_mesh: block (%d,%d) missing from the trace graph", i, j);
                kidx[((size_t)e * d.nfe + a) * d.nfe + b] = k;
                cnt[k + 1]++;
            }
        }
    for (size_t q = 0; q < einv.size(); q++)
        if (einv[q] < 0) return hfem_fail(-1, "hfem_set_mesh: elcon of element %d maps two nodes to one trace node", (int)(q / d.ndf));
    for (int k = 0; k < nblocks; k++) cnt[k + 1] += cnt[k];
    std::vector<int> list((size_t)cnt[nblocks] * 3), fillp(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < ne; e++)
        for (int a = 0; a < d.nfe; a++)
            for (int b = 0; b < d.nfe; b++) {
                int k = kidx[((size_t)e * d.nfe + a) * d.nfe + b];
                int q = fillp[k]++;
                list[3 * q] = e; list[3 * q + 1] = a; list[3 * q + 2] = b;
            }
    int rc;
    if ((rc = upload(c, &c->dgnodes, dgnodes, (size_t)ne * d.npv * d.nd))) return rc;
    if ((rc = upload(c, &c->elcon, elcon, (size_t)ne * d.ndf))) return rc;
    if ((rc = upload(c, &c->bf, bf, (size_t)ne * d.nfe))) return rc;
    if ((rc = upload(c, &c->rowstart, rowstart, (size_t)nf + 1))) return rc;
    if ((rc = upload(c, &c->cols, cols, (size_t)nblocks))) return rc;
    if ((rc = upload(c, &c->einv, einv.data(), einv.size()))) return rc;
    if ((rc = upload(c, &c->f2e, f2e.data(), f2e.size()))) return rc;
    if ((rc = upload(c, &c->blkptr, cnt.data(), cnt.size()))) return rc;
    if ((rc = upload(c, &c->blklist, list.data(), list.size()))) return rc;
    c->nblocks = nblocks;
    c->sv.rowpts.assign(rowstart, rowstart + nf + 1);
    c->sv.colind.assign(cols, cols + nblocks);
    // a new trace graph invalidates the ordering and the preconditioner built on the old one
    c->sv.order_method = -1;
    c->sv.prec = -2;
    c->sv.graph_dirty = true;
    for (int i = 0; i < nf; i++)
        if (rowstart[i + 1] <= rowstart[i] || cols[rowstart[i]] != i)
            return hfem_fail(-1, "hfem_set_mesh: row %d of the trace graph must start with its diagonal block", i);
    const long bsz = (long)d.nch * d.npf;
    if (!c->Hg) HIPCHK(dalloc(&c->Hg, (size_t)nblocks * bsz * bsz));
    if (!c->Rg) HIPCHK(dalloc(&c->Rg, (size_t)nf * bsz));
    if (!c->r) HIPCHK(dalloc(&c->r, (size_t)nf * bsz));
    if (!c->xv) HIPCHK(dalloc(&c->xv, (size_t)nf * bsz));
    if (!c->yv) HIPCHK(dalloc(&c->yv, (size_t)nf * bsz));
    c->has_mesh = true;
    return 0;
}

extern "C" int hfem_set_solution(hfem_ctx *c, const double *UDG, const double *UH, const double *SH)
{
    if (!c || !UDG || !UH) return hfem_fail(-1, "hfem_set_solution: null argument");
    HIPCHK(hipSetDevice(c->device));
    const Dims &d = c->d;
    const long ne = c->hd.ne, nuh = (long)c->hd.nf * d.npf;
    int rc;
    if ((rc = upload(c, &c->UDG, UDG, (size_t)ne * d.npv * d.nc))) return rc;
    if ((rc = upload(c, &c->UH, UH, (size_t)nuh * d.nch))) return rc;
    if (SH) {
        if ((rc = upload(c, &c->SH, SH, (size_t)ne * d.npv * d.nc))) return rc;
    } else if (!c->SH) {
        HIPCHK(dalloc(&c->SH, (size_t)ne * d.npv * d.nc));
        HIPCHK(hipMemsetAsync(c->SH, 0, sizeof(double) * ne * d.npv * d.nc, c->stream));
    }
    c->has_sol = true;
    return 0;
}

extern "C" int hfem_get_solution(hfem_ctx *c, double *UDG, double *UH)
{
    if (!c || !c->has_sol) return hfem_fail(-3, "hfem_get_solution: no solution set");
    HIPCHK(hipSetDevice(c->device));
    const Dims &d = c->d;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (UDG) HIPCHK(hipMemcpy(UDG, c->UDG, sizeof(double) * c->hd.ne * d.npv * d.nc, hipMemcpyDeviceToHost));
    if (UH) HIPCHK(hipMemcpy(UH, c->UH, sizeof(double) * (long)c->hd.nf * d.npf * d.nch, hipMemcpyDeviceToHost));
    return 0;
}

static int check_range(hfem_ctx *c, int e0, int n, const char *who)
{
    if (!c) return hfem_fail(-1, "%s: null context", who);
    if (e0 < 0 || n < 0 || (long)e0 + n > c->hd.ne) return hfem_fail(-1, "%s: element range [%d,%d) outside [0,%d)", who, e0, e0 + n, c->hd.ne);
    return 0;
}

static int assemble_range(hfem_ctx *c, int e0, int n, int res, const char *who)
{
    int rc = check_range(c, e0, n, who);
    if (rc) return rc;
    if (!c->has_master || !c->has_phys || !c->has_mesh || !c->has_sol)
        return hfem_fail(-3, "%s: master, physics, mesh and solution must be set first", who);
    if (n == 0) return 0;
    HIPCHK(hipSetDevice(c->device));
    Master m{c->shapvt, c->shapvg, c->shapft, c->shapfg, c->perm, c->invperm};
    const int grid = std::min(n, c->grid_asm);
    if (!c->capturing) HIPCHK(hipEventRecord(c->ev[0], c->stream));
    HIPCHK(launch_assemble(c->d, m, c->ph, c->dgnodes, c->UDG, c->UH, c->SH, c->elcon, c->bf, e0, n, grid, c->ldx,
                           c->fluxbuf, c->fstride, c->chunk, c->Ru, c->D, c->Fx, c->K, c->Hx, c->Rh, c->ru2,
                           c->spill, c->stream, res));
    if (!c->capturing) HIPCHK(hipEventRecord(c->ev[1], c->stream));
    return 0;
}

extern "C" int hfem_assemble(hfem_ctx *c, int e0, int n) { return assemble_range(c, e0, n, 0, "hfem_assemble"); }

// assembleElementVector (assembly/elementassembly.cpp:1501-1512): the residual
// rows only.  The reference still forms every Jacobian there (:358, :785); here
// the flux kernel evaluates the fluxes without derivatives and the assembly skips
// the matrix phases.  Ru, Rh (and the [F Ru] / [H Rh] columns, |Ru|^2) are
// written; D, F, K, H keep their values.
extern "C" int hfem_assemble_vector(hfem_ctx *c, int e0, int n)
{
    return assemble_range(c, e0, n, 1, "hfem_assemble_vector");
}

// The two batched GEMMs of the FP64 condensation, over n elements:
//   which 0: X^T = [F Ru]^T inv(D)^T  (X row-major = column-major X^T, ld ldx)
//   which 1: [AH FH] = [H Rh] - K X
struct CondGemm {
    rocblas_operation ta, tb;
    int m, n, k;
    const double *alpha, *A, *B, *beta;
    double *C;
    int lda, ldb, ldc;
    long sa, sb, sc;
    int batch;
    double *Dout = nullptr;     // output, same layout as C (nullptr: in place over C)
};

static const double kOne = 1.0, kZero = 0.0, kMinusOne = -1.0;

static CondGemm cond_gemm(const hfem_ctx *c, int which, const double *Fx, const double *dinv, const double *K,
                          double *X, double *Hx, int n)
{
    const Dims &d = c->d;
    const long sx = (long)d.na * c->ldx, sd = (long)d.na * d.na, sk = (long)d.nb * d.na,
               sh = (long)d.nb * (d.nb + 1);
    if (which == 0)
        return CondGemm{rocblas_operation_none, rocblas_operation_transpose, d.nb + 1, d.na, d.na, &kOne, Fx, dinv,
                        &kZero, X, c->ldx, d.na, c->ldx, sx, sd, sx, n};
    return CondGemm{rocblas_operation_none, rocblas_operation_transpose, d.nb, d.nb + 1, d.na, &kMinusOne, K, X,
                    &kOne, Hx, d.nb, c->ldx, d.nb, sk, sx, sh, n};
}

// sol 0: rocBLAS's own choice; otherwise a solution index from the list for this shape
static rocblas_status cond_gemm_run(rocblas_handle h, const CondGemm &g, int sol)
{
    const rocblas_datatype t = rocblas_datatype_f64_r;
    return rocblas_gemm_strided_batched_ex(h, g.ta, g.tb, g.m, g.n, g.k, g.alpha, g.A, t, g.lda, g.sa, g.B, t, g.ldb,
                                           g.sb, g.beta, g.C, t, g.ldc, g.sc, g.Dout ? g.Dout : g.C, t, g.ldc, g.sc,
                                           g.batch, t,
                                           sol ? rocblas_gemm_algo_solution_index : rocblas_gemm_algo_standard, sol,
                                           0);
}

#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wdeprecated-declarations"
static rocblas_status cond_gemm_solutions(rocblas_handle h, const CondGemm &g, std::vector<rocblas_int> &out)
{
    const rocblas_datatype t = rocblas_datatype_f64_r;
    rocblas_int n = 0;
    double *D = g.Dout ? g.Dout : g.C;
    rocblas_status st = rocblas_gemm_strided_batched_ex_get_solutions(
        h, g.ta, g.tb, g.m, g.n, g.k, g.alpha, g.A, t, g.lda, g.sa, g.B, t, g.ldb, g.sb, g.beta, g.C, t, g.ldc, g.sc,
        D, t, g.ldc, g.sc, g.batch, t, rocblas_gemm_algo_solution_index, 0, nullptr, &n);
    if (st != rocblas_status_success) return st;
    out.resize(n);
    if (n == 0) return st;
    return rocblas_gemm_strided_batched_ex_get_solutions(
        h, g.ta, g.tb, g.m, g.n, g.k, g.alpha, g.A, t, g.lda, g.sa, g.B, t, g.ldb, g.sb, g.beta, g.C, t, g.ldc, g.sc,
        D, t, g.ldc, g.sc, g.batch, t, rocblas_gemm_algo_solution_index, 0, out.data(), &n);
}
#pragma clang diagnostic pop

// The batched GEMMs of the condensation, as tuned by hfem_tune_condensation:
//   0: X = inv(D) [F Ru]                      (both paths)
//   1: [AH FH] = [H Rh] - K X                 (both paths)
//   2: R = [F Ru] - D X   (C read, R written)  (mixed refinement residual)
//   3: X += inv32 R                           (mixed refinement update)
static CondGemm tuned_gemm(const hfem_ctx *c, int which, const double *Fx, const double *Dm, const double *dinv,
                           const double *K, double *X, double *Hx, double *R, int n)
{
    if (which == 0) return cond_gemm(c, 0, Fx, dinv, nullptr, X, nullptr, n);
    if (which == 1) return cond_gemm(c, 1, nullptr, nullptr, K, X, Hx, n);
    if (which == 2) {
        CondGemm g = cond_gemm(c, 0, X, Dm, nullptr, const_cast<double *>(Fx), nullptr, n);
        g.Dout = R;
        g.alpha = &kMinusOne;
        g.beta = &kOne;
        return g;
    }
    CondGemm g = cond_gemm(c, 0, R, dinv, nullptr, X, nullptr, n);
    g.beta = &kOne;
    return g;
}

// The condensation's products in the batched-GEMM form of bgemm.hip, over the
// elements [e0, e0+n):
//   0: X = inv(D) [F Ru]              FP64 MFMA
//   1: [AH FH] = [H Rh] - K X         FP64 MFMA (computed transposed: AH^T = H^T - X^T K^T)
//   2: X = inv32 [F Ru]               FP32 MFMA (mixed X0)
//   3: R = [F Ru] - D X               FP64 MFMA (mixed refinement residual)
//   4: X += inv32 R                   FP32 MFMA (mixed correction; per-element maxima)
// inv points at the FP64 (0) or FP32 (2, 4) inverse; R at the residual buffer.
static int prod_mode(int which) { return (which == 2 || which == 4) ? 1 : 0; }

static BgemmArgs prod_args(const hfem_ctx *c, int which, int e0, int n, const void *inv, const double *Fx,
                           const double *Dm, const double *K, double *X, double *Hx, double *R)
{
    const Dims &d = c->d;
    const long sx = (long)d.na * c->ldx, sd = (long)d.na * d.na, sk = (long)d.nb * d.na,
               sh = (long)d.nb * (d.nb + 1);
    BgemmArgs g{};
    g.e0 = e0;
    g.n = n;
    if (which == 1) {
        g.M = d.nb + 1; g.N = d.nb; g.K = d.na;
        g.A = X; g.sa = sx; g.lda = c->ldx;
        g.B = K; g.sb = sk; g.ldb = d.nb;
        g.alpha = -1.0; g.beta = 1.0;
        g.C0 = Hx; g.sc0 = sh; g.ldc0 = d.nb;
        g.C = Hx; g.sc = sh; g.ldc = d.nb;
        return g;
    }
    g.M = d.na; g.N = d.nb + 1; g.K = d.na;
    g.sa = sd; g.lda = d.na;
    g.sb = sx; g.ldb = c->ldx;
    g.sc = sx; g.ldc = c->ldx;
    g.sc0 = sx; g.ldc0 = c->ldx;
    g.alpha = 1.0;
    if (which == 0 || which == 2) { g.A = inv; g.B = Fx; g.C = X; }
    else if (which == 3) { g.A = Dm; g.B = X; g.C0 = Fx; g.beta = 1.0; g.alpha = -1.0; g.C = R; }
    else { g.A = inv; g.B = R; g.C0 = X; g.beta = 1.0; g.C = X; g.dmax = c->rdmax + e0; g.cmax = c->rcmax + e0; }
    return g;
}

// Choice value of products 0 and 2 for the fused per-element solve (xsolve.hip):
// product 0 -> X = inv(D)[F Ru] with inv(D) in LDS; product 2 -> X0 and every
// refinement sweep in one kernel (products 3 and 4 then unused).
static constexpr int kFused = 64;

// rocBLAS form of products 0, 1, 3 (the measured comparison; tuned_gemm numbering)
static int prod_rocblas_index(int which) { return which == 0 ? 0 : which == 1 ? 1 : which == 3 ? 2 : -1; }

static hipError_t run_prod(hfem_ctx *c, int which, int var, int e0, int n, const void *inv, const double *Fx,
                           const double *Dm, const double *K, double *X, double *Hx, double *R, int sol = 0)
{
    if (var == kFused) {
        const Dims &d = c->d;
        if (which == 0)
            return launch_xsolve(0, d.na, d.nb + 1, c->ldx, e0, n, Dm, inv, Fx, X, 0, 0, nullptr, nullptr, c->stream);
        if (which == 2)   // X0 and one sweep
            return launch_xsolve(1, d.na, d.nb + 1, c->ldx, e0, n, Dm, inv, Fx, X, 1, 0, c->rdmax + e0, c->rcmax + e0,
                                 c->stream);
        return hipErrorInvalidValue;
    }
    if (var < 0) {
        const int rw = prod_rocblas_index(which);
        if (rw < 0) return hipErrorInvalidValue;
        const Dims &d = c->d;
        const long sx = (long)d.na * c->ldx, sd = (long)d.na * d.na, sk = (long)d.nb * d.na,
                   sh = (long)d.nb * (d.nb + 1);
        const CondGemm g = tuned_gemm(c, rw, Fx + e0 * sx, Dm ? Dm + e0 * sd : nullptr,
                                      inv ? (const double *)inv + e0 * sd : nullptr, K ? K + e0 * sk : nullptr,
                                      X + e0 * sx, Hx ? Hx + e0 * sh : nullptr, R ? R + e0 * sx : nullptr, n);
        return cond_gemm_run(c->blas, g, sol) == rocblas_status_success ? hipSuccess : hipErrorLaunchFailure;
    }
    return launch_bgemm(prod_mode(which), var, prod_args(c, which, e0, n, inv, Fx, Dm, K, X, Hx, R),
                        (size_t)c->maxlds, c->stream);
}

// Default launch choices before any tuning: the first hand-written variant that
// fits each product.
static void default_choices(hfem_ctx *c)
{
    const Dims &d = c->d;
    for (int w = 0; w < 5; w++) {
        const int M = w == 1 ? d.nb + 1 : d.na, N = w == 1 ? d.nb : d.nb + 1;
        c->prod_var[w] = -1;
        for (int v = 0; v < bgemm_num_variants(); v++)
            if (bgemm_fits(prod_mode(w), v, M, N, (size_t)c->maxlds)) { c->prod_var[w] = v; break; }
    }
    for (int w = 0; w <= 2; w += 2)
        if (xsolve_fits(w == 0 ? 0 : 1, d.na, (size_t)c->maxlds)) c->prod_var[w] = kFused;
}

// Times the launch candidates of every stage of the condensation on a sample of
// nsample elements (default min(ne, 32768); the winners change with the batch
// size) and keeps the fastest: the inverse kernels' launch shapes (FP64, FP32),
// and for each product the hand-written variants that fit its shape plus, for the
// FP64 products 0, 1 and 3, rocBLAS (its own choice and every solution it lists).
// Inputs are the context's D, [F Ru] and K; outputs go to temporary buffers, so
// no element block of the context changes.  choice[7] (may be NULL) receives
// {inverse FP64, inverse FP32, product 0..4} (product: variant >= 0, or -1 for
// rocBLAS), ms[7] (may be NULL) the best time of each in milliseconds.
extern "C" int hfem_tune_condensation(hfem_ctx *c, int nsample, int *choice, float *ms)
{
    if (!c) return hfem_fail(-1, "hfem_tune_condensation: null context");
    HIPCHK(hipSetDevice(c->device));
    const Dims &d = c->d;
    const int ne = (int)c->hd.ne;
    if (nsample < 0) return hfem_fail(-1, "hfem_tune_condensation: nsample=%d < 0", nsample);
    if (ne == 0) return 0;
    const int ns = nsample > 0 ? std::min(nsample, ne) : std::min(ne, 32768);
    double *Xt = nullptr, *Ht = nullptr, *Rt = nullptr, *It = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int rc = 0;
    // Rt holds the timed inverses (ns na^2 doubles) and then the sample's FP32 inverse
    if (hfem_dalloc(&Xt, (size_t)ns * d.na * c->ldx) != hipSuccess ||
        hfem_dalloc(&Rt, (size_t)ns * d.na * std::max(c->ldx, d.na)) != hipSuccess ||
        hfem_dalloc(&It, (size_t)ns * d.na * d.na) != hipSuccess ||
        hfem_dalloc(&Ht, (size_t)ns * d.nb * (d.nb + 1)) != hipSuccess || hipEventCreate(&t0) != hipSuccess ||
        hipEventCreate(&t1) != hipSuccess)
        rc = hfem_fail(-2, "hfem_tune_condensation: out of device memory");
    // realistic operands: the inverses of the sample's D, X = inv(D)[F Ru], H
    if (!rc && (launch_dinv(d.na, 0, ns, c->D, It, c->info, c->stream) != hipSuccess ||
                hipMemcpyAsync(Ht, c->Hx, (size_t)ns * d.nb * (d.nb + 1) * 8, hipMemcpyDeviceToDevice,
                               c->stream) != hipSuccess))
        rc = hfem_fail(-2, "hfem_tune_condensation: setup failed");
    // best of two timed runs after one untimed run; -1 if the launch fails
    auto time_it = [&](auto &&launch) -> float {
        if (launch() != hipSuccess) { (void)hipGetLastError(); return -1.f; }
        float best = 1e30f;
        for (int r = 0; r < 2; r++) {
            float t = 0;
            if (hipEventRecord(t0, c->stream) != hipSuccess || launch() != hipSuccess ||
                hipEventRecord(t1, c->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
                hipEventElapsedTime(&t, t0, t1) != hipSuccess) {
                (void)hipGetLastError();
                return -1.f;
            }
            best = std::min(best, t);
        }
        return best;
    };
    float best[7];
    for (int p = 0; p < 2 && !rc; p++) {
        int bv = 0;
        float bt = 1e30f;
        for (int v = 0; v < kDinvVariants; v++) {
            const float t = time_it([&]() {
                return p == 0 ? launch_dinv(d.na, 0, ns, c->D, Rt, c->info, c->stream, v)
                              : launch_dinv32(d.na, 0, ns, c->D, (float *)Rt, c->info, c->stream, v);
            });
            if (t >= 0 && t < bt) { bt = t; bv = v; }
        }
        c->inv_var[p] = bv;
        best[p] = bt;
    }
    float *It32 = (float *)Rt;   // FP32 inverse of the sample for the mixed products (in Rt's first half)
    double *R2 = nullptr;
    if (!rc && (launch_dinv32(d.na, 0, ns, c->D, It32, c->info, c->stream) != hipSuccess ||
                hfem_dalloc(&R2, (size_t)ns * d.na * c->ldx) != hipSuccess ||
                hipMemsetAsync(R2, 0, (size_t)ns * d.na * c->ldx * 8, c->stream) != hipSuccess ||
                launch_refine_reset(ns, c->rdmax, c->rcmax, c->rcount, c->stream) != hipSuccess))
        rc = hfem_fail(-2, "hfem_tune_condensation: setup failed");
    float fused_mixed_ms = -1.f;
    for (int w = 0; w < 5 && !rc; w++) {
        const int M = w == 1 ? d.nb + 1 : d.na, N = w == 1 ? d.nb : d.nb + 1;
        const void *inv = prod_mode(w) ? (const void *)It32 : (const void *)It;
        int bv = -2;
        float bt = 1e30f;
        for (int v = 0; v < bgemm_num_variants(); v++) {
            if (!bgemm_fits(prod_mode(w), v, M, N, (size_t)c->maxlds)) continue;
            const float t = time_it([&]() { return run_prod(c, w, v, 0, ns, inv, c->Fx, c->D, c->K, Xt, Ht, R2); });
            if (t >= 0 && t < bt) { bt = t; bv = v; }
        }
        if ((w == 0 || w == 2) && xsolve_fits(w == 0 ? 0 : 1, d.na, (size_t)c->maxlds)) {
            // the fused solve; for the mixed path it runs X0 and one sweep, so it
            // competes with X0 alone here and is weighed against the separate
            // sweep products below
            const float t = time_it([&]() { return run_prod(c, w, kFused, 0, ns, inv, c->Fx, c->D, c->K, Xt, Ht, R2); });
            if (w == 0 && t >= 0 && t < bt) { bt = t; bv = kFused; }
            if (w == 2) fused_mixed_ms = t;
        }
        const int rw = prod_rocblas_index(w);
        if (rw >= 0) {
            // rocBLAS: its own choice, then each listed solution
            const CondGemm g = tuned_gemm(c, rw, c->Fx, c->D, It, c->K, Xt, Ht, R2, ns);
            std::vector<rocblas_int> sols{0};
            std::vector<rocblas_int> more;
            if (cond_gemm_solutions(c->blas, g, more) == rocblas_status_success) sols.insert(sols.end(), more.begin(), more.end());
            for (rocblas_int sol : sols) {
                const float t = time_it([&]() {
                    return cond_gemm_run(c->blas, g, sol) == rocblas_status_success ? hipSuccess : hipErrorLaunchFailure;
                });
                if (t >= 0 && t < bt) { bt = t; bv = -1; c->gemm_sol[rw] = sol; }
            }
        }
        if (bv == -2) { rc = hfem_fail(-2, "hfem_tune_condensation: no launch candidate runs product %d", w); break; }
        c->prod_var[w] = bv;
        best[2 + w] = bt;
    }
    // the mixed solve: the fused kernel (X0 + one sweep) against X0, residual and
    // correction as separate products
    if (!rc && fused_mixed_ms >= 0 && fused_mixed_ms < best[4] + best[5] + best[6]) {
        c->prod_var[2] = kFused;
        best[4] = fused_mixed_ms;
    }
    (void)hipStreamSynchronize(c->stream);
    if (!rc) {
        if (choice) {
            choice[0] = c->inv_var[0];
            choice[1] = c->inv_var[1];
            for (int w = 0; w < 5; w++) choice[2 + w] = c->prod_var[w];
        }
        if (ms)
            for (int i = 0; i < 7; i++) ms[i] = best[i];
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    for (double *p : {Xt, Rt, Ht, It, R2})
        if (p) (void)hipFree(p);
    (void)hipGetLastError();
    return rc;
}

// Number of hand-written product variants (choices 0 .. n-1 of each product).
extern "C" int hfem_condense_variants(void) { return bgemm_num_variants(); }

// Sets the launch choices in the tuner's layout (e.g. a pick kept from an earlier
// tuning), after checking that each fits this context's shapes.
extern "C" int hfem_set_condensation_choice(hfem_ctx *c, const int *choice)
{
    if (!c || !choice) return hfem_fail(-1, "hfem_set_condensation_choice: null argument");
    const Dims &d = c->d;
    for (int p = 0; p < 2; p++)
        if (choice[p] < 0 || choice[p] >= kDinvVariants)
            return hfem_fail(-1, "hfem_set_condensation_choice: inverse variant %d out of range", choice[p]);
    for (int w = 0; w < 5; w++) {
        const int v = choice[2 + w], M = w == 1 ? d.nb + 1 : d.na, N = w == 1 ? d.nb : d.nb + 1;
        if (v == -1 && prod_rocblas_index(w) >= 0) continue;
        if (v == kFused && (w == 0 || w == 2) && xsolve_fits(w == 0 ? 0 : 1, d.na, (size_t)c->maxlds)) continue;
        if (!bgemm_fits(prod_mode(w), v, M, N, (size_t)c->maxlds))
            return hfem_fail(-1, "hfem_set_condensation_choice: product %d cannot run variant %d", w, v);
    }
    c->inv_var[0] = choice[0];
    c->inv_var[1] = choice[1];
    for (int w = 0; w < 5; w++) c->prod_var[w] = choice[2 + w];
    return 0;
}

// The two products of the FP64 condensation of [e0, e0+n) from inv(D) in dinv:
// X = inv(D)[F Ru], [AH FH] = [H Rh] - K X (the context's tuned shapes).
int condense_products_fp64(hfem_ctx *c, int e0, int n)
{
    HIPCHK(run_prod(c, 0, c->prod_var[0], e0, n, c->dinv, c->Fx, c->D, c->K, c->X, c->Hx, nullptr, c->gemm_sol[0]));
    HIPCHK(run_prod(c, 1, c->prod_var[1], e0, n, c->dinv, c->Fx, c->D, c->K, c->X, c->Hx, nullptr, c->gemm_sol[1]));
    return 0;
}

// The FP64 condensation of [e0, e0+n): inv(D), X = inv(D)[F Ru], [AH FH] = [H Rh] - K X.
static int condense_fp64(hfem_ctx *c, int e0, int n)
{
    const Dims &d = c->d;
    HIPCHK(launch_dinv(d.na, e0, n, c->D, c->dinv, c->info, c->stream, c->inv_var[0]));
    return condense_products_fp64(c, e0, n);
}

extern "C" int hfem_condense(hfem_ctx *c, int e0, int n, int precision)
{
    int rc = check_range(c, e0, n, "hfem_condense");
    if (rc) return rc;
    if (precision != HFEM_FP64 && precision != HFEM_MIXED) return hfem_fail(-1, "hfem_condense: unknown precision %d", precision);
    if (n == 0) return 0;
    HIPCHK(hipSetDevice(c->device));
    const Dims &d = c->d;
    if (!c->capturing) HIPCHK(hipEventRecord(c->ev[2], c->stream));
    c->last_sweeps = 0;
    c->last_unconverged = 0;
    c->last_fallback = 0;
    if (precision == HFEM_FP64) {
        if ((rc = condense_fp64(c, e0, n))) return rc;
    } else {
        // FP32 MFMA Gauss-Jordan inv32, X0 = inv32 [F Ru] on the FP32 MFMA, then
        // sweeps R = [F Ru] - D X (FP64) and X += inv32 R (FP32 products, FP64
        // update) until every element's estimated remaining error is below
        // mix_tol |X| (at most mix_max_sweeps), then the FP64 Schur product.
        // Elements that do not converge send the range through the FP64 path.
        if (!c->rwork && c->prod_var[2] != kFused) HIPCHK(hfem_dalloc(&c->rwork, (size_t)c->hd.ne * d.na * c->ldx));
        float *inv32 = (float *)c->dinv;
        HIPCHK(launch_dinv32(d.na, e0, n, c->D, inv32, c->info, c->stream, c->inv_var[1]));
        const bool fused = c->prod_var[2] == kFused;
        if (!fused) HIPCHK(run_prod(c, 2, c->prod_var[2], e0, n, inv32, c->Fx, c->D, c->K, c->X, c->Hx, c->rwork));
        int left = n;
        for (int sweep = 1; sweep <= c->mix_max_sweeps && left > 0; sweep++) {
            HIPCHK(launch_refine_reset(n, c->rdmax + e0, c->rcmax + e0, c->rcount, c->stream));
            if (fused) {
                // X0 and the first sweep, or one more sweep from the current X
                HIPCHK(launch_xsolve(1, d.na, d.nb + 1, c->ldx, e0, n, c->D, inv32, c->Fx, c->X, 1, sweep > 1,
                                     c->rdmax + e0, c->rcmax + e0, c->stream));
            } else {
                HIPCHK(run_prod(c, 3, c->prod_var[3], e0, n, inv32, c->Fx, c->D, c->K, c->X, c->Hx, c->rwork,
                                c->gemm_sol[2]));
                HIPCHK(run_prod(c, 4, c->prod_var[4], e0, n, inv32, c->Fx, c->D, c->K, c->X, c->Hx, c->rwork));
            }
            HIPCHK(launch_refine_count(n, c->rdmax + e0, c->rcmax + e0, c->rdprev + e0, sweep == 1, c->mix_tol,
                                       c->rcount, c->stream));
            HIPCHK(hipMemcpyAsync(&left, c->rcount, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            c->last_sweeps = sweep;
        }
        c->last_unconverged = left;
        if (left > 0) {
            c->last_fallback = 1;
            if ((rc = condense_fp64(c, e0, n))) return rc;
        } else {
            HIPCHK(run_prod(c, 1, c->prod_var[1], e0, n, nullptr, c->Fx, c->D, c->K, c->X, c->Hx, nullptr,
                            c->gemm_sol[1]));
        }
    }
    if (!c->capturing) HIPCHK(hipEventRecord(c->ev[3], c->stream));
    return 0;
}

// Refinement report of the last hfem_condense: sweeps run, elements that did not
// meet the tolerance, whether the range was redone in FP64 (mixed path only).
extern "C" int hfem_condense_report(hfem_ctx *c, int *sweeps, int *unconverged, int *fallback)
{
    if (!c) return hfem_fail(-1, "hfem_condense_report: null context");
    if (sweeps) *sweeps = c->last_sweeps;
    if (unconverged) *unconverged = c->last_unconverged;
    if (fallback) *fallback = c->last_fallback;
    return 0;
}

// Mixed-path refinement parameters: tolerance on the estimated remaining error
// relative to max |X| per element, and the sweep cap.
extern "C" int hfem_set_refinement(hfem_ctx *c, double tol, int max_sweeps)
{
    if (!c) return hfem_fail(-1, "hfem_set_refinement: null context");
    if (!(tol > 0.0) || max_sweeps < 1 || max_sweeps > 16)
        return hfem_fail(-1, "hfem_set_refinement: need tol > 0 and 1 <= max_sweeps <= 16");
    c->mix_tol = tol;
    c->mix_max_sweeps = max_sweeps;
    return 0;
}

extern "C" int hfem_last_kernel_ms(hfem_ctx *c, float *ams, float *cms)
{
    if (!c) return hfem_fail(-1, "hfem_last_kernel_ms: null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (ams) { *ams = 0; if (hipEventElapsedTime(ams, c->ev[0], c->ev[1]) != hipSuccess) *ams = -1; }
    if (cms) { *cms = 0; if (hipEventElapsedTime(cms, c->ev[2], c->ev[3]) != hipSuccess) *cms = -1; }
    (void)hipGetLastError();
    return 0;
}

extern "C" int hfem_sync(hfem_ctx *c)
{
    if (!c) return hfem_fail(-1, "hfem_sync: null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// Copy the element blocks of [e0, e0+n) to the host in the reference layouts.
// Fx (row-major, F then Ru) is transposed back to column-major F; Hx gives H
// (first nb columns) -- before condensation -- or AH after it.
static int read_rowmajor(hfem_ctx *c, const double *src, int e0, int n, double *colmajor, double *lastcol)
{
    const Dims &d = c->d;
    const long na = d.na, nb = d.nb, ldx = c->ldx;
    if (!colmajor) {
        // the last column alone: one strided copy
        if (lastcol && n > 0)
            HIPCHK(hipMemcpy2D(lastcol, sizeof(double), src + (long)e0 * na * ldx + nb, sizeof(double) * ldx,
                               sizeof(double), (size_t)n * na, hipMemcpyDeviceToHost));
        return 0;
    }
    std::vector<double> tmp((size_t)n * na * ldx);
    HIPCHK(hipMemcpy(tmp.data(), src + (long)e0 * na * ldx, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
    for (long t = 0; t < n; t++)
        for (long i = 0; i < na; i++) {
            const double *row = tmp.data() + (t * na + i) * ldx;
            if (colmajor)
                for (long j = 0; j < nb; j++) colmajor[t * na * nb + i + na * j] = row[j];
            if (lastcol) lastcol[t * na + i] = row[nb];
        }
    return 0;
}

static int read_hx(hfem_ctx *c, int e0, int n, double *H, double *lastcol)
{
    const Dims &d = c->d;
    const long nb = d.nb, es = nb * (nb + 1);
    for (long t = 0; t < n; t++) {
        const double *src = c->Hx + (e0 + t) * es;
        if (H) HIPCHK(hipMemcpy(H + t * nb * nb, src, sizeof(double) * nb * nb, hipMemcpyDeviceToHost));
        if (lastcol) HIPCHK(hipMemcpy(lastcol + t * nb, src + nb * nb, sizeof(double) * nb, hipMemcpyDeviceToHost));
    }
    return 0;
}

extern "C" int hfem_read_elements(hfem_ctx *c, int e0, int n, double *Ru, double *D, double *F, double *K, double *H,
                                  double *Rh)
{
    int rc = check_range(c, e0, n, "hfem_read_elements");
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const Dims &d = c->d;
    const long na = d.na, nb = d.nb;
    if (Ru) HIPCHK(hipMemcpy(Ru, c->Ru + e0 * na, sizeof(double) * n * na, hipMemcpyDeviceToHost));
    if (D) HIPCHK(hipMemcpy(D, c->D + e0 * na * na, sizeof(double) * n * na * na, hipMemcpyDeviceToHost));
    if (F && (rc = read_rowmajor(c, c->Fx, e0, n, F, nullptr))) return rc;
    if (K) HIPCHK(hipMemcpy(K, c->K + e0 * nb * na, sizeof(double) * n * nb * na, hipMemcpyDeviceToHost));
    if (H && (rc = read_hx(c, e0, n, H, nullptr))) return rc;
    if (Rh) HIPCHK(hipMemcpy(Rh, c->Rh + e0 * nb, sizeof(double) * n * nb, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int hfem_read_condensed(hfem_ctx *c, int e0, int n, double *DinvRu, double *DinvF, double *AH, double *FH,
                                   int *info)
{
    int rc = check_range(c, e0, n, "hfem_read_condensed");
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    if ((DinvRu || DinvF) && (rc = read_rowmajor(c, c->X, e0, n, DinvF, DinvRu))) return rc;
    if ((AH || FH) && (rc = read_hx(c, e0, n, AH, FH))) return rc;
    if (info) HIPCHK(hipMemcpy(info, c->info + e0, sizeof(int) * n, hipMemcpyDeviceToHost));
    return 0;
}

int hfem_reduce_sum(hfem_ctx *c, long n, const double *x, int square, double *out)
{
    const int nblk = 256;
    HIPCHK(launch_sumsq(n, x, square, c->partial, nblk, c->stream));
    HIPCHK(launch_sumsq(nblk, c->partial, 0, c->partial + nblk, 1, c->stream));
    HIPCHK(hipMemcpyAsync(out, c->partial + nblk, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// updateQuasiNewtonStructures (assembly/globalassembly.cpp:327-389): keep inv(D) and K
// of this Jacobian, in float (quasiNewtonAccuracy 0) or double (1).  The reference
// keeps the LU factors of D and solves with them; inv(D) is what this build's
// condensation forms, so it is kept instead.
static int store_quasi_newton(hfem_ctx *c, int precision)
{
    const Dims &d = c->d;
    const int ne = c->hd.ne;
    const long nD = (long)ne * d.na * d.na, nK = (long)ne * d.nb * d.na;
    // the mixed condensation leaves the FP32 inverse in dinv: the stored one is FP64
    if (precision != HFEM_FP64) HIPCHK(launch_dinv(d.na, 0, ne, c->D, c->dinv, c->info, c->stream));
    if (c->qn_acc == 1) {
        HIPCHK(hipMemcpyAsync(c->qn_dinv, c->dinv, nD * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->qn_K, c->K, nK * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    } else {
        HIPCHK(launch_to_f32(nD, c->dinv, (float *)c->qn_dinv, c->stream));
        HIPCHK(launch_to_f32(nK, c->K, (float *)c->qn_K, c->stream));
    }
    c->qn_valid = true;
    return 0;
}

extern "C" int hfem_quasi_newton_storage(hfem_ctx *c, int accuracy)
{
    if (!c) return hfem_fail(-1, "hfem_quasi_newton_storage: null context");
    if (accuracy < -1 || accuracy > 1) return hfem_fail(-1, "hfem_quasi_newton_storage: accuracy %d", accuracy);
    if (accuracy == c->qn_acc) return 0;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->qn_dinv) HIPCHK(hipFree(c->qn_dinv));
    if (c->qn_K) HIPCHK(hipFree(c->qn_K));
    c->qn_dinv = c->qn_K = nullptr;
    c->qn_valid = false;
    c->qn_acc = accuracy;
    if (accuracy < 0) return 0;
    const Dims &d = c->d;
    const size_t elt = accuracy == 1 ? sizeof(double) : sizeof(float);
    const size_t nD = (size_t)c->hd.ne * d.na * d.na, nK = (size_t)c->hd.ne * d.nb * d.na;
    if (hipMalloc(&c->qn_dinv, nD * elt) != hipSuccess || hipMalloc(&c->qn_K, nK * elt) != hipSuccess) {
        if (c->qn_dinv) hipFree(c->qn_dinv);
        c->qn_dinv = c->qn_K = nullptr;
        c->qn_acc = -1;
        return hfem_fail(-2, "hfem_quasi_newton_storage: out of device memory (%zu bytes)", (nD + nK) * elt);
    }
    return 0;
}

// assembleRHS (assembly/globalassembly.cpp:559-616): element vectors at the current
// state, condensed with the stored inv(D) and K (schurElementVector); Rg and r
// gathered; inv(D) Ru replaces the Ru column of X, whose inv(D) F columns stay
// those of the stored Jacobian (what updateSol then combines).
extern "C" int hfem_assemble_rhs(hfem_ctx *c, double *rNorm)
{
    if (!c) return hfem_fail(-1, "hfem_assemble_rhs: null context");
    if (c->qn_acc < 0 || !c->qn_valid)
        return hfem_fail(-3, "hfem_assemble_rhs: no stored Jacobian (enable hfem_quasi_newton_storage, then "
                             "hfem_assemble_linear_system)");
    const int ne = c->hd.ne, nf = c->hd.nf;
    const Dims &d = c->d;
    int rc = hfem_assemble_vector(c, 0, ne);
    if (rc) return rc;
    const int bsz = d.nch * d.npf;
    HIPCHK(launch_face_gather(nf, bsz, d.nch, d.npf, d.nb, d.nb, c->f2e, c->einv, c->Rh, c->r, nullptr, nullptr,
                              c->stream));
    double s1 = 0, s2 = 0;
    if ((rc = hfem_reduce_sum(c, ne, c->ru2, 0, &s1))) return rc;
    if ((rc = hfem_reduce_sum(c, (long)nf * bsz, c->r, 1, &s2))) return rc;
    if (rNorm) *rNorm = sqrt(s1 + s2);
    HIPCHK(launch_schur_vector(d.na, d.nb, c->ldx, ne, c->qn_acc == 0, c->qn_dinv, c->qn_K, c->Ru, c->Rh, c->X,
                               c->Hx, c->stream));
    const long es = (long)d.nb * (d.nb + 1);
    HIPCHK(launch_face_gather(nf, bsz, d.nch, d.npf, d.nb, es, c->f2e, c->einv, c->Hx + (long)d.nb * d.nb, c->Rg,
                              nullptr, nullptr, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// The mixed condensation of all elements with a fixed number of refinement
// sweeps and no host synchronisation: FP32 inverse, X0, `sweeps` sweeps (the
// unconverged count after sweep s left in rcount[1 + s]), the FP64 Schur
// product.  The caller reads the counts afterwards: it redoes the condensation
// kernel by kernel if an element did not converge after the last sweep, and
// runs fewer sweeps next time if an earlier one already sufficed.
static int condense_mixed_fixed(hfem_ctx *c, int sweeps)
{
    const Dims &d = c->d;
    const int n = c->hd.ne;
    float *inv32 = (float *)c->dinv;
    HIPCHK(launch_dinv32(d.na, 0, n, c->D, inv32, c->info, c->stream, c->inv_var[1]));
    const bool fused = c->prod_var[2] == kFused;
    if (!fused) HIPCHK(run_prod(c, 2, c->prod_var[2], 0, n, inv32, c->Fx, c->D, c->K, c->X, c->Hx, c->rwork));
    for (int sweep = 1; sweep <= sweeps; sweep++) {
        int *cnt = c->rcount + 1 + sweep - 1;
        HIPCHK(launch_refine_reset(n, c->rdmax, c->rcmax, cnt, c->stream));
        if (fused) {
            HIPCHK(launch_xsolve(1, d.na, d.nb + 1, c->ldx, 0, n, c->D, inv32, c->Fx, c->X, 1, sweep > 1, c->rdmax,
                                 c->rcmax, c->stream));
        } else {
            HIPCHK(run_prod(c, 3, c->prod_var[3], 0, n, inv32, c->Fx, c->D, c->K, c->X, c->Hx, c->rwork,
                            c->gemm_sol[2]));
            HIPCHK(run_prod(c, 4, c->prod_var[4], 0, n, inv32, c->Fx, c->D, c->K, c->X, c->Hx, c->rwork));
        }
        HIPCHK(launch_refine_count(n, c->rdmax, c->rcmax, c->rdprev, sweep == 1, c->mix_tol, cnt, c->stream));
    }
    HIPCHK(run_prod(c, 1, c->prod_var[1], 0, n, nullptr, c->Fx, c->D, c->K, c->X, c->Hx, nullptr, c->gemm_sol[1]));
    c->last_sweeps = sweeps;
    c->last_unconverged = 0;
    c->last_fallback = 0;
    return 0;
}

// The device work of one linearisation: element assembly, the residual norm's
// two sums of squares (left on the device in partial[512] and partial[513]), the
// condensation and the BSR scatter.  The FP64 path, and the mixed path with a
// fixed number of sweeps (mixed_sweeps > 0), have no host synchronisation in
// them, so they can be captured as a graph.
static int linearize(hfem_ctx *c, int precision, int mixed_sweeps = 0)
{
    const int ne = c->hd.ne, nf = c->hd.nf;
    const Dims &d = c->d;
    int rc = hfem_assemble(c, 0, ne);
    if (rc) return rc;
    const int bsz = d.nch * d.npf;
    // r = assembled uncondensed Rh (Rhonly)
    HIPCHK(launch_face_gather(nf, bsz, d.nch, d.npf, d.nb, d.nb, c->f2e, c->einv, c->Rh, c->r, nullptr, nullptr,
                              c->stream));
    HIPCHK(launch_sumsq(ne, c->ru2, 0, c->partial, 256, c->stream));
    HIPCHK(launch_sumsq(256, c->partial, 0, c->partial + 512, 1, c->stream));
    HIPCHK(launch_sumsq((long)nf * bsz, c->r, 1, c->partial + 256, 256, c->stream));
    HIPCHK(launch_sumsq(256, c->partial + 256, 0, c->partial + 513, 1, c->stream));
    if (precision == HFEM_MIXED && mixed_sweeps > 0) {
        if ((rc = condense_mixed_fixed(c, mixed_sweeps))) return rc;
    } else if ((rc = hfem_condense(c, 0, ne, precision))) {
        return rc;
    }
    if (c->qn_acc >= 0 && (rc = store_quasi_newton(c, precision))) return rc;
    const long es = (long)d.nb * (d.nb + 1);
    HIPCHK(launch_bsr_gather(c->nblocks, bsz, d.nch, d.npf, d.nb, es, c->blkptr, c->blklist, c->einv, c->Hx, c->Hg,
                             c->stream));
    HIPCHK(launch_face_gather(nf, bsz, d.nch, d.npf, d.nb, es, c->f2e, c->einv, c->Hx + (long)d.nb * d.nb, c->Rg,
                              nullptr, nullptr, c->stream));
    return 0;
}

static hfem_ctx::LinKey lin_key(const hfem_ctx *c, int precision, int sweeps)
{
    hfem_ctx::LinKey k;
    memset(&k, 0, sizeof k);
    k.ph = c->ph;
    k.precision = precision;
    k.sweeps = sweeps;
    if (precision == HFEM_MIXED) {
        k.inv_var1 = c->inv_var[1];
        k.prod_var2 = c->prod_var[2];
        k.prod_var3 = c->prod_var[3];
        k.prod_var4 = c->prod_var[4];
        k.gemm_sol2 = c->gemm_sol[2];
        k.mix_tol = c->mix_tol;
    }
    k.inv_var0 = c->inv_var[0];
    k.prod_var0 = c->prod_var[0];
    k.prod_var1 = c->prod_var[1];
    k.gemm_sol0 = c->gemm_sol[0];
    k.gemm_sol1 = c->gemm_sol[1];
    const void *ptr[] = {c->shapvt, c->shapvg, c->shapft, c->shapfg, c->perm, c->invperm, c->dgnodes, c->elcon, c->bf,
                         c->UDG, c->UH, c->SH, c->f2e, c->einv, c->blkptr, c->blklist, c->Hg, c->Rg, c->r, c->rwork};
    static_assert(sizeof ptr == sizeof k.ptr, "linearisation key size");
    memcpy(k.ptr, ptr, sizeof ptr);
    return k;
}

static void drop_lin_graph(hfem_ctx *c)
{
    if (c->lin_exec) (void)hipGraphExecDestroy(c->lin_exec);
    c->lin_exec = nullptr;
}

// No counterpart in the reference: whether hfem_assemble_linear_system replays the
// linearisation as a captured graph (1, the default) or launches it kernel by
// kernel (0).  The mixed condensation is captured with the number of sweeps the
// last kernel-by-kernel mixed condensation needed (the first one runs kernel by
// kernel to learn it); a replay in which an element does not converge is redone
// kernel by kernel.  Quasi-Newton storage always launches kernel by kernel.
extern "C" int hfem_set_linearization_graph(hfem_ctx *c, int use)
{
    if (!c) return hfem_fail(-1, "hfem_set_linearization_graph: null context");
    c->lin_graph = use != 0;
    if (!c->lin_graph) drop_lin_graph(c);
    return 0;
}

extern "C" int hfem_assemble_linear_system(hfem_ctx *c, int precision, double *rNorm)
{
    if (!c) return hfem_fail(-1, "hfem_assemble_linear_system: null context");
    if (!c->has_master || !c->has_phys || !c->has_mesh || !c->has_sol)
        return hfem_fail(-3, "hfem_assemble_linear_system: master, physics, mesh and solution must be set first");
    if (precision != HFEM_FP64 && precision != HFEM_MIXED)
        return hfem_fail(-1, "hfem_assemble_linear_system: unknown precision %d", precision);
    HIPCHK(hipSetDevice(c->device));
    int rc;
    const int sweeps = precision == HFEM_MIXED ? c->lin_mixed_sweeps : 0;
    const bool graph = c->lin_graph && c->qn_acc < 0 && (precision == HFEM_FP64 || sweeps > 0);
    if (graph) {
        if (precision == HFEM_MIXED && !c->rwork && c->prod_var[2] != kFused)
            HIPCHK(hfem_dalloc(&c->rwork, (size_t)c->hd.ne * c->d.na * c->ldx));
        const hfem_ctx::LinKey k = lin_key(c, precision, sweeps);
        if (c->lin_exec && memcmp(&k, &c->lin_key, sizeof k) != 0) drop_lin_graph(c);
        if (!c->lin_exec) {
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
            c->capturing = true;
            rc = linearize(c, precision, sweeps);
            c->capturing = false;
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            HIPCHK(e);
            e = hipGraphInstantiate(&c->lin_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHK(e);
            c->lin_key = k;
        }
        HIPCHK(hipGraphLaunch(c->lin_exec, c->stream));
        if (precision == HFEM_MIXED) {
            int cnt[16] = {0};
            HIPCHK(hipMemcpyAsync(cnt, c->rcount + 1, sizeof(int) * sweeps, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            if (cnt[sweeps - 1] > 0) {
                // an element needs more sweeps than the graph runs: redo kernel by kernel
                // (the graph's Schur product has overwritten H), and learn the new count
                if ((rc = linearize(c, precision))) return rc;
                c->lin_mixed_sweeps = c->last_fallback ? 0 : c->last_sweeps;
            } else {
                // the first sweep after which every element had converged: the next
                // capture runs that many
                int s1 = sweeps;
                while (s1 > 1 && cnt[s1 - 2] == 0) s1--;
                c->lin_mixed_sweeps = s1;
            }
        }
    } else {
        if ((rc = linearize(c, precision))) return rc;
        if (precision == HFEM_MIXED) c->lin_mixed_sweeps = c->last_fallback ? 0 : c->last_sweeps;
    }
    double s[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(s, c->partial + 512, sizeof s, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (rNorm) *rNorm = sqrt(s[0] + s[1]);
    return 0;
}

extern "C" int hfem_read_system(hfem_ctx *c, double *Hg, double *Rg)
{
    if (!c || !c->has_mesh) return hfem_fail(-3, "hfem_read_system: mesh not set");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const long bsz = (long)c->d.nch * c->d.npf;
    if (Hg) HIPCHK(hipMemcpy(Hg, c->Hg, sizeof(double) * c->nblocks * bsz * bsz, hipMemcpyDeviceToHost));
    if (Rg) HIPCHK(hipMemcpy(Rg, c->Rg, sizeof(double) * c->hd.nf * bsz, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int hfem_bsr_matvec(hfem_ctx *c, const double *x, double *y)
{
    if (!c || !c->has_mesh || !x || !y) return hfem_fail(-1, "hfem_bsr_matvec: bad argument or mesh not set");
    HIPCHK(hipSetDevice(c->device));
    const long bsz = (long)c->d.nch * c->d.npf, n = c->hd.nf * bsz;
    HIPCHK(hipMemcpyAsync(c->xv, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_bsr_matvec(c->hd.nf, (int)bsz, c->rowstart, c->cols, c->Hg, c->xv, c->yv, c->stream));
    HIPCHK(hipMemcpyAsync(y, c->yv, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}
