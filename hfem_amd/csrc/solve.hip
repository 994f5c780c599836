// Host side of the trace-system solver and of the Newton / DIRK drivers
// (declared in include/hfem_amd.h).  The reference binds GMRES, the
// preconditioner and the ordering from closed libraries (solvers/dgsolvers.cpp:16-20);
// their semantics here follow its MATLAB specifications (kernel/gmres_bilu0.m,
// computeBILU0.m, applyBILU0.m, computeMDFordering.m, mkBlockJacobi.m).  The
// drivers follow solveNonlinearProblem and solveUnsteadyProblem
// (solvers/dgsolvers.cpp:142-205, :221-444) with status codes in place of exit().
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <set>
#include <utility>
#include <vector>

#include "ctx.h"

#define RC(x)                  \
    do {                       \
        int rc_ = (x);         \
        if (rc_) return rc_;   \
    } while (0)

static double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static long trace_len(hfem_ctx *c) { return (long)c->hd.nf * c->d.nch * c->d.npf; }
static int trace_bsz(hfem_ctx *c) { return c->d.nch * c->d.npf; }

// Vectors and work arrays of the solver, allocated on first use.
static int ensure_vectors(hfem_ctx *c, int restart)
{
    SolverState &s = c->sv;
    const long N = trace_len(c);
    if (!s.x) HIPCHK(hfem_dalloc(&s.x, N));
    if (!s.b) HIPCHK(hfem_dalloc(&s.b, N));
    if (!s.t) HIPCHK(hfem_dalloc(&s.t, N));
    if (!s.d_info) HIPCHK(hfem_dalloc(&s.d_info, (size_t)c->hd.nf));
    for (auto &e : s.ev)
        if (!e) HIPCHK(hipEventCreate(&e));
    if (restart + 1 > s.vcols) {
        if (s.V) HIPCHK(hipFree(s.V));
        if (s.dots) HIPCHK(hipFree(s.dots));
        if (s.coef) HIPCHK(hipFree(s.coef));
        s.V = s.dots = s.coef = nullptr;
        s.vcols = restart + 1;
        HIPCHK(hfem_dalloc(&s.V, (size_t)N * s.vcols));
        HIPCHK(hfem_dalloc(&s.dots, (size_t)multidot_blocks(N) * (s.vcols + 1)));
        HIPCHK(hfem_dalloc(&s.coef, (size_t)s.vcols + 1));
    }
    return 0;
}

static int require_system(hfem_ctx *c, const char *who)
{
    if (!c) return hfem_fail(-1, "%s: null context", who);
    if (!c->has_mesh) return hfem_fail(-3, "%s: mesh not set", who);
    return 0;
}

// ---------------------------------------------------------------- ordering

// Level schedules of the two BILU0 sweeps for the current ordering: a row's
// forward level is one more than the largest forward level of its neighbours
// numbered before it, its backward level likewise over the neighbours numbered
// after it.  Rows of one level never depend on each other.
static int build_levels(hfem_ctx *c)
{
    SolverState &s = c->sv;
    const int n = c->hd.nf;
    const std::vector<int> &rp = s.rowpts, &ci = s.colind;
    std::vector<int> lv(n, 0), uv(n, 0);
    int nl = 0, nu = 0;
    for (int j = 0; j < n; j++) {
        const int r = s.o2u[j];
        int l = 0;
        for (int k = rp[r] + 1; k < rp[r + 1]; k++)
            if (s.u2o[ci[k]] < j) l = std::max(l, lv[ci[k]] + 1);
        lv[r] = l;
        nl = std::max(nl, l + 1);
    }
    for (int j = n - 1; j >= 0; j--) {
        const int r = s.o2u[j];
        int l = 0;
        for (int k = rp[r] + 1; k < rp[r + 1]; k++)
            if (s.u2o[ci[k]] > j) l = std::max(l, uv[ci[k]] + 1);
        uv[r] = l;
        nu = std::max(nu, l + 1);
    }
    auto bucket = [&](const std::vector<int> &lev, int nlev, std::vector<int> &ptr, std::vector<int> &rows) {
        ptr.assign(nlev + 1, 0);
        for (int r = 0; r < n; r++) ptr[lev[r] + 1]++;
        for (int l = 0; l < nlev; l++) ptr[l + 1] += ptr[l];
        rows.assign(n, 0);
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (int j = 0; j < n; j++) {                 // rows of a level in ordering sequence
            const int r = s.o2u[j];
            rows[fill[lev[r]]++] = r;
        }
    };
    std::vector<int> lrows, urows;
    bucket(lv, nl, s.lptr, lrows);
    bucket(uv, nu, s.uptr, urows);
    if (!s.d_lrows) HIPCHK(hfem_dalloc(&s.d_lrows, (size_t)n));
    if (!s.d_urows) HIPCHK(hfem_dalloc(&s.d_urows, (size_t)n));
    if (!s.d_o2u) HIPCHK(hfem_dalloc(&s.d_o2u, (size_t)n));
    if (!s.d_u2o) HIPCHK(hfem_dalloc(&s.d_u2o, (size_t)n));
    HIPCHK(hipMemcpy(s.d_lrows, lrows.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s.d_urows, urows.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s.d_o2u, s.o2u.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s.d_u2o, s.u2o.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    s.prec = -2;                                       // factors belong to the previous ordering
    return 0;
}

// Approximate minimum discarded fill (kernel/computeMDFordering.m): weight of row k
// = Frobenius norm of C(ik) C(kj) over pairs of distinct unnumbered neighbours
// ri, rj of k that are not neighbours of each other.
static double mdf_weight(const std::vector<int> &rp, const std::vector<int> &ci, const std::vector<double> &Cw,
                         const std::vector<char> &numbered, int k)
{
    double s = 0.0;
    for (int i = rp[k] + 1; i < rp[k + 1]; i++) {
        const int ri = ci[i];
        if (numbered[ri]) continue;
        int ik = -1;
        for (int m = rp[ri]; m < rp[ri + 1]; m++)
            if (ci[m] == k) { ik = m; break; }
        for (int j = rp[k] + 1; j < rp[k + 1]; j++) {
            if (j == i) continue;
            const int rj = ci[j];
            if (numbered[rj]) continue;
            bool adj = false;
            for (int m = rp[ri]; m < rp[ri + 1]; m++)
                if (ci[m] == rj) { adj = true; break; }
            if (adj) continue;
            const double dc = Cw[ik] * Cw[j];
            s += dc * dc;
        }
    }
    return std::sqrt(s);
}

extern "C" int hfem_compute_ordering(hfem_ctx *c, int method)
{
    RC(require_system(c, "hfem_compute_ordering"));
    if (method != 0 && method != 1) return hfem_fail(-1, "hfem_compute_ordering: method %d (0 natural, 1 MDF)", method);
    HIPCHK(hipSetDevice(c->device));
    const double t0 = now_ms();
    SolverState &s = c->sv;
    const int n = c->hd.nf, bsz = trace_bsz(c);
    s.o2u.resize(n);
    s.u2o.resize(n);
    if (method == 0) {
        for (int i = 0; i < n; i++) s.o2u[i] = s.u2o[i] = i;
    } else {
        RC(ensure_vectors(c, 1));
        const long b2 = (long)bsz * bsz;
        if (!s.Dinv) HIPCHK(hfem_dalloc(&s.Dinv, (size_t)n * b2));
        if (!s.Cw) HIPCHK(hfem_dalloc(&s.Cw, (size_t)c->nblocks));
        HIPCHK(launch_block_inverse(n, bsz, c->rowstart, c->Hg, s.Dinv, s.d_info, c->stream));
        HIPCHK(launch_mdf_weights(n, bsz, c->rowstart, c->Hg, s.Dinv, s.Cw, c->stream));
        std::vector<double> Cw(c->nblocks);
        std::vector<int> info(n);
        HIPCHK(hipMemcpyAsync(Cw.data(), s.Cw, sizeof(double) * Cw.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(info.data(), s.d_info, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < n; i++)
            if (info[i]) return hfem_fail(-4, "hfem_compute_ordering: diagonal block of row %d is singular", i);
        const std::vector<int> &rp = s.rowpts, &ci = s.colind;
        std::vector<char> numbered(n, 0);
        std::vector<double> w(n);
        // (weight, row): the first element is the minimum weight with the lowest row
        // index among ties, as MATLAB's min (computeMDFordering.m:93)
        std::set<std::pair<double, int>> heap;
        for (int k = 0; k < n; k++) {
            w[k] = mdf_weight(rp, ci, Cw, numbered, k);
            heap.insert({w[k], k});
        }
        for (int l = 0; l < n; l++) {
            const int p = heap.begin()->second;
            heap.erase(heap.begin());
            s.o2u[l] = p;
            numbered[p] = 1;
            for (int e = rp[p] + 1; e < rp[p + 1]; e++) {
                const int k = ci[e];
                if (numbered[k]) continue;
                heap.erase({w[k], k});
                w[k] = mdf_weight(rp, ci, Cw, numbered, k);
                heap.insert({w[k], k});
            }
        }
        for (int l = 0; l < n; l++) s.u2o[s.o2u[l]] = l;
    }
    s.order_method = method;
    RC(build_levels(c));
    s.ms_order = now_ms() - t0;
    return 0;
}

extern "C" int hfem_set_ordering(hfem_ctx *c, const int *o2u)
{
    RC(require_system(c, "hfem_set_ordering"));
    if (!o2u) return hfem_fail(-1, "hfem_set_ordering: null ordering");
    HIPCHK(hipSetDevice(c->device));
    SolverState &s = c->sv;
    const int n = c->hd.nf;
    std::vector<int> u2o(n, -1);
    for (int l = 0; l < n; l++) {
        if (o2u[l] < 0 || o2u[l] >= n || u2o[o2u[l]] >= 0)
            return hfem_fail(-1, "hfem_set_ordering: not a permutation of 0..%d", n - 1);
        u2o[o2u[l]] = l;
    }
    s.o2u.assign(o2u, o2u + n);
    s.u2o = u2o;
    s.order_method = 3;
    return build_levels(c);
}

extern "C" int hfem_get_ordering(hfem_ctx *c, int *o2u, int *nlevels)
{
    RC(require_system(c, "hfem_get_ordering"));
    const SolverState &s = c->sv;
    if (s.order_method < 0) return hfem_fail(-3, "hfem_get_ordering: no ordering computed");
    if (o2u) std::copy(s.o2u.begin(), s.o2u.end(), o2u);
    if (nlevels) {
        nlevels[0] = (int)s.lptr.size() - 1;
        nlevels[1] = (int)s.uptr.size() - 1;
    }
    return 0;
}

// ---------------------------------------------------------------- preconditioner

static int ensure_syncfree(hfem_ctx *c);
static int check_syncfree(hfem_ctx *c, const char *who);

extern "C" int hfem_precond_factor(hfem_ctx *c, int kind)
{
    RC(require_system(c, "hfem_precond_factor"));
    if (kind != -1 && kind != 0 && kind != 2)
        return hfem_fail(-1, "hfem_precond_factor: kind %d (-1 none, 0 BILU0, 2 block-Jacobi)", kind);
    HIPCHK(hipSetDevice(c->device));
    const double t0 = now_ms();
    SolverState &s = c->sv;
    RC(ensure_vectors(c, 1));
    const int n = c->hd.nf, bsz = trace_bsz(c);
    const long b2 = (long)bsz * bsz;
    if (kind == -1) {
        s.prec = -1;
        return 0;
    }
    if ((size_t)(kind == 0 ? bilu0_factor_lds(bsz) + 16 : block_inverse_lds(bsz)) > (size_t)c->maxlds)
        return hfem_fail(-1, "hfem_precond_factor: block size %d needs more LDS than the device offers", bsz);
    const bool fp32 = s.prec_fp32 != 0;
    s.prec = -2;   // until this factorisation has succeeded
    if (!fp32 && !s.Mg) HIPCHK(hfem_dalloc(&s.Mg, (size_t)c->nblocks * b2));
    if (fp32 && !s.Mg32) HIPCHK(hfem_dalloc(&s.Mg32, (size_t)c->nblocks * b2));
    if (kind == 2) {
        // block-Jacobi: inverses in FP64; single-precision factors are their FP32 copies
        if (!fp32) HIPCHK(launch_block_inverse(n, bsz, c->rowstart, c->Hg, s.Mg, s.d_info, c->stream));
        else {
            if (!s.Mg) HIPCHK(hfem_dalloc(&s.Mg, (size_t)c->nblocks * b2));
            HIPCHK(launch_block_inverse(n, bsz, c->rowstart, c->Hg, s.Mg, s.d_info, c->stream));
            HIPCHK(launch_to_float((long)n * b2, s.Mg, s.Mg32, c->stream));
        }
    } else {
        if (s.order_method < 0) RC(hfem_compute_ordering(c, 1));
        // (sync-free: one persistent launch over the rows in forward-level order;
        // else one launch per level)
        if (s.sweep_mode == 1) RC(ensure_syncfree(c));
        unsigned *sync = s.sweep_mode == 1 ? s.sf_flags + 2 * (size_t)n : nullptr;
        if (!fp32) {
            HIPCHK(hipMemcpyAsync(s.Mg, c->Hg, sizeof(double) * c->nblocks * b2, hipMemcpyDeviceToDevice, c->stream));
            if (s.sweep_mode == 1)
                HIPCHK(launch_bilu0_factor_syncfree(bsz, c->rowstart, c->cols, s.d_u2o, s.d_lrows, n, s.Mg, s.d_info,
                                                    s.sf_flags, sync, s.sweep_wgs, s.spin_limit, s.sf_err, c->stream));
            else
                for (size_t l = 0; l + 1 < s.lptr.size(); l++)
                    HIPCHK(launch_bilu0_factor_level(bsz, c->rowstart, c->cols, s.d_u2o, s.d_lrows + s.lptr[l],
                                                     s.lptr[l + 1] - s.lptr[l], s.Mg, s.d_info, c->stream));
        } else {
            // single precision (precPrecision 0): Hg rounded to FP32 and factored on
            // the FP32 MFMA
            HIPCHK(launch_to_float((long)c->nblocks * b2, c->Hg, s.Mg32, c->stream));
            if (s.sweep_mode == 1)
                HIPCHK(launch_bilu0_factor_syncfree(bsz, c->rowstart, c->cols, s.d_u2o, s.d_lrows, n, s.Mg32,
                                                    s.d_info, s.sf_flags, sync, s.sweep_wgs, s.spin_limit, s.sf_err,
                                                    c->stream));
            else
                for (size_t l = 0; l + 1 < s.lptr.size(); l++)
                    HIPCHK(launch_bilu0_factor_level(bsz, c->rowstart, c->cols, s.d_u2o, s.d_lrows + s.lptr[l],
                                                     s.lptr[l + 1] - s.lptr[l], s.Mg32, s.d_info, c->stream));
        }
    }
    if (kind != 2) {
        // the sweeps' form of the upper blocks
        if (!fp32) HIPCHK(launch_bilu0_scale_upper(bsz, n, c->rowstart, c->cols, s.d_u2o, s.Mg, c->stream));
        else HIPCHK(launch_bilu0_scale_upper(bsz, n, c->rowstart, c->cols, s.d_u2o, s.Mg32, c->stream));
    }
    std::vector<int> info(n);
    HIPCHK(hipMemcpyAsync(info.data(), s.d_info, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    RC(check_syncfree(c, "hfem_precond_factor"));
    for (int i = 0; i < n; i++) {
        if (info[i] < 0)
            return hfem_fail(-4, "hfem_precond_factor: row %d has %d blocks, more than the factorisation takes", i,
                             -info[i]);
        if (info[i]) return hfem_fail(-4, "hfem_precond_factor: singular pivot block in row %d", i);
    }
    s.prec = kind;
    s.graph_dirty = true;
    s.ms_factor = now_ms() - t0;
    return 0;
}

// sys.precPrecision (preprocessing/initializestructs.cpp:476-486): 0 single-precision
// preconditioner factors (Mg_sp), 1 double.  Takes effect at the next factorisation.
extern "C" int hfem_set_precond_precision(hfem_ctx *c, int prec_precision)
{
    if (!c) return hfem_fail(-1, "hfem_set_precond_precision: null context");
    if (prec_precision != 0 && prec_precision != 1)
        return hfem_fail(-1, "hfem_set_precond_precision: %d (0 single, 1 double)", prec_precision);
    if ((prec_precision == 0) != (c->sv.prec_fp32 != 0)) c->sv.prec = -2;   // factors of the other precision are stale
    c->sv.prec_fp32 = prec_precision == 0;
    return 0;
}

// Buffers of the sync-free sweeps, allocated and cleared at their first use (before
// any graph capture: the clearing is a memset).
static int ensure_syncfree(hfem_ctx *c)
{
    SolverState &s = c->sv;
    const int n = c->hd.nf;
    if (s.lptr.empty() || s.lptr.back() != n || s.uptr.empty() || s.uptr.back() != n)
        return hfem_fail(-3, "sync-free BILU0 sweeps: the level lists do not cover the %d rows", n);
    if (!s.sf_flags) {
        HIPCHK(hfem_dalloc(&s.sf_flags, (size_t)2 * n + 3));
        HIPCHK(hipMemset(s.sf_flags, 0, sizeof(unsigned) * ((size_t)2 * n + 3)));
    }
    if (!s.sf_err) {
        HIPCHK(hfem_dalloc(&s.sf_err, 1));
        HIPCHK(hipMemset(s.sf_err, 0, sizeof(int)));
    }
    return 0;
}

// Whether a bounded wait of the sync-free sweeps gave up since the last check (the
// applications since then are invalid); clears the flag.  Synchronises the stream.
static int check_syncfree(hfem_ctx *c, const char *who)
{
    SolverState &s = c->sv;
    if (!s.sf_err) return 0;
    int e = 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&e, s.sf_err, sizeof(int), hipMemcpyDeviceToHost));
    if (!e) return 0;
    HIPCHK(hipMemset(s.sf_err, 0, sizeof(int)));
    return hfem_fail(-5, "%s: a wait of the sync-free BILU0 sweeps gave up after %d polls; the preconditioner "
                     "applications are invalid", who, s.spin_limit);
}

// The two BILU0 sweeps: forward into s.t, backward into y; one launch per level
// (sweep_mode 0) or one persistent launch per sweep (sweep_mode 1).
static int bilu0_sweeps(hfem_ctx *c, const double *x, double *y)
{
    SolverState &s = c->sv;
    const int bsz = trace_bsz(c);
    if (s.sweep_mode == 1) {
        const int n = c->hd.nf;
        if (s.prec_fp32)
            HIPCHK(launch_bilu0_syncfree(bsz, c->rowstart, c->cols, s.d_u2o, s.d_lrows, s.d_urows, n, s.Mg32, x, s.t, y,
                                         s.sf_flags, s.sf_flags + 2 * (size_t)n, s.sweep_wgs, s.spin_limit, s.sf_err,
                                         c->stream));
        else
            HIPCHK(launch_bilu0_syncfree(bsz, c->rowstart, c->cols, s.d_u2o, s.d_lrows, s.d_urows, n, s.Mg, x, s.t, y,
                                         s.sf_flags, s.sf_flags + 2 * (size_t)n, s.sweep_wgs, s.spin_limit, s.sf_err,
                                         c->stream));
        return 0;
    }
    for (size_t l = 0; l + 1 < s.lptr.size(); l++) {
        const int *rows = s.d_lrows + s.lptr[l];
        const int nl = s.lptr[l + 1] - s.lptr[l];
        if (s.prec_fp32) HIPCHK(launch_bilu0_lsolve_level(bsz, c->rowstart, c->cols, s.d_u2o, rows, nl, s.Mg32, x, s.t, c->stream));
        else HIPCHK(launch_bilu0_lsolve_level(bsz, c->rowstart, c->cols, s.d_u2o, rows, nl, s.Mg, x, s.t, c->stream));
    }
    for (size_t l = 0; l + 1 < s.uptr.size(); l++) {
        const int *rows = s.d_urows + s.uptr[l];
        const int nl = s.uptr[l + 1] - s.uptr[l];
        if (s.prec_fp32) HIPCHK(launch_bilu0_usolve_level(bsz, c->rowstart, c->cols, s.d_u2o, rows, nl, s.Mg32, s.t, y, c->stream));
        else HIPCHK(launch_bilu0_usolve_level(bsz, c->rowstart, c->cols, s.d_u2o, rows, nl, s.Mg, s.t, y, c->stream));
    }
    return 0;
}

// How the BILU0 sweeps run: mode 1 (default) one persistent launch per sweep whose
// workgroups wait on per-row ready flags, each wait bounded by spin_limit polls (0:
// the default 2^20), a wait that gives up making the next hfem_gmres /
// hfem_precond_apply fail; mode 0 one launch per level.  Same bits either way.  No
// counterpart in the reference.
extern "C" int hfem_set_bilu0_sweep(hfem_ctx *c, int mode, int spin_limit)
{
    if (!c) return hfem_fail(-1, "hfem_set_bilu0_sweep: null context");
    if (mode != 0 && mode != 1) return hfem_fail(-1, "hfem_set_bilu0_sweep: mode %d (0 level launches, 1 sync-free)", mode);
    if (spin_limit < 0) return hfem_fail(-1, "hfem_set_bilu0_sweep: spin_limit %d", spin_limit);
    c->sv.sweep_mode = mode;
    c->sv.spin_limit = spin_limit ? spin_limit : (1 << 20);
    c->sv.graph_dirty = true;
    return 0;
}

// Whether the BILU0 application replays a captured graph (1, default) or launches
// its levels one by one (0).  No counterpart in the reference.
extern "C" int hfem_set_apply_graph(hfem_ctx *c, int use)
{
    if (!c) return hfem_fail(-1, "hfem_set_apply_graph: null context");
    c->sv.use_graph = use != 0;
    c->sv.graph_dirty = true;
    return 0;
}

// y = M^{-1} x on device vectors (x and y distinct).
static int precond_apply_dev(hfem_ctx *c, const double *x, double *y)
{
    SolverState &s = c->sv;
    const int n = c->hd.nf, bsz = trace_bsz(c);
    const long N = trace_len(c);
    if (s.prec == -1) {
        HIPCHK(hipMemcpyAsync(y, x, sizeof(double) * N, hipMemcpyDeviceToDevice, c->stream));
    } else if (s.prec == 2) {
        if (s.prec_fp32) HIPCHK(launch_block_jacobi(n, bsz, s.Mg32, x, y, c->stream));
        else HIPCHK(launch_block_jacobi(n, bsz, s.Mg, x, y, c->stream));
    } else if (s.prec == 0) {
        const long bytes = sizeof(double) * N;
        if (s.sweep_mode == 1) RC(ensure_syncfree(c));
        if (s.use_graph) {
            // the level launches of both sweeps as one graph from c->yv into s.gy
            // (captured after each factorisation; no memset nodes in it)
            if (s.graph_dirty || !s.apply_graph) {
                if (s.apply_graph) HIPCHK(hipGraphExecDestroy(s.apply_graph));
                s.apply_graph = nullptr;
                if (!s.gy) HIPCHK(hfem_dalloc(&s.gy, N));
                hipGraph_t g = nullptr;
                HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                const int rc = bilu0_sweeps(c, c->yv, s.gy);
                hipError_t e = hipStreamEndCapture(c->stream, &g);
                if (rc) {
                    if (g) hipGraphDestroy(g);
                    return rc;
                }
                HIPCHK(e);
                e = hipGraphInstantiate(&s.apply_graph, g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                HIPCHK(e);
                s.graph_dirty = false;
            }
            if (x != c->yv) HIPCHK(hipMemcpyAsync(c->yv, x, bytes, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipGraphLaunch(s.apply_graph, c->stream));
            HIPCHK(hipMemcpyAsync(y, s.gy, bytes, hipMemcpyDeviceToDevice, c->stream));
        } else {
            RC(bilu0_sweeps(c, x, y));
        }
    } else {
        return hfem_fail(-3, "preconditioner not factored");
    }
    return 0;
}

extern "C" int hfem_precond_apply(hfem_ctx *c, const double *x, double *y)
{
    RC(require_system(c, "hfem_precond_apply"));
    if (!x || !y) return hfem_fail(-1, "hfem_precond_apply: null vector");
    HIPCHK(hipSetDevice(c->device));
    const long N = trace_len(c);
    HIPCHK(hipMemcpyAsync(c->xv, x, sizeof(double) * N, hipMemcpyHostToDevice, c->stream));
    RC(precond_apply_dev(c, c->xv, c->yv));
    HIPCHK(hipMemcpyAsync(y, c->yv, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    RC(check_syncfree(c, "hfem_precond_apply"));
    return 0;
}

// ---------------------------------------------------------------- GMRES

static int dots_host(hfem_ctx *c, int k, const double *V, long ldv, const double *w, double *out)
{
    SolverState &s = c->sv;
    const long N = trace_len(c);
    HIPCHK(launch_multidot(N, k, V, ldv, w, s.dots, s.coef, c->stream));
    HIPCHK(hipMemcpyAsync(out, s.coef, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// Least squares min |beta e1 - Hbar y| over the (j+1) x j Hessenberg matrix by
// Givens rotations; returns the residual norm.
static double hessenberg_ls(int j, int ldh, const std::vector<double> &hh, double beta, std::vector<double> &y)
{
    std::vector<double> R(hh.begin(), hh.begin() + (size_t)ldh * j), g(j + 1, 0.0);
    g[0] = beta;
    for (int k = 0; k < j; k++) {
        const double a = R[k + ldh * k], b = R[k + 1 + ldh * k], r = std::hypot(a, b);
        const double cs = r == 0 ? 1.0 : a / r, sn = r == 0 ? 0.0 : b / r;
        for (int col = k; col < j; col++) {
            const double u = R[k + ldh * col], v = R[k + 1 + ldh * col];
            R[k + ldh * col] = cs * u + sn * v;
            R[k + 1 + ldh * col] = -sn * u + cs * v;
        }
        const double u = g[k], v = g[k + 1];
        g[k] = cs * u + sn * v;
        g[k + 1] = -sn * u + cs * v;
    }
    y.assign(j, 0.0);
    for (int k = j - 1; k >= 0; k--) {
        double t = g[k];
        for (int c2 = k + 1; c2 < j; c2++) t -= R[k + ldh * c2] * y[c2];
        y[k] = t / R[k + ldh * k];
    }
    return std::fabs(g[j]);
}

// Left-preconditioned restarted GMRES on Hg x = Rg (kernel/gmres_bilu0.m), x in the
// solver's device vector, zero initial guess.  orthog 0: modified Gram-Schmidt,
// 1: classical Gram-Schmidt applied twice.  flag: 0 converged, 1 iteration limit.
extern "C" int hfem_gmres(hfem_ctx *c, int restart, double tol, int maxit, int orthog, int *iters, double *rev,
                          int *flag)
{
    RC(require_system(c, "hfem_gmres"));
    if (restart < 1 || maxit < 1 || !(tol > 0)) return hfem_fail(-1, "hfem_gmres: restart, maxit and tol must be positive");
    if (orthog != 0 && orthog != 1) return hfem_fail(-1, "hfem_gmres: orthog %d (0 MGS, 1 ICGS)", orthog);
    if (c->sv.prec < -1) return hfem_fail(-3, "hfem_gmres: call hfem_precond_factor first");
    HIPCHK(hipSetDevice(c->device));
    const double t0 = now_ms();
    RC(ensure_vectors(c, restart));
    SolverState &s = c->sv;
    s.ms_matvec = s.ms_prec = 0;
    s.n_matvec = s.n_prec = 0;
    const long N = trace_len(c);
    const int nf = c->hd.nf, bsz = trace_bsz(c);
    auto matvec = [&](const double *x, double *y) -> int {
        HIPCHK(launch_bsr_matvec(nf, bsz, c->rowstart, c->cols, c->Hg, x, y, c->stream));
        return 0;
    };
    RC(precond_apply_dev(c, c->Rg, s.b));                       // b = M \ Rg
    double nb2;
    RC(dots_host(c, 1, s.b, N, s.b, &nb2));
    const double nrmb = std::sqrt(nb2);
    HIPCHK(launch_fill(N, 0.0, s.x, c->stream));
    const int ldh = restart + 1;
    std::vector<double> hh((size_t)ldh * restart), y, hc(restart + 1);
    int it = 0, fl = -1;
    double res = 0.0;
    while (true) {
        double *V0 = s.V;
        RC(matvec(s.x, c->yv));                                  // r = b - M \ (Hg x)
        RC(precond_apply_dev(c, c->yv, V0));
        HIPCHK(launch_axpby(N, 1.0, s.b, -1.0, V0, c->stream));
        double beta2;
        RC(dots_host(c, 1, V0, N, V0, &beta2));
        const double beta = std::sqrt(beta2);
        if (beta == 0.0 || nrmb == 0.0) { fl = 0; res = 0.0; break; }
        HIPCHK(launch_axpby(N, 0.0, V0, 1.0 / beta, V0, c->stream));
        std::fill(hh.begin(), hh.end(), 0.0);
        int jdone = 0;
        for (int j = 0; j < restart; j++) {
            double *vj = s.V + (long)j * N, *w = s.V + (long)(j + 1) * N;
            HIPCHK(hipEventRecord(s.ev[0], c->stream));
            RC(matvec(vj, c->yv));
            HIPCHK(hipEventRecord(s.ev[1], c->stream));
            RC(precond_apply_dev(c, c->yv, w));
            HIPCHK(hipEventRecord(s.ev[2], c->stream));
            if (orthog == 0) {
                for (int i = 0; i <= j; i++) {
                    double h;
                    RC(dots_host(c, 1, s.V + (long)i * N, N, w, &h));
                    hh[i + (size_t)ldh * j] = h;
                    HIPCHK(launch_axpby(N, -h, s.V + (long)i * N, 1.0, w, c->stream));
                }
            } else {
                for (int pass = 0; pass < 2; pass++) {
                    RC(dots_host(c, j + 1, s.V, N, w, hc.data()));   // leaves the coefficients in s.coef
                    for (int i = 0; i <= j; i++) hh[i + (size_t)ldh * j] += hc[i];
                    HIPCHK(launch_multi_axpy(N, j + 1, s.V, N, s.coef, -1.0, w, c->stream));
                }
            }
            double hn2;
            RC(dots_host(c, 1, w, N, w, &hn2));
            const double hn = std::sqrt(hn2);
            float e1 = 0, e2 = 0;
            HIPCHK(hipEventElapsedTime(&e1, s.ev[0], s.ev[1]));
            HIPCHK(hipEventElapsedTime(&e2, s.ev[1], s.ev[2]));
            s.ms_matvec += e1;
            s.ms_prec += e2;
            s.n_matvec++;
            s.n_prec++;
            hh[j + 1 + (size_t)ldh * j] = hn;
            jdone = j + 1;
            if (hn != 0.0) HIPCHK(launch_axpby(N, 0.0, w, 1.0 / hn, w, c->stream));
            res = hessenberg_ls(j + 1, ldh, hh, beta, y);
            it++;
            if (rev) rev[it - 1] = res / nrmb;
            if (hn == 0.0 || res / nrmb <= tol) { fl = 0; break; }
            if (it >= maxit) { fl = 1; break; }
        }
        HIPCHK(hipMemcpyAsync(s.coef, y.data(), sizeof(double) * jdone, hipMemcpyHostToDevice, c->stream));
        HIPCHK(launch_multi_axpy(N, jdone, s.V, N, s.coef, 1.0, s.x, c->stream));   // x += V y
        if (fl >= 0) break;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    RC(check_syncfree(c, "hfem_gmres"));
    if (iters) *iters = it;
    if (flag) *flag = fl;
    s.ms_gmres = now_ms() - t0;
    return 0;
}

extern "C" int hfem_solver_stats(hfem_ctx *c, double *out)
{
    if (!c || !out) return hfem_fail(-1, "hfem_solver_stats: null argument");
    const SolverState &s = c->sv;
    const double v[7] = {s.ms_order, s.ms_factor, s.ms_gmres, s.ms_matvec, (double)s.n_matvec, s.ms_prec,
                         (double)s.n_prec};
    std::copy(v, v + 7, out);
    return 0;
}

extern "C" int hfem_read_trace_solution(hfem_ctx *c, double *x)
{
    RC(require_system(c, "hfem_read_trace_solution"));
    if (!c->sv.x || !x) return hfem_fail(-3, "hfem_read_trace_solution: no solution");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(x, c->sv.x, sizeof(double) * trace_len(c), hipMemcpyDeviceToHost));
    return 0;
}

// ---------------------------------------------------------------- solution update and residual

extern "C" int hfem_update_solution(hfem_ctx *c, double alpha)
{
    RC(require_system(c, "hfem_update_solution"));
    if (!c->has_sol || !c->sv.x) return hfem_fail(-3, "hfem_update_solution: no solution or trace update");
    HIPCHK(hipSetDevice(c->device));
    const Dims &d = c->d;
    // UH += alpha duh (solvers/solutionupdate.cpp:75-77)
    HIPCHK(launch_axpby(trace_len(c), alpha, c->sv.x, 1.0, c->UH, c->stream));
    // getU on every element (:6-67), then getQ (assembly/elementassembly.cpp:1414-1479)
    HIPCHK(launch_update_u(c->hd.ne, d.npv, d.nc, d.nch, d.ndf, d.na, d.nb, c->ldx, c->elcon, c->X, c->sv.x, alpha,
                           c->UDG, c->stream));
    if (c->ph.fc_q != 0.0) {
        Master m{c->shapvt, c->shapvg, c->shapft, c->shapfg, c->perm, c->invperm};
        HIPCHK(launch_getq(c->d, m, c->ph, c->dgnodes, c->UDG, c->UH, c->SH, c->elcon, c->bf, 0, c->hd.ne, c->stream));
    }
    return 0;
}

extern "C" int hfem_get_q(hfem_ctx *c)
{
    RC(require_system(c, "hfem_get_q"));
    if (!c->has_sol || !c->has_master) return hfem_fail(-3, "hfem_get_q: master and solution must be set");
    HIPCHK(hipSetDevice(c->device));
    Master m{c->shapvt, c->shapvg, c->shapft, c->shapfg, c->perm, c->invperm};
    HIPCHK(launch_getq(c->d, m, c->ph, c->dgnodes, c->UDG, c->UH, c->SH, c->elcon, c->bf, 0, c->hd.ne, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hfem_residual_norm(hfem_ctx *c, double *rNorm)
{
    RC(require_system(c, "hfem_residual_norm"));
    const int ne = c->hd.ne, nf = c->hd.nf;
    const Dims &d = c->d;
    RC(hfem_assemble_vector(c, 0, ne));   // residual rows only
    HIPCHK(launch_face_gather(nf, trace_bsz(c), d.nch, d.npf, d.nb, d.nb, c->f2e, c->einv, c->Rh, c->r, nullptr,
                              nullptr, c->stream));
    double s1 = 0, s2 = 0;
    RC(hfem_reduce_sum(c, ne, c->ru2, 0, &s1));
    RC(hfem_reduce_sum(c, trace_len(c), c->r, 1, &s2));
    if (rNorm) *rNorm = std::sqrt(s1 + s2);
    return 0;
}

// ---------------------------------------------------------------- Newton and DIRK

static int check_params(const hfem_solver_params *p)
{
    if (!p) return hfem_fail(-1, "solver parameters missing");
    if (!(p->newton_tol > 0) || p->newton_maxiter < 1 || p->restart < 1 || !(p->gmres_tol > 0) || p->gmres_maxit < 1)
        return hfem_fail(-1, "solver parameters: tolerances and iteration limits must be positive");
    if (p->precond != -1 && p->precond != 0 && p->precond != 2) return hfem_fail(-1, "solver parameters: precond %d", p->precond);
    if (p->prec_precision != 0 && p->prec_precision != 1)
        return hfem_fail(-1, "solver parameters: prec_precision %d (0 single, 1 double)", p->prec_precision);
    if (p->ordering != 0 && p->ordering != 1) return hfem_fail(-1, "solver parameters: ordering %d", p->ordering);
    if (p->orthog != 0 && p->orthog != 1) return hfem_fail(-1, "solver parameters: orthog %d", p->orthog);
    if (p->precision != HFEM_FP64 && p->precision != HFEM_MIXED) return hfem_fail(-1, "solver parameters: precision");
    if (p->quasi_newton && (p->qn_accuracy < 0 || p->qn_accuracy > 1 || !(p->qn_reuse_ratio >= 0)))
        return hfem_fail(-1, "solver parameters: quasi-Newton accuracy %d, reuse ratio", p->qn_accuracy);
    return 0;
}

extern "C" int hfem_solve_steady(hfem_ctx *c, const hfem_solver_params *p, int *converged, int *niter,
                                 double *history)
{
    RC(require_system(c, "hfem_solve_steady"));
    RC(check_params(p));
    const bool qn = p->quasi_newton != 0;
    if (qn) RC(hfem_quasi_newton_storage(c, p->qn_accuracy));
    const double alpha_min = 1.0e-4;
    const int true_max = p->true_newton_maxiter > 0 ? p->true_newton_maxiter : INT_MAX;
    double rNorm = p->newton_tol + 1.0;
    int it = 0, true_it = 0;
    // robustMode: reset at the start of a nonlinear solve; a failed line search goes
    // back to the previous iterate only while it is 0, or with quasi-Newton
    // (solvers/dgsolvers.cpp:114, :175-177)
    int robust = 0;
    // quasi-Newton bookkeeping (sys.linearSolvesClocks, sys.lastAssemblyAndPrecClocks)
    bool reuse = false;
    double lin_time = 0.0, jac_time = 0.0;
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto secs = [](std::chrono::steady_clock::time_point t0) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    auto recompute_jacobian = [&] { reuse = false; lin_time = 0.0; };
    while (rNorm > p->newton_tol && it < p->newton_maxiter && true_it < true_max) {
        it++;
        const bool reused = reuse;
        if (!reused) true_it++;
        double oldNorm;
        auto t0 = now();
        if (!reused) {
            RC(hfem_assemble_linear_system(c, p->precision, &oldNorm));
            jac_time = secs(t0);
        } else {
            RC(hfem_assemble_rhs(c, &oldNorm));
        }
        if (!std::isfinite(oldNorm)) return hfem_fail(-5, "hfem_solve_steady: residual norm is NaN or Inf");
        if (c->sv.order_method < 0 || (p->reuse_ordering == 0) || c->sv.order_method != p->ordering)
            RC(hfem_compute_ordering(c, p->ordering));
        if (!reused) {
            auto t1 = now();
            RC(hfem_set_precond_precision(c, p->prec_precision));
            RC(hfem_precond_factor(c, p->precond));
            jac_time += secs(t1);
        }
        int gi = 0, gf = 0;
        auto t2 = now();
        RC(hfem_gmres(c, p->restart, p->gmres_tol, p->gmres_maxit, p->orthog, &gi, nullptr, &gf));
        lin_time += secs(t2);
        if (history) {
            history[HFEM_HIST * (it - 1)] = oldNorm;
            history[HFEM_HIST * (it - 1) + 2] = reused ? -gi : gi;
        }
        if (qn) {
            if (lin_time < p->qn_reuse_ratio * jac_time) reuse = true;
            else recompute_jacobian();
        }
        if (it == 1 && oldNorm < p->newton_tol) {
            rNorm = oldNorm;
            if (history) { history[1] = rNorm; history[3] = 0.0; history[4] = 0.0; }
            break;
        }
        double alpha = 1.0, applied = 0.0;
        int steps = 0;
        while (true) {
            RC(hfem_update_solution(c, alpha));
            applied += alpha;
            steps++;
            RC(hfem_residual_norm(c, &rNorm));
            if (!std::isfinite(rNorm) || (rNorm > oldNorm && rNorm > p->newton_tol)) {
                const double a = std::fabs(alpha / 2);
                alpha = -a;
                if (qn) recompute_jacobian();
                if (a < alpha_min) {
                    if (robust == 0 || qn) {
                        // back to the previous iterate (dgsolvers.cpp:175-180)
                        robust = 1;
                        alpha = -2.0 * std::fabs(alpha);
                        RC(hfem_update_solution(c, alpha));
                        applied += alpha;
                        steps++;
                    }
                    break;
                }
            } else {
                break;
            }
        }
        if (history) {
            history[HFEM_HIST * (it - 1) + 1] = rNorm;
            history[HFEM_HIST * (it - 1) + 3] = applied;
            history[HFEM_HIST * (it - 1) + 4] = steps;
        }
    }
    if (converged) *converged = rNorm < p->newton_tol;
    if (niter) *niter = it;
    return 0;
}

// DIRK coefficients of solvers/DIRKcoeff.cpp (d = inv(A) column-major, c = b inv(A), t).
static bool dirk_table(int nstage, int order, double *c, double *d, double *t)
{
    struct Tab { int s, o; double c[3], d[9], t[3]; };
    static const Tab tabs[] = {
        {1, 1, {1.0}, {1.0}, {1.0}},
        {1, 2, {2.0}, {2.0}, {0.5}},
        {2, 2, {0.0, 1.0}, {3.414213562373096, -8.242640687119289, 0.0, 3.414213562373096}, {0.292893218813452, 1.0}},
        {2, 3, {1.098076211353316, 0.633974596215561}, {1.267949192431123, 0.928203230275509, 0.0, 1.267949192431123},
         {0.788675134594813, 0.211324865405187}},
        {3, 3, {0.0, 0.0, 1.0},
         {2.294280360323567, -1.484721005721436, -8.556127801733982, 0.0, 2.294280360323568, 3.391748836909791, 0.0,
          0.0, 2.294280360323567},
         {0.435866521500000, 0.717933260750000, 1.0}},
        {3, 4, {0.445622407287714, 1.064177772475912, 0.120614758428183},
         {0.935822227524088, 0.497940606760015, -0.345865915800969, 0.0, 0.935822227524088, 2.867525668568206, 0.0,
          0.0, 0.935822227524088},
         {1.068579021301629, 0.5, -0.068579021301629}},
    };
    for (const Tab &tb : tabs)
        if (tb.s == nstage && tb.o == order) {
            for (int i = 0; i < nstage; i++) { c[i] = tb.c[i]; t[i] = tb.t[i]; }
            for (int i = 0; i < nstage * nstage; i++) d[i] = tb.d[i];
            return true;
        }
    return false;
}

extern "C" int hfem_dirk_step(hfem_ctx *c, double dt, int nstage, int order, const hfem_solver_params *p,
                              int *converged, int *newton_iters)
{
    RC(require_system(c, "hfem_dirk_step"));
    RC(check_params(p));
    double cc[3], d[9], tt[3];
    if (!(dt > 0)) return hfem_fail(-1, "hfem_dirk_step: dt must be positive");
    if (!dirk_table(nstage, order, cc, d, tt)) return hfem_fail(-1, "hfem_dirk_step: DIRK(%d,%d) not implemented", nstage, order);
    HIPCHK(hipSetDevice(c->device));
    const Dims &dm = c->d;
    const long ne = c->hd.ne, nu = ne * dm.npv * dm.nc;
    SolverState &s = c->sv;
    if (!s.UDGi) HIPCHK(hfem_dalloc(&s.UDGi, (size_t)nu));
    HIPCHK(hipMemcpyAsync(s.UDGi, c->UDG, sizeof(double) * nu, hipMemcpyDeviceToDevice, c->stream));
    const Phys saved = c->ph;
    auto D = [&](int i, int j) { return d[i + nstage * j]; };
    int total = 0, ok = 1;
    for (int st = 0; st < nstage; st++) {
        c->ph.fc_u = D(st, st) / dt;
        c->ph.fc_q = 1.0;
        c->ph.tdep = 1;
        // SH of the stage (solvers/dgsolvers.cpp:316-341) and zero q source (:344-349)
        double a, b, cf;
        if (st == 0) { a = D(0, 0) / dt; b = 0.0; cf = 0.0; }
        else if (st == 1) { a = (D(1, 0) + D(1, 1)) / dt; b = -D(1, 0) / dt; cf = 0.0; }
        else {
            const double r = D(2, 0) / D(1, 0);
            a = (D(2, 0) + D(2, 1) + D(2, 2)) / dt - r * (D(1, 0) + D(1, 1)) / dt;
            b = -D(2, 1) / dt;
            cf = r;
        }
        HIPCHK(launch_stage_source(ne, dm.npv, dm.nc, dm.ncu, a, s.UDGi, b, c->UDG, cf, c->SH, c->stream));
        int conv = 0, it = 0;
        int rc = hfem_solve_steady(c, p, &conv, &it, nullptr);
        total += it;
        if (rc) { c->ph = saved; return rc; }
        if (!conv) { ok = 0; break; }
    }
    c->ph = saved;
    if (converged) *converged = ok;
    if (newton_iters) *newton_iters = total;
    if (!ok) return hfem_fail(-6, "hfem_dirk_step: a DIRK stage did not converge");
    return 0;
}
