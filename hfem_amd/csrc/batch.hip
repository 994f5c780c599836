// Ragged element batches: the elements of several contexts (one per element
// family: polynomial orders, tets and hexes) assembled and condensed by single
// persistent launches with ragged per-bucket offsets.
//
// The reference runs one element family per process (main/digasoser.cpp:62-63);
// each element here still goes through exactly the per-element work of
// assembleElementMatrixVector (assembly/elementassembly.cpp:1481-1499) and
// schurElementMatrixVector (assembly/elementschur.cpp:331-343), with the results
// in each context's own buffers, bit for bit what the per-context entry points
// write.
//
// A bucket is one context's elements.  The buckets' items (elements, or rounds of
// elements for the inverse) are laid end to end and handed out as tickets from a
// device counter, most expensive bucket first, so that a workgroup that finishes
// early takes the next item whatever its bucket and the launch drains without a
// per-family tail.  A workgroup runs the kernel body of the item's family (the
// same bodies as the per-family kernels) and stages that family's master tables
// when it enters a bucket.  Per bucket, the tuner picks the waves per element of
// the inverse (8, 4 or 2 waves, several elements per workgroup in lockstep) and
// the condensation products' launch shapes (per context), and orders the buckets
// by their measured cost.
#include "asm_body.h"
#include "ctx.h"
#include "dinv_body.h"

#include <algorithm>
#include <numeric>
#include <vector>

namespace hfem {

constexpr int kBatchMax = 6;

// 3D element families with batch kernels: slot, npv, npf, spill placement
// (tets of orders 1-4, hexes of orders 1-2)
#define HFEM_BATCH3D(X) X(0, 4, 3, false) X(1, 10, 6, false) X(2, 20, 10, false) X(3, 35, 15, true) \
                        X(4, 8, 4, false) X(5, 27, 9, false)

static int batch_family_slot(const Dims &d)
{
    if (d.nd != 3) return -1;
    const bool sp = asm_layout(d).spill != 0;
#define X(i, a, b, s) if (d.npv == a && d.npf == b && sp == s) return i;
    HFEM_BATCH3D(X)
#undef X
    return -1;
}

// One bucket: a context's elements and buffers.
struct BatchElems {
    Dims d;
    Master m;
    Phys ph;
    const double *dgnodes, *UDG, *UH, *SH;
    const int *elcon, *bf;
    double *fluxbuf;
    long fstride;
    double *Ru, *D, *Fx, *K, *Hx, *Rh, *ru2, *dinv;
    int *info;
    int n, ldx;
};

// A bucket is a family slot (at most one context per family), so that a kernel
// body reads its bucket's arguments at fixed kernel-argument offsets, as the
// per-family kernels do.
struct BatchArgs {
    int total;
    int lo[kBatchMax], hi[kBatchMax];   // ticket range of slot i (empty: lo == hi), end to end
    int code[kBatchMax];                // inverse variant code of slot i
    BatchElems b[kBatchMax];            // slot i's context
    unsigned *ctr;                      // this launch's ticket counter
    double *spill;                      // per-workgroup spill regions of the spill placement
};

// The slot whose range holds ticket t (a wave-uniform scan), or -1.
__device__ __forceinline__ int batch_bucket(const BatchArgs &A, int t)
{
    for (int i = 0; i < kBatchMax; i++)
        if (t >= A.lo[i] && t < A.hi[i]) return i;
    return -1;
}

__global__ void __launch_bounds__(256) batch_flux_kernel(BatchArgs A)
{
    __shared__ int slot;
    TicketSource src{A.ctr, &slot, 0, 0, 0};
    src.carry = src.draw();
    while (src.carry < A.total) {
        const int b = batch_bucket(A, src.carry);
        if (b < 0) break;   // not reached: the ranges cover [0, total)
        src.lo = A.lo[b];
        src.n = A.hi[b] - A.lo[b];
        switch (b) {
#define X(i, a, bb, s)                                                                                          \
        case i:                                                                                                 \
            flux_body<3, a, bb>(A.b[i].d, A.b[i].m, A.b[i].ph, A.b[i].dgnodes, A.b[i].UDG, A.b[i].UH,            \
                                A.b[i].elcon, A.b[i].bf, 0, src.n, A.b[i].fluxbuf, A.b[i].fstride, 0, src);     \
            break;
            HFEM_BATCH3D(X)
#undef X
        default: src.carry = A.total;   // not reached: the host checks every slot
        }
    }
}

__global__ void __launch_bounds__(ASM_THREADS) batch_assemble_kernel(BatchArgs A)
{
    __shared__ int slot;
    TicketSource src{A.ctr, &slot, 0, 0, 0};
    src.carry = src.draw();
    while (src.carry < A.total) {
        const int b = batch_bucket(A, src.carry);
        if (b < 0) break;
        src.lo = A.lo[b];
        src.n = A.hi[b] - A.lo[b];
        switch (b) {
#define X(i, a, bb, s)                                                                                          \
        case i:                                                                                                 \
            assemble_body<3, a, bb, s>(A.b[i].d, A.b[i].m, A.b[i].ph, A.b[i].dgnodes, A.b[i].UDG, A.b[i].UH,     \
                                       A.b[i].SH, A.b[i].elcon, A.b[i].bf, 0, src.n, A.b[i].ldx, A.b[i].fluxbuf, \
                                       A.b[i].fstride, A.b[i].Ru, A.b[i].D, A.b[i].Fx, A.b[i].K, A.b[i].Hx,      \
                                       A.b[i].Rh, A.b[i].ru2, A.spill, 0, blockIdx.x, src);                     \
            break;
            HFEM_BATCH3D(X)
#undef X
        default: src.carry = A.total;
        }
    }
}

// Inverse variants of the 512-thread batch workgroup: (waves per element W,
// elements per workgroup GR) = (8, 1), (4, 2), (2, 4).  Tile counts of the
// batch families: 2, 4, 7, 11 (tets), 3, 9 (hexes); beyond 8 tiles only the
// 8-wave form keeps the tiles in registers.
constexpr int kBatchInvVariants = 3;
static const int kInvW[kBatchInvVariants] = {8, 4, 2};
static bool batch_inv_allowed(int NT, int v)
{
    if (NT > 8) return v == 0;
    if (NT <= 4) return v != 0;
    return v >= 0 && v < kBatchInvVariants;
}
static int batch_inv_code(int NT, int v) { return NT * 4 + v; }

#define HFEM_BATCH_INV(X) X(2, 1, 4, 2) X(2, 2, 2, 4) X(3, 1, 4, 2) X(3, 2, 2, 4) X(4, 1, 4, 2) X(4, 2, 2, 4) \
                          X(7, 0, 8, 1) X(7, 1, 4, 2) X(7, 2, 2, 4) X(9, 0, 8, 1) X(11, 0, 8, 1)

__global__ void __launch_bounds__(512) batch_dinv_kernel(BatchArgs A)
{
    extern __shared__ double bi_lds[];
    __shared__ int slot;
    TicketSource src{A.ctr, &slot, 0, 0, 0};
    src.carry = src.draw();
    while (src.carry < A.total) {
        const int b = batch_bucket(A, src.carry);
        if (b < 0) break;
        src.lo = A.lo[b];
        src.n = A.hi[b] - A.lo[b];
        const BatchElems &E = A.b[b];
        switch (A.code[b]) {
#define X(NT, v, W, GR)                                                                                          \
        case NT * 4 + v:                                                                                         \
            dinv_body<double, double, NT, W, GR>(E.d.na, 0, E.n, E.D, E.dinv, E.info, (char *)bi_lds, src);      \
            break;
            HFEM_BATCH_INV(X)
#undef X
        default: src.carry = A.total;
        }
    }
}

__global__ void batch_reset_kernel(unsigned *ctr)
{
    if (threadIdx.x < 4) ctr[threadIdx.x] = 0u;
}

}  // namespace hfem

struct hfem_batch {
    int device = 0, nb = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t join = nullptr;
    std::vector<hipEvent_t> done;
    std::vector<hfem_ctx *> ctx;     // in the caller's order
    std::vector<int> nelem;          // elements [0, nelem[k]) of ctx[k]
    std::vector<int> order;          // bucket k is ctx[order[k]]
    std::vector<int> inv_var;        // per context: inverse variant
    // 0: one persistent launch over all buckets, 1: one launch per bucket (the
    // per-family kernels, on the batch stream); the assembly's and the inverse's
    int mode_asm = 0, mode_inv = 0;
    float tuned_ms[4] = {0, 0, 0, 0};   // last tuning: assembly one launch / per bucket, inverse likewise
    unsigned *ctr = nullptr;
    double *spill = nullptr;
    int grid_flux = 0, grid_asm = 0, grid_inv = 0;
    size_t lds_flux = 0, lds_asm = 0;
};

static BatchElems batch_elems(hfem_ctx *c, int n)
{
    BatchElems E{};
    E.d = c->d;
    E.m = Master{c->shapvt, c->shapvg, c->shapft, c->shapfg, c->perm, c->invperm};
    E.ph = c->ph;
    E.dgnodes = c->dgnodes; E.UDG = c->UDG; E.UH = c->UH; E.SH = c->SH;
    E.elcon = c->elcon; E.bf = c->bf;
    E.fluxbuf = c->fluxbuf; E.fstride = c->fstride;
    E.Ru = c->Ru; E.D = c->D; E.Fx = c->Fx; E.K = c->K; E.Hx = c->Hx; E.Rh = c->Rh; E.ru2 = c->ru2;
    E.dinv = c->dinv; E.info = c->info;
    E.n = n;
    E.ldx = c->ldx;
    return E;
}

// The launch arguments over the buckets `which` (indices into b->ctx, in ticket
// order); inv: tickets are rounds of the inverse variant's GR elements.
static BatchArgs batch_args(const hfem_batch *b, const std::vector<int> &which, bool inv)
{
    BatchArgs A{};
    int t = 0;
    for (int k : which) {
        hfem_ctx *c = b->ctx[k];
        const int i = batch_family_slot(c->d);
        A.b[i] = batch_elems(c, b->nelem[k]);
        int items = A.b[i].n;
        if (inv) {
            const int NT = (c->d.na + 15) / 16, v = b->inv_var[k], GR = 512 / (64 * kInvW[v]);
            A.code[i] = batch_inv_code(NT, v);
            items = (items + GR - 1) / GR;
        }
        A.lo[i] = t;
        A.hi[i] = t += items;
    }
    A.total = t;
    A.spill = b->spill;
    return A;
}

static size_t batch_inv_lds(const hfem_batch *b, const std::vector<int> &which)
{
    size_t m = 0;
    for (int k : which) {
        const int NT = (b->ctx[k]->d.na + 15) / 16, v = b->inv_var[k], GR = 512 / (64 * kInvW[v]);
        m = std::max(m, (size_t)GR * dinv_group_lds(8, NT));
    }
    return m;
}

// The batch stream waits for every context's queued work; afterwards every
// context stream waits for the batch's.
static int batch_join_in(hfem_batch *b)
{
    for (size_t k = 0; k < b->ctx.size(); k++) {
        HIPCHK(hipEventRecord(b->done[k], b->ctx[k]->stream));
        HIPCHK(hipStreamWaitEvent(b->stream, b->done[k], 0));
    }
    return 0;
}
static int batch_join_out(hfem_batch *b)
{
    HIPCHK(hipEventRecord(b->join, b->stream));
    for (hfem_ctx *c : b->ctx) HIPCHK(hipStreamWaitEvent(c->stream, b->join, 0));
    return 0;
}

static hipError_t batch_launch_inv(hfem_batch *b, const std::vector<int> &which, int mode)
{
    if (mode == 1) {
        for (int k : which) {
            hfem_ctx *c = b->ctx[k];
            hipError_t e = launch_dinv(c->d.na, 0, b->nelem[k], c->D, c->dinv, c->info, b->stream, c->inv_var[0]);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    BatchArgs A = batch_args(b, which, true);
    if (A.total == 0) return hipSuccess;
    A.ctr = b->ctr + 2;
    const size_t lds = batch_inv_lds(b, which);
    hipError_t e = hipFuncSetAttribute((const void *)batch_dinv_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, b->stream, b->ctr);
    hipLaunchKernelGGL(batch_dinv_kernel, dim3(std::min(b->grid_inv, A.total)), dim3(512), lds, b->stream, A);
    return hipGetLastError();
}

static hipError_t batch_launch_asm(hfem_batch *b, const std::vector<int> &which, int mode)
{
    if (mode == 1) {
        for (int k : which) {
            hfem_ctx *c = b->ctx[k];
            const int n = b->nelem[k];
            if (n == 0) continue;
            Master m{c->shapvt, c->shapvg, c->shapft, c->shapfg, c->perm, c->invperm};
            hipError_t e = launch_assemble(c->d, m, c->ph, c->dgnodes, c->UDG, c->UH, c->SH, c->elcon, c->bf, 0, n,
                                           std::min(n, c->grid_asm), c->ldx, c->fluxbuf, c->fstride, c->chunk, c->Ru,
                                           c->D, c->Fx, c->K, c->Hx, c->Rh, c->ru2, c->spill, b->stream, 0);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    BatchArgs A = batch_args(b, which, false);
    if (A.total == 0) return hipSuccess;
    hipError_t e = hipFuncSetAttribute((const void *)batch_flux_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)b->lds_flux);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)batch_assemble_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)b->lds_asm);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, b->stream, b->ctr);
    A.ctr = b->ctr;
    hipLaunchKernelGGL(batch_flux_kernel, dim3(std::min(b->grid_flux, A.total)), dim3(256), b->lds_flux, b->stream, A);
    A.ctr = b->ctr + 1;
    hipLaunchKernelGGL(batch_assemble_kernel, dim3(std::min(b->grid_asm, A.total)), dim3(ASM_THREADS), b->lds_asm,
                       b->stream, A);
    return hipGetLastError();
}

extern "C" int hfem_batch_create(hfem_ctx *const *ctxs, int nctx, const int *nelem, hfem_batch **out)
{
    if (!ctxs || !out) return hfem_fail(-1, "hfem_batch_create: null argument");
    *out = nullptr;
    if (nctx < 1 || nctx > kBatchMax) return hfem_fail(-1, "hfem_batch_create: nctx=%d outside [1,%d]", nctx, kBatchMax);
    for (int k = 0; k < nctx; k++) {
        hfem_ctx *c = ctxs[k];
        if (!c) return hfem_fail(-1, "hfem_batch_create: context %d is null", k);
        for (int j = 0; j < k; j++)
            if (ctxs[j] == c) return hfem_fail(-1, "hfem_batch_create: context %d given twice", k);
        if (c->device != ctxs[0]->device) return hfem_fail(-1, "hfem_batch_create: contexts on different devices");
        if (!c->has_master || !c->has_phys || !c->has_mesh)
            return hfem_fail(-3, "hfem_batch_create: context %d needs master, physics and mesh set", k);
        if (batch_family_slot(c->d) < 0)
            return hfem_fail(-1, "hfem_batch_create: context %d (nd=%d npv=%d npf=%d) is not a batch family "
                                 "(3D tets of orders 1-4, hexes of orders 1-2)", k, c->d.nd, c->d.npv, c->d.npf);
        for (int j = 0; j < k; j++)
            if (batch_family_slot(ctxs[j]->d) == batch_family_slot(c->d))
                return hfem_fail(-1, "hfem_batch_create: contexts %d and %d hold the same element family", j, k);
        const long n = nelem ? nelem[k] : c->hd.ne;
        if (n < 0 || n > c->hd.ne)
            return hfem_fail(-1, "hfem_batch_create: nelem[%d]=%ld outside [0,%ld]", k, n, (long)c->hd.ne);
        if (n > c->chunk)
            return hfem_fail(-1, "hfem_batch_create: context %d: %ld elements, more than one flux chunk (%d)", k, n,
                             c->chunk);
    }
    HIPCHK(hipSetDevice(ctxs[0]->device));
    hfem_batch *b = new hfem_batch();
    b->device = ctxs[0]->device;
    b->nb = nctx;
    b->ctx.assign(ctxs, ctxs + nctx);
    for (int k = 0; k < nctx; k++) b->nelem.push_back(nelem ? nelem[k] : (int)ctxs[k]->hd.ne);
    b->done.assign(nctx, nullptr);
    b->inv_var.assign(nctx, 0);
    b->order.resize(nctx);
    std::iota(b->order.begin(), b->order.end(), 0);
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    for (int i = 0; i < 4 && e == hipSuccess; i++) e = hipEventCreate(&b->ev[i]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming);
    for (int k = 0; k < nctx && e == hipSuccess; k++) e = hipEventCreateWithFlags(&b->done[k], hipEventDisableTiming);
    if (e == hipSuccess) e = hfem_dalloc(&b->ctr, 4);
    int ncu = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device);
    long spill = 0;
    for (int k = 0; k < nctx; k++) {
        const Dims &d = b->ctx[k]->d;
        b->lds_flux = std::max(b->lds_flux, flux_lds_bytes(d));
        b->lds_asm = std::max(b->lds_asm, assemble_lds_bytes(d));
        spill = std::max(spill, assemble_spill_doubles(d));
        // default inverse shape: the most elements per workgroup that the tiles allow
        const int NT = (d.na + 15) / 16;
        b->inv_var[k] = NT > 8 ? 0 : (NT <= 4 ? 2 : 1);
    }
    // persistent grids: the workgroups that are resident at once (one assembly,
    // one inverse workgroup per CU; flux workgroups at their register occupancy)
    b->grid_asm = std::max(ncu, 1);
    b->grid_inv = std::max(ncu, 1);
    b->grid_flux = std::max(ncu, 1) * 2;
    if (e == hipSuccess && spill > 0) e = hfem_dalloc(&b->spill, (size_t)b->grid_asm * spill);
    // most expensive bucket first: the larger local block (npv) first
    std::stable_sort(b->order.begin(), b->order.end(),
                     [&](int x, int y) { return b->ctx[x]->d.na > b->ctx[y]->d.na; });
    if (e != hipSuccess) {
        hfem_batch_destroy(b);
        return hfem_fail(-2, "hfem_batch_create: %s", hipGetErrorString(e));
    }
    *out = b;
    return 0;
}

extern "C" int hfem_batch_destroy(hfem_batch *b)
{
    if (!b) return 0;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (hipEvent_t &v : b->ev)
        if (v) (void)hipEventDestroy(v);
    if (b->join) (void)hipEventDestroy(b->join);
    for (hipEvent_t v : b->done)
        if (v) (void)hipEventDestroy(v);
    if (b->ctr) (void)hipFree(b->ctr);
    if (b->spill) (void)hipFree(b->spill);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

static int batch_check_ready(hfem_batch *b, const char *who)
{
    if (!b) return hfem_fail(-1, "%s: null batch", who);
    for (size_t k = 0; k < b->ctx.size(); k++)
        if (!b->ctx[k]->has_sol) return hfem_fail(-3, "%s: context %zu has no solution set", who, k);
    return 0;
}

extern "C" int hfem_batch_assemble(hfem_batch *b)
{
    int rc = batch_check_ready(b, "hfem_batch_assemble");
    if (rc) return rc;
    HIPCHK(hipSetDevice(b->device));
    if ((rc = batch_join_in(b))) return rc;
    HIPCHK(hipEventRecord(b->ev[0], b->stream));
    HIPCHK(batch_launch_asm(b, b->order, b->mode_asm));
    HIPCHK(hipEventRecord(b->ev[1], b->stream));
    return batch_join_out(b);
}

extern "C" int hfem_batch_condense(hfem_batch *b, int precision)
{
    int rc = batch_check_ready(b, "hfem_batch_condense");
    if (rc) return rc;
    if (precision != HFEM_FP64 && precision != HFEM_MIXED)
        return hfem_fail(-1, "hfem_batch_condense: unknown precision %d", precision);
    HIPCHK(hipSetDevice(b->device));
    if ((rc = batch_join_in(b))) return rc;
    HIPCHK(hipEventRecord(b->ev[2], b->stream));
    if (precision == HFEM_MIXED) {
        // each bucket's mixed condensation (FP32 inverse, adaptive refinement with
        // its convergence read-backs, FP64 Schur product) on its own stream
        HIPCHK(hipEventRecord(b->join, b->stream));
        for (int k : b->order) {
            hfem_ctx *c = b->ctx[k];
            if (b->nelem[k] == 0) continue;
            HIPCHK(hipStreamWaitEvent(c->stream, b->join, 0));
            if ((rc = hfem_condense(c, 0, b->nelem[k], HFEM_MIXED))) return rc;
            HIPCHK(hipEventRecord(b->done[k], c->stream));
            HIPCHK(hipStreamWaitEvent(b->stream, b->done[k], 0));
        }
        HIPCHK(hipEventRecord(b->ev[3], b->stream));
        return batch_join_out(b);
    }
    HIPCHK(batch_launch_inv(b, b->order, b->mode_inv));
    HIPCHK(hipEventRecord(b->join, b->stream));
    // the products of each context on its own stream (its tuned shapes), joined
    // back into the batch stream
    for (int k : b->order) {
        hfem_ctx *c = b->ctx[k];
        HIPCHK(hipStreamWaitEvent(c->stream, b->join, 0));
        c->last_sweeps = c->last_unconverged = c->last_fallback = 0;
        if ((rc = condense_products_fp64(c, 0, b->nelem[k]))) return rc;
        HIPCHK(hipEventRecord(b->done[k], c->stream));
        HIPCHK(hipStreamWaitEvent(b->stream, b->done[k], 0));
    }
    HIPCHK(hipEventRecord(b->ev[3], b->stream));
    return batch_join_out(b);
}

extern "C" int hfem_batch_tune(hfem_batch *b, int *choice, float *ms)
{
    int rc = batch_check_ready(b, "hfem_batch_tune");
    if (rc) return rc;
    HIPCHK(hipSetDevice(b->device));
    const int n = b->nb;
    std::vector<float> tasm(n, 0.f), tinv(n, 0.f);
    // assembled blocks to time on, then the products: each context's own tuner
    // (rocBLAS and hand-written shapes) on its batch elements
    if ((rc = batch_join_in(b))) return rc;
    HIPCHK(batch_launch_asm(b, b->order, 1));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int k = 0; k < n; k++)
        if (b->nelem[k] > 0 && (rc = hfem_tune_condensation(b->ctx[k], b->nelem[k], nullptr, nullptr))) return rc;
    if ((rc = batch_join_in(b))) return rc;
    hipEvent_t t0 = b->ev[0], t1 = b->ev[1];
    auto time_it = [&](auto &&launch, float *out) -> hipError_t {
        hipError_t e = launch();
        float best = 1e30f;
        for (int r = 0; r < 2 && e == hipSuccess; r++) {
            float t = 0;
            if ((e = hipEventRecord(t0, b->stream)) != hipSuccess || (e = launch()) != hipSuccess ||
                (e = hipEventRecord(t1, b->stream)) != hipSuccess || (e = hipEventSynchronize(t1)) != hipSuccess ||
                (e = hipEventElapsedTime(&t, t0, t1)) != hipSuccess)
                break;
            best = std::min(best, t);
        }
        *out = best;
        return e;
    };
    for (int k = 0; k < n; k++) {
        const std::vector<int> one{k};
        HIPCHK(time_it([&]() { return batch_launch_asm(b, one, 0); }, &tasm[k]));
        // the inverse variants the bucket's tile count allows
        const int NT = (b->ctx[k]->d.na + 15) / 16;
        int bv = b->inv_var[k];
        float bt = 1e30f;
        for (int v = 0; v < kBatchInvVariants; v++) {
            if (!batch_inv_allowed(NT, v)) continue;
            b->inv_var[k] = v;
            float t = 0;
            HIPCHK(time_it([&]() { return batch_launch_inv(b, one, 0); }, &t));
            if (t < bt) { bt = t; bv = v; }
        }
        b->inv_var[k] = bv;
        tinv[k] = bt;
    }
    // tickets of the most expensive bucket first
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
        return (tasm[x] + tinv[x]) / std::max(1, b->nelem[x]) > (tasm[y] + tinv[y]) / std::max(1, b->nelem[y]);
    });
    // one persistent launch or one launch per bucket, whichever is faster over
    // the whole batch in the order just chosen
    float t1a = 0, tna = 0, t1i = 0, tni = 0;
    HIPCHK(time_it([&]() { return batch_launch_asm(b, b->order, 0); }, &t1a));
    HIPCHK(time_it([&]() { return batch_launch_asm(b, b->order, 1); }, &tna));
    HIPCHK(time_it([&]() { return batch_launch_inv(b, b->order, 0); }, &t1i));
    HIPCHK(time_it([&]() { return batch_launch_inv(b, b->order, 1); }, &tni));
    b->mode_asm = tna < t1a ? 1 : 0;
    b->mode_inv = tni < t1i ? 1 : 0;
    b->tuned_ms[0] = t1a;
    b->tuned_ms[1] = tna;
    b->tuned_ms[2] = t1i;
    b->tuned_ms[3] = tni;
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int k = 0; k < n; k++) {
        if (choice) choice[k] = kInvW[b->inv_var[k]];
        if (ms) {
            ms[k] = tasm[k];
            ms[n + k] = tinv[k];
        }
    }
    return batch_join_out(b);
}

extern "C" int hfem_batch_set_inverse_waves(hfem_batch *b, const int *waves)
{
    if (!b || !waves) return hfem_fail(-1, "hfem_batch_set_inverse_waves: null argument");
    std::vector<int> v(b->nb);
    for (int k = 0; k < b->nb; k++) {
        const int NT = (b->ctx[k]->d.na + 15) / 16;
        v[k] = -1;
        for (int i = 0; i < kBatchInvVariants; i++)
            if (kInvW[i] == waves[k] && batch_inv_allowed(NT, i)) v[k] = i;
        if (v[k] < 0)
            return hfem_fail(-1, "hfem_batch_set_inverse_waves: %d waves per element not available for context %d "
                                 "(%d tiles)", waves[k], k, NT);
    }
    b->inv_var = v;
    return 0;
}

extern "C" int hfem_batch_set_launch(hfem_batch *b, int assemble_mode, int inverse_mode)
{
    if (!b) return hfem_fail(-1, "hfem_batch_set_launch: null batch");
    if (assemble_mode < 0 || assemble_mode > 1 || inverse_mode < 0 || inverse_mode > 1)
        return hfem_fail(-1, "hfem_batch_set_launch: modes must be 0 or 1 (got %d, %d)", assemble_mode, inverse_mode);
    b->mode_asm = assemble_mode;
    b->mode_inv = inverse_mode;
    return 0;
}

extern "C" int hfem_batch_get_launch(hfem_batch *b, int *modes, float *tuned_ms)
{
    if (!b) return hfem_fail(-1, "hfem_batch_get_launch: null batch");
    if (modes) {
        modes[0] = b->mode_asm;
        modes[1] = b->mode_inv;
    }
    if (tuned_ms)
        for (int i = 0; i < 4; i++) tuned_ms[i] = b->tuned_ms[i];
    return 0;
}

extern "C" int hfem_batch_kernel_ms(hfem_batch *b, float *assemble_ms, float *condense_ms)
{
    if (!b) return hfem_fail(-1, "hfem_batch_kernel_ms: null batch");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (assemble_ms) HIPCHK(hipEventElapsedTime(assemble_ms, b->ev[0], b->ev[1]));
    if (condense_ms) HIPCHK(hipEventElapsedTime(condense_ms, b->ev[2], b->ev[3]));
    return 0;
}

extern "C" int hfem_batch_sync(hfem_batch *b)
{
    if (!b) return hfem_fail(-1, "hfem_batch_sync: null batch");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (hfem_ctx *c : b->ctx) HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}
