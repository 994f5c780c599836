// Static condensation on caller-owned device blocks that allocates nothing: the
// FP64 path of hfem_condense (the Gauss-Jordan inverse on the MFMA, then the two
// hand-written batched products) with inv(D) kept in a fixed file-scope device
// array instead of context memory.  For callers that cannot allocate device
// memory (code under graph capture, serving loops).  Every call on one device
// shares that device's array: calls that overlap on one device race.
#include "ctx.h"

#include <mutex>

// inv(D) of the elements of one call: at most kScratchDoubles doubles (32 MiB),
// e.g. 419 elements of na = 100.
static constexpr long kScratchDoubles = 4L << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};
int g_maxlds[kMaxDevices] = {};
}  // namespace

// The device address of this device's scratch array (the host symbol is only a
// stand-in), looked up once per device with that device current.
static int scratch_for_device(int dev, double **ptr, int *maxlds)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return hfem_fail(-2, "hfem_condense_noalloc: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
        int lds = 65536;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) lds = 65536;
        g_scratch[dev] = (double *)p;
        g_maxlds[dev] = lds;
    }
    *ptr = g_scratch[dev];
    *maxlds = g_maxlds[dev];
    return 0;
}

extern "C" int hfem_condense_noalloc(int na, int nb, int ldx, int n, const double *D, const double *Fx,
                                     const double *K, double *Hx, double *X, int *info, void *stream)
{
    if (na < 1 || na > 176 || nb < 1 || ldx < nb + 1 || n < 0)
        return hfem_fail(-1, "hfem_condense_noalloc: need 1 <= na <= 176, nb >= 1, ldx >= nb+1, n >= 0");
    if ((long)n * na * na > kScratchDoubles)
        return hfem_fail(-1, "hfem_condense_noalloc: %d elements of na=%d need %ld doubles of scratch, the limit is %ld "
                         "(condense at most %ld elements per call)", n, na, (long)n * na * na, kScratchDoubles,
                         kScratchDoubles / ((long)na * na));
    if (n == 0) return 0;
    if (!D || !Fx || !K || !Hx || !X || !info) return hfem_fail(-1, "hfem_condense_noalloc: null pointer");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return hfem_fail(-1, "hfem_condense_noalloc: device %d out of range", dev);
    double *inv = nullptr;
    int maxlds = 0;
    int rc = scratch_for_device(dev, &inv, &maxlds);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    // the first hand-written variant that fits each product (the context's
    // default choice)
    auto pick = [&](int M, int N) {
        for (int v = 0; v < bgemm_num_variants(); v++)
            if (bgemm_fits(0, v, M, N, (size_t)maxlds)) return v;
        return -1;
    };
    const int v0 = pick(na, nb + 1), v1 = pick(nb + 1, nb);
    if (v0 < 0 || v1 < 0) return hfem_fail(-1, "hfem_condense_noalloc: no product kernel fits na=%d nb=%d", na, nb);
    HIPCHK(launch_dinv(na, 0, n, D, inv, info, s, 0));
    const long sx = (long)na * ldx, sd = (long)na * na, sk = (long)nb * na, sh = (long)nb * (nb + 1);
    BgemmArgs g{};
    g.e0 = 0;
    g.n = n;
    // X = inv(D) [F Ru]
    g.M = na; g.N = nb + 1; g.K = na;
    g.A = inv; g.sa = sd; g.lda = na;
    g.B = Fx; g.sb = sx; g.ldb = ldx;
    g.alpha = 1.0; g.beta = 0.0;
    g.C = X; g.sc = sx; g.ldc = ldx;
    HIPCHK(launch_bgemm(0, v0, g, (size_t)maxlds, s));
    // [AH FH] = [H Rh] - K X, as AH^T = H^T - X^T K^T
    g = BgemmArgs{};
    g.e0 = 0;
    g.n = n;
    g.M = nb + 1; g.N = nb; g.K = na;
    g.A = X; g.sa = sx; g.lda = ldx;
    g.B = K; g.sb = sk; g.ldb = nb;
    g.alpha = -1.0; g.beta = 1.0;
    g.C0 = Hx; g.sc0 = sh; g.ldc0 = nb;
    g.C = Hx; g.sc = sh; g.ldc = nb;
    HIPCHK(launch_bgemm(0, v1, g, (size_t)maxlds, s));
    return 0;
}
