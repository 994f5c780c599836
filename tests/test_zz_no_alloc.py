"""hfem_condense_noalloc (GPU): the condensation of caller-owned device blocks
with its scratch in a fixed device array equals hfem_condense of the same
blocks (FP64, 1e-13 relative), on two calls in a row, the second with fewer
elements, and rejects a batch beyond its scratch with a clear error."""
import ctypes

import numpy as np
import pytest

from hfem_amd.problem import make_problem

pytestmark = pytest.mark.gpu

RTOL = 1e-13


def relerr(a, b):
    a = np.asarray(a).reshape(len(b), -1)
    b = np.asarray(b).reshape(len(b), -1)
    return float((np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)).max())


def internal_blocks(ctx, n):
    """The assembled blocks of elements [0, n) in the library's internal device
    layouts: D (na x na column-major), [F Ru] (na x ldx row-major), K (nb x na
    column-major), [H Rh] (nb x (nb+1) column-major)."""
    d = ctx.d
    na, nb = d["na"], d["nb"]
    ldx = (nb + 1 + 15) // 16 * 16
    g = ctx.read_elements(0, n)
    D = g["D"].reshape(n, na * na)
    Fx = np.zeros((n, na, ldx))
    Fx[:, :, :nb] = g["F"].reshape(n, nb, na).transpose(0, 2, 1)
    Fx[:, :, nb] = g["Ru"]
    K = g["K"].reshape(n, nb * na)
    Hx = np.concatenate([g["H"].reshape(n, nb * nb), g["Rh"]], axis=1)
    return ldx, D, Fx.reshape(n, -1), K, Hx


class DeviceBuffers:
    """Device memory through the HIP runtime the library is linked against (the
    test's own plumbing; hfem_condense_noalloc itself allocates nothing)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def test_condense_noalloc_matches_condense():
    from hfem_amd.lib import HDGContext, FP64, load
    L = load()
    pr = make_problem("tet", 2, 3, seed=11)
    d = pr.dims
    na, nb = d["na"], d["nb"]
    ctx = HDGContext(pr)
    buf = DeviceBuffers()
    for n in (pr.ne, pr.ne // 3):
        print(f"noalloc: tet p3 na={na} nb={nb} n={n}", flush=True)
        ctx.assemble_element_matrix_vector()
        ldx, D, Fx, K, Hx = internal_blocks(ctx, n)
        ctx.schur_element_matrix_vector(0, n, FP64)
        want = ctx.read_condensed(0, n)
        dD, dF, dK, dH = (buf.upload(a) for a in (D, Fx, K, Hx))
        X0 = np.full((n, na * ldx), np.nan)
        dX = buf.upload(X0)
        dI = buf.upload(np.full(n, -7, dtype=np.int32))
        rc = L.hfem_condense_noalloc(na, nb, ldx, n, dD, dF, dK, dH, dX, dI, None)
        assert rc == 0, L.hfem_last_error()
        Xh = buf.download(dX, X0).reshape(n, na, ldx)
        Hh = buf.download(dH, Hx)
        info = buf.download(dI, np.zeros(n, dtype=np.int32))
        buf.free()
        got = dict(DinvF=Xh[:, :, :nb].transpose(0, 2, 1).reshape(n, -1), DinvRu=Xh[:, :, nb],
                   AH=Hh[:, :nb * nb], FH=Hh[:, nb * nb:])
        assert np.all(info == 0)
        for k in ("DinvF", "DinvRu", "AH", "FH"):
            e = relerr(got[k], want[k])
            print(f"  {k}: {e:.2e}", flush=True)
            assert e < RTOL, k
    # beyond the scratch limit: refused before any work
    big = (4 << 20) // (na * na) + 1
    rc = L.hfem_condense_noalloc(na, nb, (nb + 16) // 16 * 16, big, None, None, None, None, None, None, None)
    assert rc != 0 and b"limit" in L.hfem_last_error()
    ctx.close()
