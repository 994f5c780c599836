"""ctypes front end of the CPU oracle (test infrastructure only).

Loads ``oracle/liboracle.so`` (built by ``oracle/Makefile``) and, when built,
``oracle/_ref/libhfemref.so`` (the reference's own element path compiled from
its sources).  Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` import this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        L.hfo_create.restype = ctypes.c_void_p
        L.hfo_create.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int]
        L.hfo_destroy.argtypes = [ctypes.c_void_p]
        L.hfo_assemble.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _ip, _ip,
                                   _dp, _dp, _dp, _dp, _dp, _dp]
        L.hfo_schur.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                _dp, _dp, _dp, _dp]
        L.hfo_getq.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _ip]
        L.hfo_flux.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.hfo_fbou.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.hfo_scatter.argtypes = [ctypes.c_int] * 4 + [_ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp]
        L.hfo_bsr_matvec.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, _dp, _dp]
        L.hfo_dense_inverse.argtypes = [ctypes.c_int, _dp]
        L.hfo_mdf_order.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, _ip, _ip]
        L.hfo_bilu0_factor.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _ip, _dp]
        L.hfo_bilu0_apply.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, _ip, _dp, _dp, _dp]
        L.hfo_block_jacobi_factor.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, _dp]
        L.hfo_gmres.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, _dp, _ip, _ip, _dp, _dp,
                                ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _ip, _dp]
        L.hfo_update_u.argtypes = [ctypes.c_int] * 6 + [_ip, _dp, _dp, _dp, ctypes.c_double, _dp]
        _lib = L
    return _lib


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """FP64 CPU restatement of the element path for one Problem."""

    def __init__(self, prob):
        self.prob = prob
        m = prob.master
        L = lib()
        self._keep = dict(
            dims=_c32([prob.nd, prob.ncu, m.npv, m.ngv, m.npf, m.ngf, m.nfe]),
            shapvt=_c64(m.flat("shapvt")), shapvg=_c64(m.flat("shapvg")), sgl=_c64(m.flat("shapvgdotshapvl")),
            shapft=_c64(m.flat("shapft")), shapfg=_c64(m.flat("shapfg")), sfgl=_c64(m.flat("shapfgdotshapfc")),
            perm=_c32(m.perm.reshape(-1, order="F")), param=_c64(prob.param), ui=_c64(prob.ui))
        k = self._keep
        self.ctx = L.hfo_create(_i(k["dims"]), _d(k["shapvt"]), _d(k["shapvg"]), _d(k["sgl"]), _d(k["shapft"]),
                                _d(k["shapfg"]), _d(k["sfgl"]), _i(k["perm"]), _d(k["param"]), _d(k["ui"]),
                                prob.fc_u, prob.fc_q, prob.tdep)
        self.dgnodes = _c64(prob.dgnodes_ref())
        self.elcon = _c32(prob.mesh.elcon)
        self.bf = _c32(prob.mesh.bf)

    def __del__(self):
        if getattr(self, "ctx", None):
            lib().hfo_destroy(self.ctx)
            self.ctx = None

    def assemble(self, e0=0, n=None, UDG=None, UH=None, SH=None):
        p = self.prob
        d = p.dims
        n = p.ne - e0 if n is None else n
        na, nb = d["na"], d["nb"]
        UDG = _c64(p.UDG if UDG is None else UDG)
        UH = _c64(p.UH if UH is None else UH)
        SH = _c64(p.SH if SH is None else SH)
        out = dict(Ru=np.zeros((n, na)), D=np.zeros((n, na * na)), F=np.zeros((n, na * nb)),
                   K=np.zeros((n, nb * na)), H=np.zeros((n, nb * nb)), Rh=np.zeros((n, nb)))
        lib().hfo_assemble(self.ctx, e0, n, _d(self.dgnodes), _d(UDG), _d(UH), _d(SH), _i(self.elcon),
                           _i(self.bf), _d(out["Ru"]), _d(out["D"]), _d(out["F"]), _d(out["K"]), _d(out["H"]),
                           _d(out["Rh"]))
        # rows hold column-major element matrices (flat index i + rows*j); see mat()
        return out

    def getq(self, UDG, UH=None, SH=None):
        p = self.prob
        UDG = _c64(UDG).copy()
        UH = _c64(p.UH if UH is None else UH)
        SH = _c64(p.SH if SH is None else SH)
        lib().hfo_getq(self.ctx, 0, p.ne, _d(self.dgnodes), _d(UDG), _d(UH), _d(SH), _i(self.elcon))
        return UDG

    def flux(self, udg):
        """udg: (np, nc) -> f (np, ncu, nd), f_udg (np, ncu, nd, nc)."""
        p = self.prob
        npt = udg.shape[0]
        u = _c64(udg.T)
        f = np.zeros(npt * p.ncu * p.nd)
        fu = np.zeros(npt * p.ncu * p.nd * p.nc)
        lib().hfo_flux(self.ctx, npt, _d(u), _d(f), _d(fu))
        return f.reshape(p.nd, p.ncu, npt).transpose(2, 1, 0), fu.reshape(p.nc, p.nd, p.ncu, npt).transpose(3, 2, 1, 0)

    def fbou(self, udg, uh, nl):
        p = self.prob
        npt = udg.shape[0]
        fb = np.zeros(npt * p.nch)
        fbu = np.zeros(npt * p.nch * p.nc)
        fbuh = np.zeros(npt * p.nch * p.nch)
        lib().hfo_fbou(self.ctx, npt, _d(_c64(udg.T)), _d(_c64(uh.T)), _d(_c64(nl.T)), _d(fb), _d(fbu), _d(fbuh))
        return (fb.reshape(p.nch, npt).T, fbu.reshape(p.nc, p.nch, npt).transpose(2, 1, 0),
                fbuh.reshape(p.nch, p.nch, npt).transpose(2, 1, 0))


def schur(impl, npv, ncu, blocks):
    """Condense element blocks in place (dict from Oracle.assemble); returns the dict."""
    n, na = blocks["Ru"].shape
    nb = blocks["Rh"].shape[1]
    info = lib().hfo_schur(impl, n, npv, ncu, nb, _d(blocks["D"]), _d(blocks["F"]), _d(blocks["K"]),
                           _d(blocks["H"]), _d(blocks["Ru"]), _d(blocks["Rh"]))
    return info


def mat(a, rows, cols):
    """View element-major flat column-major blocks as (n, rows, cols) matrices."""
    return a.reshape(-1, cols, rows).transpose(0, 2, 1)


def scatter(prob, AH, FH):
    m = prob.master
    msh = prob.mesh
    bsz = prob.nch * m.npf
    Hg = np.zeros(len(msh.ent2ent) * bsz * bsz)
    Rg = np.zeros(msh.nf * bsz)
    lib().hfo_scatter(prob.ne, m.nfe, m.npf, prob.nch, _i(_c32(msh.t2f)), _i(_c32(msh.elcon)),
                      _i(_c32(msh.ent2entStart)), _i(_c32(msh.ent2ent)), _d(_c64(AH)), _d(_c64(FH)), _d(Hg), _d(Rg))
    return Hg, Rg


def bsr_matvec(prob, Hg, x):
    m = prob.master
    msh = prob.mesh
    bsz = prob.nch * m.npf
    y = np.zeros(msh.nf * bsz)
    lib().hfo_bsr_matvec(msh.nf, bsz, _i(_c32(msh.ent2entStart)), _i(_c32(msh.ent2ent)), _d(_c64(Hg)),
                         _d(_c64(x)), _d(y))
    return y


def ref_lib():
    """The reference's element path compiled from its sources (oracle/_ref), or None."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libhfemref.so")
        if not os.path.exists(path):
            return None
        try:
            L = ctypes.CDLL(path)
        except OSError:
            return None
        L.ref_element.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _ip, _ip, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int] + [_dp] * 10
        L.ref_getq.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp, ctypes.c_double, ctypes.c_int,
                               ctypes.c_int, _dp, _dp, _dp, _dp, _ip, _ip]
        L.ref_time.restype = ctypes.c_double
        L.ref_time.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp,
                               _dp, _dp, _ip, _ip, ctypes.c_int, ctypes.c_int]
        if hasattr(L, "ref_time_mt"):
            L.ref_time_mt.restype = ctypes.c_double
            L.ref_time_mt.argtypes = L.ref_time.argtypes + [ctypes.c_int]
        _ref = L
    return _ref


class Reference(Oracle):
    """Runs the compiled reference element path on the same inputs as Oracle."""

    def __init__(self, prob):
        if ref_lib() is None:
            raise RuntimeError("oracle/_ref/libhfemref.so is not built")
        super().__init__(prob)

    def _common(self):
        k = self._keep
        return (_i(k["dims"]), _d(k["shapvt"]), _d(k["shapvg"]), _d(k["sgl"]), _d(k["shapft"]), _d(k["shapfg"]),
                _d(k["sfgl"]), _i(k["perm"]), _d(k["param"]), _d(k["ui"]))

    def element(self, impl=1, e0=0, n=None, UDG=None, UH=None, SH=None):
        p = self.prob
        d = p.dims
        n = p.ne - e0 if n is None else n
        na, nb = d["na"], d["nb"]
        UDG = _c64(p.UDG if UDG is None else UDG)
        UH = _c64(p.UH if UH is None else UH)
        SH = _c64(p.SH if SH is None else SH)
        out = dict(Ru=np.zeros((n, na)), D=np.zeros((n, na * na)), F=np.zeros((n, na * nb)),
                   K=np.zeros((n, nb * na)), H=np.zeros((n, nb * nb)), Rh=np.zeros((n, nb)),
                   DinvRu=np.zeros((n, na)), DinvF=np.zeros((n, na * nb)), AH=np.zeros((n, nb * nb)),
                   FH=np.zeros((n, nb)))
        ref_lib().ref_element(*self._common(), p.fc_u, p.fc_q, p.tdep, p.ne, UH.shape[0], _d(self.dgnodes), _d(UDG),
                              _d(UH), _d(SH), _i(self.elcon), _i(self.bf), e0, n, impl,
                              *[_d(out[k]) for k in ["Ru", "D", "F", "K", "H", "Rh", "DinvRu", "DinvF", "AH", "FH"]])
        return out

    def getq(self, UDG, UH=None, SH=None):
        p = self.prob
        UDG = _c64(UDG).copy()
        UH = _c64(p.UH if UH is None else UH)
        SH = _c64(p.SH if SH is None else SH)
        ref_lib().ref_getq(*self._common(), p.fc_q, p.ne, UH.shape[0], _d(self.dgnodes), _d(UDG), _d(UH), _d(SH),
                           _i(self.elcon), _i(self.bf))
        return UDG

    def time(self, nelem, impl=1):
        """Seconds for nelem assemble+condense calls of the reference path (1 thread)."""
        p = self.prob
        return ref_lib().ref_time(*self._common(), p.ne, p.UH.shape[0], _d(self.dgnodes), _d(_c64(p.UDG)),
                                  _d(_c64(p.UH)), _d(_c64(p.SH)), _i(self.elcon), _i(self.bf), nelem, impl)

    def time_mt(self, nelem, impl=1, threads=1):
        """Wall-clock seconds for nelem assemble+condense calls on `threads` host
        threads (each with its own reference structs)."""
        p = self.prob
        return ref_lib().ref_time_mt(*self._common(), p.ne, p.UH.shape[0], _d(self.dgnodes), _d(_c64(p.UDG)),
                                     _d(_c64(p.UH)), _d(_c64(p.SH)), _i(self.elcon), _i(self.bf), nelem, impl,
                                     threads)


# --------------------------------------------------------------------------------------------
# trace-system solver (restated from the reference's MATLAB specifications; see solver_oracle.c)

class BSR:
    """A BSR matrix in the reference layout: rows = faces, the diagonal block first."""

    def __init__(self, rowpts, colind, blocks, bsz):
        self.rowpts = _c32(rowpts)
        self.colind = _c32(colind)
        self.bsz = int(bsz)
        self.nrows = len(self.rowpts) - 1
        self.blocks = _c64(blocks).reshape(-1)

    @classmethod
    def from_problem(cls, prob, Hg):
        msh = prob.mesh
        return cls(msh.ent2entStart, msh.ent2ent, Hg, prob.nch * prob.master.npf)

    def dense(self):
        n, b = self.nrows, self.bsz
        A = np.zeros((n * b, n * b))
        blk = self.blocks.reshape(-1, b, b).transpose(0, 2, 1)
        for i in range(n):
            for k in range(self.rowpts[i], self.rowpts[i + 1]):
                j = self.colind[k]
                A[i * b:(i + 1) * b, j * b:(j + 1) * b] = blk[k]
        return A

    def matvec(self, x):
        y = np.zeros(self.nrows * self.bsz)
        lib().hfo_bsr_matvec(self.nrows, self.bsz, _i(self.rowpts), _i(self.colind), _d(self.blocks), _d(_c64(x)),
                             _d(y))
        return y


def dense_inverse(A):
    A = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    info = lib().hfo_dense_inverse(A.shape[0], A.ctypes.data_as(_dp))
    return A, info


def mdf_order(M, exact=False):
    o2u = np.zeros(M.nrows, dtype=np.int32)
    u2o = np.zeros(M.nrows, dtype=np.int32)
    lib().hfo_mdf_order(M.nrows, M.bsz, _i(M.rowpts), _i(M.colind), _d(M.blocks), int(exact), _i(o2u), _i(u2o))
    return o2u, u2o


def bilu0_factor(M, o2u, u2o):
    Mg = M.blocks.copy()
    info = lib().hfo_bilu0_factor(M.nrows, M.bsz, _i(M.rowpts), _i(M.colind), _i(_c32(o2u)), _i(_c32(u2o)), _d(Mg))
    return Mg, info


def bilu0_apply(M, o2u, u2o, Mg, x):
    v = np.zeros(M.nrows * M.bsz)
    lib().hfo_bilu0_apply(M.nrows, M.bsz, _i(M.rowpts), _i(M.colind), _i(_c32(o2u)), _i(_c32(u2o)), _d(_c64(Mg)),
                          _d(_c64(x)), _d(v))
    return v


def block_jacobi_factor(M):
    Dinv = np.zeros(M.nrows * M.bsz * M.bsz)
    info = lib().hfo_block_jacobi_factor(M.nrows, M.bsz, _i(M.rowpts), _d(M.blocks), _d(Dinv))
    return Dinv, info


PREC_NONE, PREC_BILU0, PREC_BJ = -1, 0, 2


def gmres(M, b, prec=PREC_BILU0, Mg=None, o2u=None, u2o=None, x0=None, restart=30, tol=1e-10, maxit=1000,
          orthog=1):
    """Left-preconditioned restarted GMRES (kernel/gmres_bilu0.m). Returns x, flag, iters, rev."""
    n = M.nrows
    x = np.zeros(n * M.bsz) if x0 is None else _c64(x0).copy()
    o2u = np.arange(n, dtype=np.int32) if o2u is None else _c32(o2u)
    u2o = np.arange(n, dtype=np.int32) if u2o is None else _c32(u2o)
    Mg = np.zeros(1) if Mg is None else _c64(Mg)
    iters = ctypes.c_int()
    rev = np.zeros(maxit)
    flag = lib().hfo_gmres(n, M.bsz, _i(M.rowpts), _i(M.colind), _d(M.blocks), prec, _d(Mg), _i(o2u), _i(u2o),
                           _d(_c64(b)), _d(x), restart, tol, maxit, orthog, ctypes.byref(iters), _d(rev))
    return x, flag, iters.value, rev[:iters.value]


def update_u(prob, DinvF, DinvRu, duh, alpha, UDG):
    """getU (solvers/solutionupdate.cpp:6-67) on all elements; returns the new UDG (u part updated)."""
    d = prob.dims
    UDG = _c64(UDG).copy()
    lib().hfo_update_u(prob.ne, d["npv"], d["nc"], d["ncu"], d["nch"], d["ndf"], _i(_c32(prob.mesh.elcon)),
                       _d(_c64(DinvF)), _d(_c64(DinvRu)), _d(_c64(duh)), alpha, _d(UDG))
    return UDG


def trace_system(prob, UDG=None, UH=None, SH=None):
    """assembleLinearSystem on the CPU: element blocks, condensation (LU with partial
    pivoting), scatter.  Returns (BSR Hg, Rg, rNorm, condensed blocks)."""
    orc = Oracle(prob)
    blocks = orc.assemble(0, prob.ne, UDG=UDG, UH=UH, SH=SH)
    Ru0, Rh0 = blocks["Ru"].copy(), blocks["Rh"].copy()
    msh = prob.mesh
    d = prob.dims
    bsz = prob.nch * prob.master.npf
    # rNorm = sqrt(|Ru|^2 + |r|^2), r the assembled uncondensed Rh (assembly/globalassembly.cpp:538-541)
    zeros = np.zeros((prob.ne, d["nb"] * d["nb"]))
    _, r = scatter(prob, zeros, Rh0)
    rnorm = float(np.sqrt((Ru0 ** 2).sum() + (r ** 2).sum()))
    schur(0, d["npv"], d["ncu"], blocks)
    Hg, Rg = scatter(prob, blocks["H"], blocks["Rh"])
    return BSR(msh.ent2entStart, msh.ent2ent, Hg, bsz), Rg, rnorm, blocks
