"""Host-side front end of the HIP library (ctypes over the C ABI in include/hfem_amd.h).

``HDGContext`` mirrors the reference's element-path interface for one mesh:
the method names follow the reference functions they replace
(assembleElementMatrixVector -> ``assemble_element_matrix_vector``,
schurElementMatrixVector -> ``schur_element_matrix_vector``,
assembleLinearSystem -> ``assemble_linear_system`` ...), batched over element
ranges with the data resident on the GPU.  Errors raise ``HfemError`` with the
library's message (the reference prints and exits, preprocessing/errormsg.cpp).
There is no CPU fallback: without the built library or a GPU this raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhfem_amd.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

FP64 = 0
MIXED = 1

PREC_NONE, PREC_BILU0, PREC_BLOCK_JACOBI = -1, 0, 2
ORDER_NATURAL, ORDER_MDF = 0, 1
MGS, ICGS = 0, 1
HIST = 5            # doubles per Newton iteration in the solve_steady history (HFEM_HIST)


class HfemError(RuntimeError):
    pass


class hfem_dims(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("nd", "ncu", "npv", "ngv", "npf", "ngf", "nfe", "ne", "nf")]


EXPORTS = {
    "hfem_last_error": (ctypes.c_char_p, []),
    "hfem_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(hfem_dims), ctypes.c_int]),
    "hfem_destroy": (ctypes.c_int, [_vp]),
    "hfem_set_master": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _ip]),
    "hfem_set_physics": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "hfem_set_mesh": (ctypes.c_int, [_vp, _dp, _ip, _ip, _ip, _ip, _ip, ctypes.c_int]),
    "hfem_set_solution": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "hfem_get_solution": (ctypes.c_int, [_vp, _dp, _dp]),
    "hfem_assemble": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "hfem_assemble_vector": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "hfem_condense": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "hfem_tune_condensation": (ctypes.c_int, [_vp, ctypes.c_int, _ip, ctypes.POINTER(ctypes.c_float)]),
    "hfem_set_refinement": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int]),
    "hfem_condense_report": (ctypes.c_int, [_vp, _ip, _ip, _ip]),
    "hfem_set_condensation_choice": (ctypes.c_int, [_vp, _ip]),
    "hfem_condense_variants": (ctypes.c_int, []),
    "hfem_condense_noalloc": (ctypes.c_int, [ctypes.c_int] * 4 + [ctypes.c_void_p] * 7),
    "hfem_read_elements": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
    "hfem_read_condensed": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _ip]),
    "hfem_assemble_linear_system": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "hfem_quasi_newton_storage": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hfem_assemble_rhs": (ctypes.c_int, [_vp, _dp]),
    "hfem_read_system": (ctypes.c_int, [_vp, _dp, _dp]),
    "hfem_bsr_matvec": (ctypes.c_int, [_vp, _dp, _dp]),
    "hfem_compute_ordering": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hfem_set_ordering": (ctypes.c_int, [_vp, _ip]),
    "hfem_get_ordering": (ctypes.c_int, [_vp, _ip, _ip]),
    "hfem_precond_factor": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hfem_precond_apply": (ctypes.c_int, [_vp, _dp, _dp]),
    "hfem_set_precond_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hfem_set_apply_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hfem_set_bilu0_sweep": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "hfem_gmres": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _ip, _dp, _ip]),
    "hfem_read_trace_solution": (ctypes.c_int, [_vp, _dp]),
    "hfem_update_solution": (ctypes.c_int, [_vp, ctypes.c_double]),
    "hfem_get_q": (ctypes.c_int, [_vp]),
    "hfem_residual_norm": (ctypes.c_int, [_vp, _dp]),
    "hfem_solve_steady": (ctypes.c_int, [_vp, ctypes.c_void_p, _ip, _ip, _dp]),
    "hfem_dirk_step": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, _ip, _ip]),
    "hfem_solver_stats": (ctypes.c_int, [_vp, _dp]),
    "hfem_last_kernel_ms": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "hfem_sync": (ctypes.c_int, [_vp]),
    "hfem_set_linearization_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hfem_batch_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _ip, ctypes.POINTER(_vp)]),
    "hfem_batch_destroy": (ctypes.c_int, [_vp]),
    "hfem_batch_assemble": (ctypes.c_int, [_vp]),
    "hfem_batch_condense": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hfem_batch_tune": (ctypes.c_int, [_vp, _ip, ctypes.POINTER(ctypes.c_float)]),
    "hfem_batch_set_inverse_waves": (ctypes.c_int, [_vp, _ip]),
    "hfem_batch_set_launch": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "hfem_batch_get_launch": (ctypes.c_int, [_vp, _ip, ctypes.POINTER(ctypes.c_float)]),
    "hfem_batch_kernel_ms": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "hfem_batch_sync": (ctypes.c_int, [_vp]),
}
EXTRA_EXPORTS = {}


class hfem_solver_params(ctypes.Structure):
    """Newton / GMRES parameters (include/hfem_amd.h); defaults follow the reference's
    MATLAB-side defaults (preprocessing/digasopre.m:424-438) with GMRES restart 30."""
    _fields_ = [("newton_tol", ctypes.c_double), ("newton_maxiter", ctypes.c_int), ("precond", ctypes.c_int),
                ("ordering", ctypes.c_int), ("reuse_ordering", ctypes.c_int), ("restart", ctypes.c_int),
                ("gmres_tol", ctypes.c_double), ("gmres_maxit", ctypes.c_int), ("orthog", ctypes.c_int),
                ("precision", ctypes.c_int), ("quasi_newton", ctypes.c_int), ("qn_accuracy", ctypes.c_int),
                ("true_newton_maxiter", ctypes.c_int), ("qn_reuse_ratio", ctypes.c_double),
                ("prec_precision", ctypes.c_int)]

    @classmethod
    def make(cls, newton_tol=1e-7, newton_maxiter=10, precond=0, ordering=1, reuse_ordering=1, restart=30,
             gmres_tol=1e-6, gmres_maxit=1000, orthog=1, precision=0, quasi_newton=0, qn_accuracy=1,
             true_newton_maxiter=0, qn_reuse_ratio=3.0, prec_precision=1):
        return cls(newton_tol, newton_maxiter, precond, ordering, reuse_ordering, restart, gmres_tol, gmres_maxit,
                   orthog, precision, quasi_newton, qn_accuracy, true_newton_maxiter, qn_reuse_ratio,
                   prec_precision)


def load():
    """Load the built library (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HfemError(f"{LIB_PATH} is not built; run python -m hfem_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in {**EXPORTS, **EXTRA_EXPORTS}.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise HfemError(load().hfem_last_error().decode())


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class HDGContext:
    """Device-resident HDG element path for one Problem (hfem_amd.problem)."""

    def __init__(self, prob, device=0, master_tables=None):
        """``master_tables``: dict(shapvt, shapvg, shapft, shapfg, perm) of flat
        column-major tables to use instead of the problem's master element (for
        instance those of a reference master.bin, hfem_amd.binio.master_tables);
        their sizes must be those of the problem's element."""
        L = load()
        self.prob = prob
        m = prob.master
        d = prob.dims
        self.d = d
        dims = hfem_dims(prob.nd, prob.ncu, m.npv, m.ngv, m.npf, m.ngf, m.nfe, prob.ne, prob.mesh.nf)
        t = master_tables or dict(shapvt=m.flat("shapvt"), shapvg=m.flat("shapvg"), shapft=m.flat("shapft"),
                                  shapfg=m.flat("shapfg"), perm=m.perm.reshape(-1, order="F"))
        for k, n in (("shapvt", m.npv * m.ngv * (prob.nd + 1)), ("shapvg", m.npv * m.ngv * (prob.nd + 1)),
                     ("shapft", m.npf * m.ngf * prob.nd), ("shapfg", m.npf * m.ngf * prob.nd),
                     ("perm", m.npf * m.nfe)):
            if np.asarray(t[k]).size != n:
                raise HfemError(f"master table {k} has {np.asarray(t[k]).size} entries, the element needs {n}")
        h = _vp()
        _check(L.hfem_create(ctypes.byref(h), ctypes.byref(dims), device))
        self.h = h
        _check(L.hfem_set_master(h, _d(_f64(t["shapvt"])), _d(_f64(t["shapvg"])), _d(_f64(t["shapft"])),
                                 _d(_f64(t["shapfg"])), _i(_i32(t["perm"]))))
        self.set_physics(prob.fc_u, prob.fc_q, prob.tdep)
        msh = prob.mesh
        _check(L.hfem_set_mesh(h, _d(_f64(prob.dgnodes_ref())), _i(_i32(msh.elcon)), _i(_i32(msh.bf)),
                               _i(_i32(msh.t2f)), _i(_i32(msh.ent2entStart)), _i(_i32(msh.ent2ent)),
                               len(msh.ent2ent)))
        self.set_solution(prob.UDG, prob.UH, prob.SH)

    def close(self):
        if getattr(self, "h", None):
            load().hfem_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # -- state ---------------------------------------------------------------
    def set_physics(self, fc_u, fc_q=1.0, tdep=0):
        p = self.prob
        param, ui = np.asarray(p.param).ravel(), np.asarray(p.ui).ravel()
        if param.size < 6:
            raise HfemError(f"physics parameters: {param.size} values, {{gamma, eps, Re, Pr, Minf, tau}} needs 6")
        if ui.size != self.d["nch"]:
            raise HfemError(f"far-field state: {ui.size} values, nch={self.d['nch']} needed")
        _check(load().hfem_set_physics(self.h, _d(_f64(p.param)), _d(_f64(p.ui)), fc_u, fc_q, tdep))

    def set_solution(self, UDG, UH, SH=None):
        _check(load().hfem_set_solution(self.h, _d(_f64(UDG)), _d(_f64(UH)),
                                        _d(_f64(SH)) if SH is not None else None))

    def get_solution(self):
        p = self.prob
        UDG = np.zeros_like(p.UDG)
        UH = np.zeros_like(p.UH)
        _check(load().hfem_get_solution(self.h, _d(UDG), _d(UH)))
        return UDG, UH

    # -- element path ----------------------------------------------------------
    def assemble_element_matrix_vector(self, e0=0, n=None):
        n = self.prob.ne - e0 if n is None else n
        _check(load().hfem_assemble(self.h, e0, n))

    def assemble_element_vector(self, e0=0, n=None):
        """assembleElementVector: Ru and Rh only (no Jacobians)."""
        n = self.prob.ne - e0 if n is None else n
        _check(load().hfem_assemble_vector(self.h, e0, n))

    def schur_element_matrix_vector(self, e0=0, n=None, precision=FP64):
        n = self.prob.ne - e0 if n is None else n
        _check(load().hfem_condense(self.h, e0, n, precision))

    def tune_condensation(self, nsample=0):
        """Time the launch candidates of every condensation stage on a sample and
        keep the fastest; returns (choice[7], ms[7]) for {inverse FP64, inverse
        FP32, X = inv(D)[F Ru], Schur product, mixed X0, mixed residual, mixed
        correction} (product choice: hand-written variant >= 0, -1 rocBLAS)."""
        ch = (ctypes.c_int * 7)()
        ms = (ctypes.c_float * 7)()
        _check(load().hfem_tune_condensation(self.h, nsample, ch, ms))
        return list(ch), list(ms)

    def set_condensation_choice(self, choice, strict=True):
        """Set the seven launch choices (the tuner's `choice` layout).  With
        strict=False a choice that does not fit returns False instead of raising."""
        ch = (ctypes.c_int * 7)(*choice)
        rc = load().hfem_set_condensation_choice(self.h, ch)
        if rc != 0 and not strict:
            return False
        _check(rc)
        return True

    def set_refinement(self, tol=1e-11, max_sweeps=4):
        _check(load().hfem_set_refinement(self.h, tol, max_sweeps))

    def condense_report(self):
        """(sweeps, unconverged elements, fallback to FP64) of the last condensation."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(load().hfem_condense_report(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def read_elements(self, e0=0, n=None, which=("Ru", "D", "F", "K", "H", "Rh")):
        d = self.d
        n = self.prob.ne - e0 if n is None else n
        na, nb = d["na"], d["nb"]
        sizes = dict(Ru=na, D=na * na, F=na * nb, K=nb * na, H=nb * nb, Rh=nb)
        out = {k: np.zeros((n, sizes[k])) for k in which}
        args = [_d(out[k]) if k in out else None for k in ("Ru", "D", "F", "K", "H", "Rh")]
        _check(load().hfem_read_elements(self.h, e0, n, *args))
        return out

    def read_condensed(self, e0=0, n=None):
        """DinvRu, DinvF, AH, FH (element-major, reference layouts) and the pivot flags."""
        d = self.d
        n = self.prob.ne - e0 if n is None else n
        na, nb = d["na"], d["nb"]
        out = dict(DinvRu=np.zeros((n, na)), DinvF=np.zeros((n, na * nb)), AH=np.zeros((n, nb * nb)),
                   FH=np.zeros((n, nb)))
        info = np.zeros(n, dtype=np.int32)
        _check(load().hfem_read_condensed(self.h, e0, n, _d(out["DinvRu"]), _d(out["DinvF"]), _d(out["AH"]),
                                          _d(out["FH"]), _i(info)))
        out["info"] = info
        return out

    def kernel_ms(self):
        a = ctypes.c_float()
        c = ctypes.c_float()
        _check(load().hfem_last_kernel_ms(self.h, ctypes.byref(a), ctypes.byref(c)))
        return a.value, c.value

    def sync(self):
        _check(load().hfem_sync(self.h))

    # -- global trace system ----------------------------------------------------
    def assemble_linear_system(self, precision=FP64):
        r = ctypes.c_double()
        _check(load().hfem_assemble_linear_system(self.h, precision, ctypes.byref(r)))
        return r.value

    def quasi_newton_storage(self, accuracy=1):
        """Keep inv(D) and K of every Jacobian: accuracy 0 float, 1 double, -1 off."""
        _check(load().hfem_quasi_newton_storage(self.h, accuracy))

    def assemble_rhs(self):
        """assembleRHS with the stored Jacobian; returns rNorm."""
        r = ctypes.c_double()
        _check(load().hfem_assemble_rhs(self.h, ctypes.byref(r)))
        return r.value

    def read_system(self):
        msh = self.prob.mesh
        bsz = self.d["nch"] * self.d["npf"]
        Hg = np.zeros(len(msh.ent2ent) * bsz * bsz)
        Rg = np.zeros(msh.nf * bsz)
        _check(load().hfem_read_system(self.h, _d(Hg), _d(Rg)))
        return Hg, Rg

    def bsr_matvec(self, x):
        x = _f64(x)
        y = np.zeros_like(x)
        _check(load().hfem_bsr_matvec(self.h, _d(x), _d(y)))
        return y

    # -- trace-system solver (computeOrdering / computePreconditioner / gmres) ---
    def compute_ordering(self, method=ORDER_MDF):
        _check(load().hfem_compute_ordering(self.h, method))

    def set_ordering(self, o2u):
        _check(load().hfem_set_ordering(self.h, _i(_i32(o2u))))

    def get_ordering(self):
        o2u = np.zeros(self.prob.mesh.nf, dtype=np.int32)
        nlev = np.zeros(2, dtype=np.int32)
        _check(load().hfem_get_ordering(self.h, _i(o2u), _i(nlev)))
        return o2u, tuple(int(v) for v in nlev)

    def compute_preconditioner(self, kind=PREC_BILU0, prec_precision=1):
        """computePreconditioner; prec_precision 0 keeps the factors in single
        precision (sys.precPrecision)."""
        _check(load().hfem_set_precond_precision(self.h, prec_precision))
        _check(load().hfem_precond_factor(self.h, kind))

    def set_linearization_graph(self, use=True):
        """Replay the FP64 linearisation of assemble_linear_system as a HIP graph."""
        _check(load().hfem_set_linearization_graph(self.h, int(bool(use))))

    def set_apply_graph(self, use=True):
        """Replay the BILU0 application as a captured HIP graph, or launch its
        levels one by one (default)."""
        _check(load().hfem_set_apply_graph(self.h, int(use)))

    def set_bilu0_sweep(self, mode, spin_limit=0):
        """BILU0 sweeps as one persistent launch per sweep with bounded per-row
        waits (mode 1, default) or level by level (mode 0)."""
        _check(load().hfem_set_bilu0_sweep(self.h, int(mode), int(spin_limit)))

    def apply_preconditioner(self, x):
        x = _f64(x)
        y = np.zeros_like(x)
        _check(load().hfem_precond_apply(self.h, _d(x), _d(y)))
        return y

    def gmres(self, restart=30, tol=1e-6, maxit=1000, orthog=ICGS):
        """Returns (flag, iterations, relative residual history); x via trace_solution()."""
        it = ctypes.c_int()
        fl = ctypes.c_int()
        rev = np.zeros(maxit)
        _check(load().hfem_gmres(self.h, restart, tol, maxit, orthog, ctypes.byref(it), _d(rev), ctypes.byref(fl)))
        return fl.value, it.value, rev[:it.value]

    def trace_solution(self):
        x = np.zeros(self.prob.UH.size)
        _check(load().hfem_read_trace_solution(self.h, _d(x)))
        return x

    def update_solution(self, alpha=1.0):
        _check(load().hfem_update_solution(self.h, alpha))

    def get_q(self):
        _check(load().hfem_get_q(self.h))

    def residual_norm(self):
        r = ctypes.c_double()
        _check(load().hfem_residual_norm(self.h, ctypes.byref(r)))
        return r.value

    def solver_stats(self):
        v = np.zeros(7)
        _check(load().hfem_solver_stats(self.h, _d(v)))
        return dict(order_ms=v[0], factor_ms=v[1], gmres_ms=v[2], matvec_ms=v[3], matvecs=int(v[4]),
                    precond_ms=v[5], preconds=int(v[6]))

    # -- drivers (solveNonlinearProblem, solveUnsteadyProblem) -------------------
    def solve_steady(self, params=None, **kw):
        """Newton with line search; returns (converged, iterations, history[it, 5]):
        residual before and after, GMRES iterations (negative for iterations that
        reused the Jacobian), accumulated step, line-search steps applied."""
        p = params or hfem_solver_params.make(**kw)
        conv = ctypes.c_int()
        nit = ctypes.c_int()
        hist = np.zeros(HIST * p.newton_maxiter)
        _check(load().hfem_solve_steady(self.h, ctypes.byref(p), ctypes.byref(conv), ctypes.byref(nit), _d(hist)))
        return bool(conv.value), nit.value, hist.reshape(-1, HIST)[:nit.value]

    def dirk_step(self, dt, nstage=3, order=3, params=None, **kw):
        p = params or hfem_solver_params.make(**kw)
        conv = ctypes.c_int()
        nit = ctypes.c_int()
        _check(load().hfem_dirk_step(self.h, dt, nstage, order, ctypes.byref(p), ctypes.byref(conv),
                                     ctypes.byref(nit)))
        return bool(conv.value), nit.value


class HDGBatch:
    """Ragged batch over contexts of different 3D element families (tets of orders
    1-4, hexes of orders 1-2, one context per family): the element path of every
    context in single persistent launches (hfem_batch_* in include/hfem_amd.h).
    Each context's buffers receive what its own assemble / schur calls write, and
    the contexts stay usable on their own (read_elements, read_condensed ...)."""

    def __init__(self, contexts, nelem=None):
        """``nelem``: the batch takes elements [0, nelem[k]) of contexts[k] (default all)."""
        L = load()
        self.contexts = list(contexts)
        arr = (_vp * len(self.contexts))(*[c.h for c in self.contexts])
        ne = None if nelem is None else _i32(nelem)
        h = _vp()
        _check(L.hfem_batch_create(arr, len(self.contexts), None if ne is None else _i(ne), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().hfem_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def assemble_element_matrix_vector(self):
        _check(load().hfem_batch_assemble(self.h))

    def schur_element_matrix_vector(self, precision=FP64):
        _check(load().hfem_batch_condense(self.h, precision))

    def tune(self):
        """Per bucket: the inverse's waves per element and the bucket order; returns
        (waves per context, assembly ms per context, inverse ms per context)."""
        n = len(self.contexts)
        w = np.zeros(n, dtype=np.int32)
        ms = (ctypes.c_float * (2 * n))()
        _check(load().hfem_batch_tune(self.h, _i(w), ms))
        return w.tolist(), list(ms[:n]), list(ms[n:])

    def set_inverse_waves(self, waves):
        _check(load().hfem_batch_set_inverse_waves(self.h, _i(_i32(waves))))

    def set_launch(self, assemble_mode, inverse_mode):
        """0: one persistent launch over all buckets, 1: one launch per bucket."""
        _check(load().hfem_batch_set_launch(self.h, assemble_mode, inverse_mode))

    def launch(self):
        """(assembly mode, inverse mode) and the last tuning's times {assemble_one_launch,
        assemble_per_bucket, inverse_one_launch, inverse_per_bucket} in ms."""
        m = np.zeros(2, dtype=np.int32)
        t = (ctypes.c_float * 4)()
        _check(load().hfem_batch_get_launch(self.h, _i(m), t))
        return (int(m[0]), int(m[1])), dict(zip(("assemble_one_launch", "assemble_per_bucket", "inverse_one_launch",
                                                  "inverse_per_bucket"), list(t)))

    def kernel_ms(self):
        a = ctypes.c_float()
        c = ctypes.c_float()
        _check(load().hfem_batch_kernel_ms(self.h, ctypes.byref(a), ctypes.byref(c)))
        return a.value, c.value

    def sync(self):
        _check(load().hfem_batch_sync(self.h))
