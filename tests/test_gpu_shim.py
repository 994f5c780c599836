"""The reference-signature entry points (hfem_amd/shim: assembleLinearSystem,
computeResidualNorm, computeOrdering, computePreconditioner, gmres with the
reference's struct arguments, preprocessing/digasopre.h) driven from reference
structs by tests/shim/shim_harness.cpp, the way solveLinearProblem calls them
(solvers/dgsolvers.cpp:40-113).  They must hand back exactly what the C ABI
gives for the same problem: Hg, Rg, inv(D)F, inv(D)Ru, the ordering, the GMRES
solution and both residual norms, bit for bit (same kernels, same order) -- so
the perm packing, the boundary state app.bcs and every layout conversion of the
shim are right.  The C ABI itself is pinned against the reference's element
path by test_gpu_parity.py.  The harness is built where the reference tree is
present; elsewhere the test is skipped."""
import ctypes
import os

import numpy as np
import pytest

from hfem_amd.problem import make_problem

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "shim", "libshim_harness.so")


def _p(a, t):
    return np.ascontiguousarray(a, dtype=t).ctypes.data_as(ctypes.c_void_p)


@pytest.mark.skipif(not os.path.exists(HARNESS), reason="shim harness not built (needs the reference tree)")
@pytest.mark.parametrize("kind,nc,p,prec", [("tet", 1, 3, 0), ("tri", 3, 2, 0), ("tri", 3, 2, 2)])
def test_reference_signatures_match_c_abi(kind, nc, p, prec):
    from hfem_amd.lib import HDGContext, FP64
    H = ctypes.CDLL(HARNESS)
    H.shim_run.restype = ctypes.c_int
    prob = make_problem(kind, nc, p, seed=7, angle=0.2)
    m, d, msh = prob.master, prob.dims, prob.mesh
    nf, bsz = msh.nf, d["nch"] * m.npf
    nblocks = len(msh.ent2ent)
    print(f"shim {kind}{nc} p{p} prec {prec}: ne={prob.ne} nf={nf}", flush=True)
    dims = np.array([prob.nd, prob.ncu, m.npv, m.ngv, m.npf, m.ngf, m.nfe, prob.ne, nf, nblocks], np.int32)
    solver = np.array([prec, 1, 1, 1, 30, 400], np.int32)
    tol = 1e-10
    out = dict(Hg=np.zeros(nblocks * bsz * bsz), Rg=np.zeros(nf * bsz), DinvRu=np.zeros(prob.ne * d["na"]),
               DinvF=np.zeros(prob.ne * d["na"] * d["nb"]), x=np.zeros(nf * bsz))
    scal = np.zeros(3)
    o2u = np.zeros(nf, np.int32)
    keep = []

    def arr(a, t):
        a = np.ascontiguousarray(a, dtype=t)
        keep.append(a)
        return a.ctypes.data_as(ctypes.c_void_p)

    rc = H.shim_run(arr(dims, np.int32), arr(m.flat("shapvt"), np.float64), arr(m.flat("shapvg"), np.float64),
                    arr(m.flat("shapft"), np.float64), arr(m.flat("shapfg"), np.float64),
                    arr(m.perm.reshape(-1, order="F"), np.int32), arr(prob.param, np.float64),
                    arr(prob.ui, np.float64), ctypes.c_double(prob.fc_u), ctypes.c_double(prob.fc_q),
                    ctypes.c_int(prob.tdep), arr(prob.dgnodes_ref(), np.float64), arr(msh.elcon, np.int32),
                    arr(msh.bf, np.int32), arr(msh.t2f, np.int32), arr(msh.ent2entStart, np.int32),
                    arr(msh.ent2ent, np.int32), arr(prob.UDG, np.float64), arr(prob.UH, np.float64),
                    arr(prob.SH, np.float64), arr(solver, np.int32), ctypes.c_double(tol),
                    *(o.ctypes.data_as(ctypes.c_void_p) for o in out.values()),
                    scal.ctypes.data_as(ctypes.c_void_p), o2u.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    # the same steps through the C ABI
    ctx = HDGContext(prob)
    rn0 = ctx.assemble_linear_system(FP64)
    Hg, Rg = ctx.read_system()
    c = ctx.read_condensed()
    ctx.compute_ordering(1)
    want_o2u, _ = ctx.get_ordering()
    ctx.compute_preconditioner(prec)
    flag, it, _ = ctx.gmres(restart=30, tol=tol, maxit=400, orthog=1)
    x = ctx.trace_solution()
    rn1 = ctx.residual_norm()
    ctx.close()
    print(f"  rNorm {scal[0]:.6e} / {rn0:.6e}, gmres flag {scal[2]:.0f} / {flag} ({it} it)", flush=True)
    assert np.array_equal(out["Hg"], Hg) and np.array_equal(out["Rg"], Rg)
    assert np.array_equal(out["DinvRu"], c["DinvRu"].ravel())
    assert np.array_equal(out["DinvF"], c["DinvF"].ravel())
    assert np.array_equal(o2u, want_o2u)
    assert np.array_equal(out["x"], x)
    assert scal[0] == rn0 and scal[1] == rn1 and int(scal[2]) == flag == 0
