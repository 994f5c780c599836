"""Element path on a batch large enough that every workgroup assembles and
condenses several elements in turn (GPU): a sample of elements, including ones
late in each workgroup's sequence and in later assembly chunks, against the
CPU oracle.  Tolerances as in test_gpu_parity.py."""
import numpy as np
import pytest

from hfem_amd.problem import make_problem

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float((np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)).max())


@pytest.mark.parametrize("kind,nc,p", [("tet", 8, 3), ("tri", 40, 2)])
def test_large_batch_sample_vs_oracle(oracle_mod, kind, nc, p):
    from hfem_amd.lib import HDGContext, FP64, MIXED
    O = oracle_mod
    pr = make_problem(kind, nc, p, seed=2)
    d = pr.dims
    print(f"case {kind}{nc} p{p}: ne={pr.ne}", flush=True)
    assert pr.ne > 2048
    sample = np.unique(np.r_[0, 1, 511, 512, 513, 1023, 1024, 2047, 2048, pr.ne - 1,
                             np.linspace(0, pr.ne - 1, 22).astype(int)])
    orc = O.Oracle(pr)
    ref = {k: [] for k in ("D", "F", "K", "H", "Rh")}
    cond = {k: [] for k in ("DinvF", "AH")}
    for e in sample:
        b = orc.assemble(int(e), 1)
        for k in ref:
            ref[k].append(b[k][0].copy())
        O.schur(0, d["npv"], d["ncu"], b)
        cond["DinvF"].append(b["F"][0])
        cond["AH"].append(b["H"][0])
    ref = {k: np.array(v) for k, v in ref.items()}
    cond = {k: np.array(v) for k, v in cond.items()}
    ctx = HDGContext(pr)
    ctx.assemble_element_matrix_vector()
    got = {k: [] for k in ref}
    for e in sample:
        g = ctx.read_elements(int(e), 1)
        for k in ref:
            got[k].append(g[k][0])
    for k in ref:
        err = relerr(np.array(got[k]), ref[k])
        print(f"  assemble {k}: {err:.2e}", flush=True)
        assert err < 1e-13, k
    for prec, tol in ((FP64, 1e-13), (MIXED, 1e-10)):
        ctx.assemble_element_matrix_vector()
        ctx.schur_element_matrix_vector(precision=prec)
        for k in cond:
            vals = np.array([ctx.read_condensed(int(e), 1)[k][0] for e in sample])
            err = relerr(vals, cond[k])
            print(f"  precision {prec} {k}: {err:.2e}", flush=True)
            assert err < tol, (prec, k)
    ctx.close()
