"""HIP quasi-Newton path (stored inv(D) and K, right-hand-side condensation,
Jacobian reuse in the Newton driver) vs the CPU oracle (GPU).

Tolerances: FP64 storage 1e-10 relative (different summation orders); float
storage 1e-6 relative (the GPU's and the oracle's inv(D) differ in the last
FP64 bits, which can move a float rounding by one unit); GMRES iteration counts
and which iterations reuse the Jacobian must agree exactly.
"""
import numpy as np
import pytest

from hfem_amd.problem import make_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


CASES = [("tri", 3, 2), ("tet", 1, 2)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}{c[1]}p{c[2]}")
@pytest.mark.parametrize("accuracy", [1, 0])
def test_assemble_rhs_matches_oracle(oracle_mod, case, accuracy):
    import newton_oracle as N
    from hfem_amd.lib import HDGContext
    kind, nc, p = case
    prob = make_problem(kind, nc, p, seed=11, angle=0.3, elevation=0.2)
    say(f"assemble_rhs {kind}{nc} p{p} accuracy {accuracy}")
    ctx = HDGContext(prob)
    ctx.quasi_newton_storage(accuracy)
    ctx.assemble_linear_system()
    Hg0, Rg0 = ctx.read_system()
    rng = np.random.default_rng(2)
    UDG1 = prob.UDG + 1e-3 * rng.standard_normal(prob.UDG.shape)
    UH1 = prob.UH + 1e-3 * rng.standard_normal(prob.UH.shape)
    ctx.set_solution(UDG1, UH1, prob.SH)
    rn = ctx.assemble_rhs()
    Hg1, Rg1 = ctx.read_system()
    y = ctx.read_condensed()["DinvRu"]
    _, _, _, _, stored = N.trace_system_qn(prob, prob.UDG, prob.UH, prob.SH, accuracy=accuracy)
    Rg_o, rn_o, y_o = N.assemble_rhs(prob, UDG1, UH1, prob.SH, stored)
    tol = 1e-10 if accuracy == 1 else 1e-6
    say(f"  Rg {rel(Rg1, Rg_o):.2e} DinvRu {rel(y, y_o):.2e} rNorm {rn:.6e} / {rn_o:.6e}")
    assert np.array_equal(Hg1, Hg0)
    assert rel(Rg1, Rg_o) < tol
    assert rel(y, y_o) < tol
    assert abs(rn - rn_o) <= 1e-10 * rn_o
    ctx.close()


def test_assemble_rhs_needs_a_stored_jacobian():
    from hfem_amd.lib import HDGContext, HfemError
    prob = make_problem("tri", 2, 2, seed=1)
    ctx = HDGContext(prob)
    with pytest.raises(HfemError):
        ctx.assemble_rhs()
    ctx.quasi_newton_storage(1)
    with pytest.raises(HfemError):
        ctx.assemble_rhs()          # storage on, but no Jacobian assembled yet
    ctx.close()


@pytest.mark.parametrize("accuracy", [1, 0])
def test_quasi_newton_solve_matches_oracle(oracle_mod, accuracy):
    import newton_oracle as N
    from hfem_amd.lib import HDGContext
    prob = make_problem("tri", 3, 2, seed=5, angle=0.3)
    ctx = HDGContext(prob)
    # a reuse ratio no timing can exceed: the Jacobian is always reused until a damped step
    conv, nit, hist = ctx.solve_steady(newton_tol=1e-10, newton_maxiter=30, restart=30, gmres_tol=1e-12,
                                       quasi_newton=1, qn_accuracy=accuracy, qn_reuse_ratio=1e300)
    U, UH = ctx.get_solution()
    Ur, UHr, okr, hr = N.newton_solve_qn(prob, prob.UDG, prob.UH, prob.SH, newton_tol=1e-10, maxiter=30,
                                         restart=30, gmres_tol=1e-12, accuracy=accuracy)
    say(f"  quasi-Newton gpu {nit} it {hist[:, [0, 2]].tolist()}; oracle {[(h['old'], h['gmres_iters']) for h in hr]}")
    assert conv and okr and nit == len(hr)
    for k, h in enumerate(hr):
        assert (hist[k, 2] < 0) == h["reused"]
        assert abs(hist[k, 2]) == h["gmres_iters"]
        assert abs(hist[k, 0] - h["old"]) <= 1e-6 * h["old"] + 1e-12
    assert rel(U, Ur) < 1e-8 and rel(UH, UHr) < 1e-8
    ctx.close()
