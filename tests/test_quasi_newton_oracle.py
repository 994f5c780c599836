"""The oracle's quasi-Newton path pinned by identities the algebra fixes.

assembleRHS with schurElementVector (assembly/globalassembly.cpp:559-616,
assembly/elementschur.cpp:345-458) condenses the element vectors of the current
state with the inv(D) and K stored at the last Jacobian
(updateQuasiNewtonStructures, assembly/globalassembly.cpp:327-389).  Pinned by
* the stored operands at the same state: the right-hand side equals the full
  condensation's (FH = Rh - K inv(D) Ru is the same formula);
* a dense numpy solve per element at a different state (layouts and the
  column-major reshapes of D and K);
* float storage (quasiNewtonAccuracy 0): the stored values are exactly
  representable in float and the right-hand side moves by a float-sized amount;
* the Newton iteration with a reused Jacobian converges linearly to the same
  solution as full Newton.
"""
import numpy as np
import pytest

from hfem_amd.problem import make_problem


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def case(oracle_mod):
    import newton_oracle as N
    prob = make_problem("tri", 3, 2, seed=5, angle=0.3)
    M, Rg, rnorm, blocks, stored = N.trace_system_qn(prob, prob.UDG, prob.UH, prob.SH, accuracy=1)
    return prob, M, Rg, rnorm, blocks, stored


def perturbed(prob, seed=2, eps=1e-3):
    rng = np.random.default_rng(seed)
    UDG = prob.UDG + eps * rng.standard_normal(prob.UDG.shape)
    UH = prob.UH + eps * rng.standard_normal(prob.UH.shape)
    return UDG, UH


def test_rhs_at_the_stored_state_is_the_full_condensation(case):
    import newton_oracle as N
    prob, M, Rg, rnorm, blocks, stored = case
    Rg2, rnorm2, y = N.assemble_rhs(prob, prob.UDG, prob.UH, prob.SH, stored)
    assert rel(Rg2, Rg) < 1e-12
    assert abs(rnorm2 - rnorm) <= 1e-14 * rnorm
    assert rel(y, blocks["Ru"]) < 1e-12


def test_rhs_equals_dense_solve_at_another_state(oracle_mod, case):
    import newton_oracle as N
    O = oracle_mod
    prob, M, Rg, rnorm, blocks, stored = case
    d = prob.dims
    na, nb = d["na"], d["nb"]
    raw0 = O.Oracle(prob).assemble(0, prob.ne)
    UDG1, UH1 = perturbed(prob)
    raw1 = O.Oracle(prob).assemble(0, prob.ne, UDG=UDG1, UH=UH1, SH=prob.SH)
    Rg1, _, y = N.assemble_rhs(prob, UDG1, UH1, prob.SH, stored)
    FH = np.empty((prob.ne, nb))
    for e in range(prob.ne):
        D0 = raw0["D"][e].reshape(na, na, order="F")
        K0 = raw0["K"][e].reshape(nb, na, order="F")
        ye = np.linalg.solve(D0, raw1["Ru"][e])
        assert rel(y[e], ye) < 1e-11
        FH[e] = raw1["Rh"][e] - K0 @ ye
    _, Rg_dense = O.scatter(prob, np.zeros((prob.ne, nb * nb)), FH)
    assert rel(Rg1, Rg_dense) < 1e-11
    # the state moved: the right-hand side is not the stored state's
    assert rel(Rg1, Rg) > 1e-6


def test_float_storage_rounds_the_stored_operands(case):
    import newton_oracle as N
    prob, M, Rg, rnorm, blocks, stored = case
    _, _, _, _, st32 = N.trace_system_qn(prob, prob.UDG, prob.UH, prob.SH, accuracy=0)
    for k in ("Dinv", "K"):
        assert np.array_equal(st32[k], st32[k].astype(np.float32).astype(np.float64))
        d = np.abs(st32[k] - stored[k])
        assert d.max() <= 2.0 ** -24 * np.abs(stored[k]).max() and d.max() > 0
    Rg32, _, _ = N.assemble_rhs(prob, prob.UDG, prob.UH, prob.SH, st32)
    e = rel(Rg32, Rg)
    assert 1e-12 < e < 1e-5


def test_quasi_newton_converges_to_the_newton_solution(oracle_mod):
    import newton_oracle as N
    prob = make_problem("tri", 3, 2, seed=5, angle=0.3)
    U0, UH0, ok0, h0 = N.newton_solve(prob, prob.UDG, prob.UH, prob.SH, newton_tol=1e-11, maxiter=8, restart=30,
                                      gmres_tol=1e-12)
    U1, UH1, ok1, h1 = N.newton_solve_qn(prob, prob.UDG, prob.UH, prob.SH, newton_tol=1e-10, maxiter=30,
                                         restart=30, gmres_tol=1e-12)
    assert ok0 and ok1
    assert not h1[0]["reused"] and any(h["reused"] for h in h1)
    # a reused Jacobian converges linearly, so it takes more iterations than Newton
    assert len(h1) > len(h0)
    assert rel(U1, U0) < 1e-8 and rel(UH1, UH0) < 1e-8
