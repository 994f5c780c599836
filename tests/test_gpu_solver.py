"""HIP trace-system solver and Newton / DIRK drivers vs the CPU oracle (GPU).

The oracle restates the reference's MATLAB specifications of the closed solver
pieces (oracle/solver_oracle.c) and its C++ drivers (oracle/newton_oracle.py);
those are pinned by the properties in tests/test_solver_oracle.py.
Tolerances: FP64 throughout; orderings and GMRES iteration counts must agree
exactly, vectors to 1e-10 relative (different summation orders), Newton
residual histories to 1e-10 relative above the 1e-12 level.
"""
import dataclasses

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


@pytest.fixture(params=CASES, ids=lambda c: f"{c[0]}{c[1]}p{c[2]}")
def system(request, oracle_mod):
    from hfem_amd.lib import HDGContext
    kind, nc, p = request.param
    prob = make_problem(kind, nc, p, seed=11, angle=0.3, elevation=0.2)
    say(f"case {kind}{nc} p{p}: nf={prob.mesh.nf}")
    ctx = HDGContext(prob)
    ctx.assemble_linear_system()
    Hg, Rg = ctx.read_system()
    M = oracle_mod.BSR.from_problem(prob, Hg)
    yield prob, ctx, M, Rg
    ctx.close()


def test_get_q_matches_oracle(oracle_mod):
    from hfem_amd.lib import HDGContext
    for kind, nc, p in [("tri", 2, 3), ("tet", 1, 3)]:
        prob = make_problem(kind, nc, p, seed=4)
        say(f"getq {kind}{nc} p{p}")
        ctx = HDGContext(prob)
        ctx.get_q()
        U, _ = ctx.get_solution()
        ref = oracle_mod.Oracle(prob).getq(prob.UDG)
        assert rel(U, ref) < 1e-12
        ctx.close()


def test_mdf_ordering_matches_oracle(oracle_mod, system):
    prob, ctx, M, Rg = system
    ctx.compute_ordering(1)
    o2u, nlev = ctx.get_ordering()
    ref, _ = oracle_mod.mdf_order(M)
    say(f"  levels {nlev}")
    assert np.array_equal(o2u, ref)


@pytest.mark.parametrize("kind", ["bilu0", "bj"])
def test_preconditioner_matches_oracle(oracle_mod, system, kind):
    from hfem_amd.lib import PREC_BILU0, PREC_BLOCK_JACOBI
    O = oracle_mod
    prob, ctx, M, Rg = system
    x = np.random.default_rng(3).standard_normal(len(Rg))
    if kind == "bilu0":
        o2u, u2o = O.mdf_order(M)
        ctx.set_ordering(o2u)
        ctx.compute_preconditioner(PREC_BILU0)
        Mg, _ = O.bilu0_factor(M, o2u, u2o)
        want = O.bilu0_apply(M, o2u, u2o, Mg, x)
    else:
        ctx.compute_preconditioner(PREC_BLOCK_JACOBI)
        Dinv, _ = O.block_jacobi_factor(M)
        blocks = Dinv.reshape(-1, M.bsz, M.bsz).transpose(0, 2, 1)
        want = np.concatenate([blocks[i] @ x[i * M.bsz:(i + 1) * M.bsz] for i in range(M.nrows)])
    got = ctx.apply_preconditioner(x)
    say(f"  {kind} apply rel err {rel(got, want):.2e}")
    assert rel(got, want) < 1e-10


def test_bilu0_apply_graph_replays(oracle_mod, system):
    """The BILU0 application replayed from its captured HIP graph equals the
    level-by-level launches bit for bit, on every replay, with other device work
    (a matvec, a GMRES solve, a re-factorisation) between the replays."""
    from hfem_amd.lib import PREC_BILU0
    O = oracle_mod
    prob, ctx, M, Rg = system
    rng = np.random.default_rng(5)
    xs = [rng.standard_normal(len(Rg)) for _ in range(3)]
    o2u, u2o = O.mdf_order(M)
    ctx.set_ordering(o2u)
    ctx.set_apply_graph(False)
    ctx.compute_preconditioner(PREC_BILU0)
    want = [ctx.apply_preconditioner(x) for x in xs]
    ctx.set_apply_graph(True)
    ctx.compute_preconditioner(PREC_BILU0)
    for rep in range(3):
        for x, w in zip(xs, want):
            assert np.array_equal(ctx.apply_preconditioner(x), w), rep
        ctx.bsr_matvec(xs[rep])
        if rep == 1:
            ctx.gmres(restart=10, tol=1e-8, maxit=50)
        if rep == 2:
            ctx.compute_preconditioner(PREC_BILU0)   # the graph is captured again
    say("  graph replays equal the level launches")


@pytest.mark.parametrize("kind", ["bilu0", "bj"])
def test_single_precision_preconditioner(oracle_mod, system, kind):
    """precPrecision 0: factors in FP32 (BILU0 factored on the FP32 MFMA).  The
    application equals the oracle's FP64 preconditioner to FP32 accuracy (a
    relative 1e-4 leaves room for the block condition numbers; a wrong block or
    index is O(1)), and GMRES with it reaches the same FP64 tolerance and the same
    solution as with FP64 factors."""
    from hfem_amd.lib import PREC_BILU0, PREC_BLOCK_JACOBI
    O = oracle_mod
    prob, ctx, M, Rg = system
    x = np.random.default_rng(3).standard_normal(len(Rg))
    pk = PREC_BILU0 if kind == "bilu0" else PREC_BLOCK_JACOBI
    if kind == "bilu0":
        o2u, u2o = O.mdf_order(M)
        ctx.set_ordering(o2u)
        Mg, _ = O.bilu0_factor(M, o2u, u2o)
        want = O.bilu0_apply(M, o2u, u2o, Mg, x)
    else:
        Dinv, _ = O.block_jacobi_factor(M)
        blocks = Dinv.reshape(-1, M.bsz, M.bsz).transpose(0, 2, 1)
        want = np.concatenate([blocks[i] @ x[i * M.bsz:(i + 1) * M.bsz] for i in range(M.nrows)])
    ctx.compute_preconditioner(pk, prec_precision=0)
    got = ctx.apply_preconditioner(x)
    e32 = rel(got, want)
    say(f"  {kind} FP32 factors apply rel err {e32:.2e}")
    assert 1e-12 < e32 < 1e-4
    flag32, it32, _ = ctx.gmres(restart=30, tol=1e-10, maxit=500)
    x32 = ctx.trace_solution()
    ctx.compute_preconditioner(pk, prec_precision=1)
    assert rel(ctx.apply_preconditioner(x), want) < 1e-10
    flag64, it64, _ = ctx.gmres(restart=30, tol=1e-10, maxit=500)
    x64 = ctx.trace_solution()
    say(f"  gmres FP32 factors {it32} it, FP64 factors {it64} it, x rel {rel(x32, x64):.2e}")
    assert flag32 == flag64 == 0 and abs(it32 - it64) <= max(2, it64 // 10)
    assert rel(x32, x64) < 1e-8


@pytest.mark.parametrize("orthog", [0, 1])
def test_gmres_matches_oracle(oracle_mod, system, orthog):
    from hfem_amd.lib import PREC_BILU0
    O = oracle_mod
    prob, ctx, M, Rg = system
    o2u, u2o = O.mdf_order(M)
    ctx.set_ordering(o2u)
    ctx.compute_preconditioner(PREC_BILU0)
    flag, it, rev = ctx.gmres(restart=10, tol=1e-10, maxit=500, orthog=orthog)
    x = ctx.trace_solution()
    Mg, _ = O.bilu0_factor(M, o2u, u2o)
    xr, fr, itr, revr = O.gmres(M, Rg, O.PREC_BILU0, Mg, o2u, u2o, restart=10, tol=1e-10, maxit=500, orthog=orthog)
    say(f"  gmres orthog={orthog}: gpu {it} it flag {flag}, oracle {itr} it flag {fr}, x rel {rel(x, xr):.2e}")
    assert flag == fr == 0 and it == itr
    np.testing.assert_allclose(rev, revr, rtol=1e-6, atol=1e-12)
    assert rel(x, xr) < 1e-8
    assert rel(M.matvec(x), Rg) < 1e-6


@pytest.mark.parametrize("oblique", [False, True], ids=["axis_flow", "oblique_flow"])
def test_restarted_gmres_stall_belongs_to_the_system(oracle_mod, oblique):
    """GMRES(30) with BILU0 and approximate MDF on 162 p=3 tets (the solve that
    follows the element-path bench case).  With the freestream along the box
    axes the flow is tangential to four box faces, where the normal eigenvalue
    of the far-field flux crosses zero: the trace system is nearly singular and
    indefinite, and restarted GMRES stagnates -- on the CPU oracle exactly as on
    the GPU, with the same residual history.  With the flow turned off the axes
    the same mesh converges, in the same number of iterations on both."""
    from hfem_amd.lib import HDGContext, PREC_BILU0
    O = oracle_mod
    kw = dict(angle=0.3, elevation=0.2) if oblique else {}
    prob = make_problem("tet", 3, 3, seed=0, **kw)
    ctx = HDGContext(prob)
    ctx.assemble_linear_system()
    Hg, Rg = ctx.read_system()
    M = O.BSR.from_problem(prob, Hg)
    o2u, u2o = O.mdf_order(M)
    ctx.set_ordering(o2u)
    ctx.compute_preconditioner(PREC_BILU0)
    flag, it, rev = ctx.gmres(restart=30, tol=1e-8, maxit=90)
    ctx.close()
    Mg, _ = O.bilu0_factor(M, o2u, u2o)
    _, fr, itr, revr = O.gmres(M, Rg, O.PREC_BILU0, Mg, o2u, u2o, restart=30, tol=1e-8, maxit=90)
    say(f"  {'oblique' if oblique else 'axis'} flow: gpu {it} it flag {flag} final {rev[-1]:.3e}; "
        f"oracle {itr} it flag {fr} final {revr[-1]:.3e}")
    if oblique:
        assert flag == fr == 0 and it == itr
    else:
        assert flag == fr == 1 and it == itr == 90 and rev[-1] > 1e-5 and revr[-1] > 1e-5
    # (the nearly singular axis-flow system amplifies the rounding of two summation
    # orders to a few 1e-6 within the first cycle)
    np.testing.assert_allclose(rev[:30], revr[:30], rtol=1e-6 if oblique else 1e-4)
    np.testing.assert_allclose(rev, revr, rtol=1e-3)


@pytest.mark.parametrize("prec_precision", [1, 0], ids=["fp64_factors", "fp32_factors"])
def test_syncfree_sweeps_match_level_sweeps(oracle_mod, system, prec_precision):
    """The BILU0 factorisation and application with one persistent launch each
    (rows taken in level order, per-row ready flags, bounded waits) return the
    level-by-level results bit for bit, over repeated applications (the flags carry
    each launch's epoch) and inside a replayed graph; GMRES gives the same history."""
    from hfem_amd.lib import PREC_BILU0
    O = oracle_mod
    prob, ctx, M, Rg = system
    o2u, u2o = O.mdf_order(M)
    ctx.set_ordering(o2u)
    rng = np.random.default_rng(3)
    xs = [rng.standard_normal(Rg.size) for _ in range(3)]
    ctx.set_bilu0_sweep(0)
    ctx.compute_preconditioner(PREC_BILU0, prec_precision=prec_precision)
    ref = [ctx.apply_preconditioner(x) for x in xs]
    flag0, it0, rev0 = ctx.gmres(restart=10, tol=1e-10, maxit=500)
    try:
        ctx.set_bilu0_sweep(1)
        ctx.compute_preconditioner(PREC_BILU0, prec_precision=prec_precision)
        for use_graph in (False, True):
            ctx.set_apply_graph(use_graph)
            for rep in range(2):
                for x, r in zip(xs, ref):
                    assert np.array_equal(ctx.apply_preconditioner(x), r), (use_graph, rep)
        ctx.set_apply_graph(False)
        flag1, it1, rev1 = ctx.gmres(restart=10, tol=1e-10, maxit=500)
    finally:
        ctx.set_bilu0_sweep(1)
        ctx.set_apply_graph(False)
    say(f"  sync-free sweeps: gmres {it1} it (level sweeps {it0})")
    assert flag0 == flag1 == 0 and it0 == it1 and np.array_equal(rev0, rev1)


@pytest.mark.parametrize("case", [("tet", 4, 3), ("hex", 3, 2)], ids=lambda c: f"{c[0]}{c[1]}p{c[2]}")
def test_syncfree_sweeps_many_levels(case):
    """Sync-free against level-launched factorisation and sweeps where the level
    chain is long (384 p=3 tets, block 50; periodic p=2 hexes, block 45):
    bit-identical factors, applications and GMRES histories.  Then the give-up
    path, forced from the host with a one-poll wait limit, in the application and
    in the factorisation: the kernels still end, the call reports the failure, and
    the next one with the default limit is exact again."""
    from hfem_amd.lib import HDGContext, PREC_BILU0, HfemError
    kind, nc, p = case
    kw = dict(angle=0.3, elevation=0.2) if kind == "tet" else dict(state="taylor_green", amp=1e-3)
    prob = make_problem(kind, nc, p, seed=2, **kw)
    ctx = HDGContext(prob)
    try:
        ctx.assemble_linear_system()
        ctx.compute_ordering(1)
        _, nlev = ctx.get_ordering()
        n = prob.mesh.nf * prob.nch * prob.master.npf
        xs = [np.random.default_rng(s).standard_normal(n) for s in range(2)]
        ctx.set_bilu0_sweep(0)
        ctx.compute_preconditioner(PREC_BILU0)
        t_level = ctx.solver_stats()["factor_ms"]
        ref = [ctx.apply_preconditioner(x) for x in xs]
        f0, it0, rev0 = ctx.gmres(restart=30, tol=1e-8, maxit=200)
        ctx.set_bilu0_sweep(1)
        ctx.compute_preconditioner(PREC_BILU0)
        say(f"  factorisation: level launches {t_level:.1f} ms, sync-free {ctx.solver_stats()['factor_ms']:.1f} ms")
        for rep in range(3):
            for x, r in zip(xs, ref):
                assert np.array_equal(ctx.apply_preconditioner(x), r), rep
        f1, it1, rev1 = ctx.gmres(restart=30, tol=1e-8, maxit=200)
        say(f"  {kind}{nc} p{p}: levels {nlev}, gmres {it1} it (level sweeps {it0})")
        assert f0 == f1 and it0 == it1 and np.array_equal(rev0, rev1)
        ctx.set_bilu0_sweep(1, spin_limit=1)
        with pytest.raises(HfemError, match="gave up"):
            ctx.apply_preconditioner(xs[0])
        with pytest.raises(HfemError, match="gave up"):
            ctx.compute_preconditioner(PREC_BILU0)
        ctx.set_bilu0_sweep(1)
        ctx.compute_preconditioner(PREC_BILU0)
        assert np.array_equal(ctx.apply_preconditioner(xs[0]), ref[0])
    finally:
        ctx.close()


def test_update_and_residual_match_oracle(oracle_mod, system):
    import newton_oracle as N
    from hfem_amd.lib import PREC_BILU0
    O = oracle_mod
    prob, ctx, M, Rg = system
    ctx.compute_ordering(1)
    ctx.compute_preconditioner(PREC_BILU0)
    ctx.gmres(restart=30, tol=1e-12, maxit=500)
    x = ctx.trace_solution()
    ctx.update_solution(0.5)
    U, UH = ctx.get_solution()
    _, _, _, blocks = O.trace_system(prob)
    Ur, UHr = N.update_solution(prob, prob.UDG, prob.UH, prob.SH, blocks, x, 0.5)
    say(f"  update: UDG {rel(U, Ur):.2e} UH {rel(UH, UHr):.2e}")
    assert rel(UH, UHr) < 1e-14
    assert rel(U, Ur) < 1e-11
    r = ctx.residual_norm()
    rr = N.residual_norm(prob, Ur, UHr, prob.SH)
    say(f"  residual norm gpu {r:.6e} oracle {rr:.6e}")
    assert abs(r - rr) <= 1e-11 * rr


def test_newton_matches_oracle(oracle_mod):
    import newton_oracle as N
    from hfem_amd.lib import HDGContext
    prob = make_problem("tri", 3, 2, seed=5, angle=0.3)
    ctx = HDGContext(prob)
    conv, nit, hist = ctx.solve_steady(newton_tol=1e-11, newton_maxiter=6, restart=30, gmres_tol=1e-12, ordering=1)
    U, UH = ctx.get_solution()
    Ur, UHr, okr, hr = N.newton_solve(prob, prob.UDG, prob.UH, prob.SH, newton_tol=1e-11, maxiter=6, restart=30,
                                      gmres_tol=1e-12)
    say(f"  newton gpu {nit} it {hist[:, :2].tolist()}; oracle {[(h['old'], h['new']) for h in hr]}")
    assert conv and okr and nit == len(hr)
    for k, h in enumerate(hr):
        assert abs(hist[k, 0] - h["old"]) <= 1e-10 * h["old"] + 1e-12
        assert hist[k, 2] == h["gmres_iters"]
    assert rel(U, Ur) < 1e-10 and rel(UH, UHr) < 1e-10
    ctx.close()


def test_dirk_step_matches_oracle(oracle_mod):
    import newton_oracle as N
    from hfem_amd.lib import HDGContext
    prob = make_problem("tri", 2, 2, seed=9, angle=0.3)
    ctx = HDGContext(prob)
    conv, nit = ctx.dirk_step(0.05, 3, 3, newton_tol=1e-11, newton_maxiter=8, restart=30, gmres_tol=1e-12)
    U, UH = ctx.get_solution()
    Ur, UHr, SHr, hist = N.dirk_step(prob, prob.UDG, prob.UH, prob.SH, 0.05, newton_tol=1e-11, maxiter=8,
                                     restart=30, gmres_tol=1e-12)
    say(f"  dirk: newton its gpu {nit} oracle {sum(len(h) for h in hist)}; UDG {rel(U, Ur):.2e} UH {rel(UH, UHr):.2e}")
    assert conv
    assert rel(U, Ur) < 1e-10 and rel(UH, UHr) < 1e-10
    ctx.close()


def test_newton_line_search_retraction_matches_oracle(oracle_mod):
    """A state far from the solution (perturbation 0.5 on u and q) whose full
    Newton steps raise the residual: the line search retracts (alphas 1, -1/2 in
    the first iterations, solvers/dgsolvers.cpp:152-187).  The GPU driver's
    per-iteration line-search steps, accumulated step, GMRES counts and residual
    history equal the oracle's."""
    import newton_oracle as N
    from hfem_amd.lib import HDGContext
    prob = make_problem("tri", 2, 2, seed=5, angle=0.3, amp=0.5, qamp=0.5)
    kw = dict(restart=30, gmres_tol=1e-12)
    Ur, UHr, okr, hr = N.newton_solve(prob, prob.UDG, prob.UH, prob.SH, newton_tol=1e-10, maxiter=10, **kw)
    assert okr and any(len(h["alphas"]) > 1 for h in hr), "the case must exercise the retraction"
    ctx = HDGContext(prob)
    conv, nit, hist = ctx.solve_steady(newton_tol=1e-10, newton_maxiter=10, ordering=1, **kw)
    say(f"  retraction gpu steps {hist[:, 4].tolist()} applied {hist[:, 3].tolist()}; "
        f"oracle {[h['alphas'] for h in hr]}")
    assert conv and nit == len(hr)
    for k, h in enumerate(hr):
        assert hist[k, 4] == len(h["alphas"])
        assert hist[k, 3] == sum(h["alphas"])
        assert hist[k, 2] == h["gmres_iters"]
        assert abs(hist[k, 0] - h["old"]) <= 1e-10 * h["old"] + 1e-12
    U, UH = ctx.get_solution()
    assert rel(U, Ur) < 1e-10 and rel(UH, UHr) < 1e-10
    ctx.close()


def test_single_precision_preconditioner_newton():
    """Newton with single-precision BILU0 factors (precPrecision 0) takes the same
    iterations as with FP64 factors, the linear solves running to the same FP64
    tolerance; the converged state agrees to 1e-10."""
    from hfem_amd.lib import HDGContext
    prob = make_problem("tri", 3, 2, seed=5, angle=0.3)
    out = {}
    for pp in (1, 0):
        ctx = HDGContext(prob)
        conv, nit, hist = ctx.solve_steady(newton_tol=1e-11, newton_maxiter=8, restart=30, gmres_tol=1e-12,
                                           ordering=1, prec_precision=pp)
        U, UH = ctx.get_solution()
        out[pp] = (conv, nit, hist, U, UH)
        ctx.close()
    (c1, n1, h1, U1, UH1), (c0, n0, h0, U0, UH0) = out[1], out[0]
    say(f"  newton FP64 factors {h1[:, [0, 2]].tolist()} FP32 factors {h0[:, [0, 2]].tolist()}")
    assert c0 and c1 and n0 == n1
    assert rel(U0, U1) < 1e-10 and rel(UH0, UH1) < 1e-10


def test_mixed_newton_same_iterations_as_fp64():
    """The mixed condensation (FP32 factorisation, FP64 refinement) inside Newton:
    the same Newton and GMRES iteration counts as the FP64 path and a residual
    history within 1e-10 relative of it (above the 1e-12 level)."""
    from hfem_amd.lib import HDGContext, FP64, MIXED
    prob = make_problem("tet", 1, 3, seed=5, angle=0.3, elevation=0.2)
    runs = {}
    for prec in (FP64, MIXED):
        ctx = HDGContext(prob)
        conv, nit, hist = ctx.solve_steady(newton_tol=1e-11, newton_maxiter=8, restart=30, gmres_tol=1e-12,
                                           ordering=1, precision=prec)
        U, UH = ctx.get_solution()
        runs[prec] = (conv, nit, hist, U, UH)
        ctx.close()
    (c0, n0, h0, U0, UH0), (c1, n1, h1, U1, UH1) = runs[FP64], runs[MIXED]
    say(f"  newton fp64 {h0[:, [0, 2]].tolist()} mixed {h1[:, [0, 2]].tolist()}")
    assert c0 and c1 and n0 == n1
    assert np.array_equal(h0[:, 2], h1[:, 2]) and np.array_equal(h0[:, 4], h1[:, 4])
    for k in range(n0):
        assert abs(h1[k, 0] - h0[k, 0]) <= 1e-10 * h0[k, 0] + 1e-12
    assert rel(U1, U0) < 1e-10 and rel(UH1, UH0) < 1e-10


def test_mixed_dirk_step_same_iterations_as_fp64():
    from hfem_amd.lib import HDGContext, FP64, MIXED
    prob = make_problem("tri", 2, 3, seed=9, angle=0.3)
    runs = {}
    for prec in (FP64, MIXED):
        ctx = HDGContext(prob)
        conv, nit = ctx.dirk_step(0.05, 3, 3, newton_tol=1e-11, newton_maxiter=8, restart=30, gmres_tol=1e-12,
                                  precision=prec)
        runs[prec] = (conv, nit) + ctx.get_solution()
        ctx.close()
    (c0, n0, U0, UH0), (c1, n1, U1, UH1) = runs[FP64], runs[MIXED]
    say(f"  dirk newton its fp64 {n0} mixed {n1}; UDG {rel(U1, U0):.2e} UH {rel(UH1, UH0):.2e}")
    assert c0 and c1 and n0 == n1
    assert rel(U1, U0) < 1e-10 and rel(UH1, UH0) < 1e-10


def test_dirk_step_periodic_hex_matches_oracle(oracle_mod):
    """The DIRK(3,3) configuration on p=2 hexes at 3^3 periodic cells (the
    Taylor-Green state of the 24^3 benchmark case, 45 x 45 trace blocks, rows of
    11 blocks): Newton iteration counts per stage and the final state equal the
    oracle's."""
    import newton_oracle as N
    from hfem_amd.lib import HDGContext
    prob = make_problem("hex", 3, 2, seed=0, state="taylor_green", amp=1e-3)
    say(f"  hex3 p2 periodic: ne={prob.ne} nf={prob.mesh.nf}")
    Ur, UHr, SHr, hist = N.dirk_step(prob, prob.UDG, prob.UH, prob.SH, 0.1, newton_tol=1e-10, maxiter=8,
                                     restart=30, gmres_tol=1e-12)
    ctx = HDGContext(prob)
    conv, nit = ctx.dirk_step(0.1, 3, 3, newton_tol=1e-10, newton_maxiter=8, restart=30, gmres_tol=1e-12)
    U, UH = ctx.get_solution()
    say(f"  dirk hex: newton its gpu {nit} oracle {sum(len(h) for h in hist)}; "
        f"UDG {rel(U, Ur):.2e} UH {rel(UH, UHr):.2e}")
    assert conv and nit == sum(len(h) for h in hist)
    assert rel(U, Ur) < 1e-10 and rel(UH, UHr) < 1e-10
    ctx.close()
