"""Linearisation graph (hfem_set_linearization_graph, GPU): the FP64 device work of
assemble_linear_system -- fluxes, element assembly, the residual norm's sums,
condensation, BSR scatter -- captured once and replayed.

* FP64 replays give bit for bit what the kernel-by-kernel launches give, over
  several replays on changed states with eager element work between them, and
  after a physics change (which recaptures).  The mixed condensation is
  captured with a fixed number of refinement sweeps, the count the last
  condensation needed, where the kernel-by-kernel one stops adaptively: a
  replay may run one sweep more than the adaptive loop would (its result is
  then the better refined one), so mixed replays are compared within the
  mixed path's 1e-10 bound; an element that would need more sweeps than the
  graph runs makes the library redo the iteration kernel by kernel.
* A Newton solve and a DIRK(3,3) step on the periodic p=2 hexes (the DIRK
  benchmark's element type) run from the graph with the same iterations and the
  same final state as without it; so does the mixed configuration (FP32-MFMA
  condensation with FP64 refinement, captured with the sweep count the first
  kernel-by-kernel condensation needed, and FP32 BILU0 factors), which also
  keeps the FP64 path's Newton iteration count.
"""
import numpy as np
import pytest

from hfem_amd.problem import make_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("products", ["hand-written", "rocblas", "fused"])
def test_replays_equal_kernel_by_kernel_launches(products, precision):
    from hfem_amd.lib import HDGContext, FP64, MIXED
    prec = FP64 if precision == "fp64" else MIXED
    pr = make_problem("tet", 2, 2, seed=4)
    say(f"  tet2 p2: ne={pr.ne}, products {products}, {precision}")
    a = HDGContext(pr)
    b = HDGContext(pr)
    # condensation choices {inverse FP64, FP32, X, Schur, mixed ...}: -1 rocBLAS, 64 the fused solve
    x, schur = {"hand-written": (0, 0), "rocblas": (-1, -1), "fused": (64, 4)}[products]
    mixed = {"hand-written": (4, 4, 4), "rocblas": (4, -1, 4), "fused": (64, 4, 4)}[products]
    for c in (a, b):
        c.set_condensation_choice([1, 1, x, schur, *mixed])
    a.set_linearization_graph(False)
    b.set_linearization_graph(True)
    rng = np.random.default_rng(0)
    for it in range(5):
        UDG = pr.UDG * (1 + 1e-3 * rng.standard_normal(pr.UDG.shape))
        UH = pr.UH * (1 + 1e-3 * rng.standard_normal(pr.UH.shape))
        a.set_solution(UDG, UH)
        b.set_solution(UDG, UH)
        if it == 3:
            a.set_physics(0.5)
            b.set_physics(0.5)
        ra = a.assemble_linear_system(prec)
        rb = b.assemble_linear_system(prec)
        Hga, Rga = a.read_system()
        Hgb, Rgb = b.read_system()
        say(f"  replay {it}: norm {ra:.6e} / {rb:.6e}")
        assert ra == rb
        if prec == FP64:
            assert np.array_equal(Hga, Hgb) and np.array_equal(Rga, Rgb)
        else:
            eh = np.linalg.norm(Hga - Hgb) / np.linalg.norm(Hga)
            er = np.linalg.norm(Rga - Rgb) / np.linalg.norm(Rga)
            say(f"    mixed: Hg {eh:.1e} Rg {er:.1e}")
            assert eh < 1e-10 and er < 1e-10
        if prec == MIXED:
            ca, cb = a.condense_report(), b.condense_report()
            say(f"    mixed sweeps {ca} / {cb}")
        # kernel-by-kernel element work on the graph's context between replays
        b.assemble_element_matrix_vector()
        b.schur_element_matrix_vector()
        b.sync()
    a.close()
    b.close()


def test_newton_and_dirk_from_graphs():
    from hfem_amd.lib import HDGContext
    pr = make_problem("tri", 3, 2, seed=5, angle=0.3)
    res = []
    for use in (False, True):
        ctx = HDGContext(pr)
        ctx.set_linearization_graph(use)
        res.append(ctx.solve_steady(newton_tol=1e-11, newton_maxiter=6, restart=30, gmres_tol=1e-12, ordering=1))
        ctx.close()
    (c0, n0, h0), (c1, n1, h1) = res
    say(f"  newton: {n0} / {n1} iterations, final residual {h0[-1, 1]:.3e} / {h1[-1, 1]:.3e}")
    assert c0 and c1 and n0 == n1 and np.array_equal(h0, h1)

    pr = make_problem("hex", 3, 2, seed=0, state="taylor_green", amp=1e-3)
    out = []
    for use in (False, True):
        ctx = HDGContext(pr)
        ctx.set_linearization_graph(use)
        conv, nit = ctx.dirk_step(0.1, 3, 3, newton_tol=1e-10, newton_maxiter=8, restart=30, gmres_tol=1e-12)
        out.append((conv, nit) + tuple(ctx.get_solution()))
        ctx.close()
    say(f"  dirk hex3 p2: Newton iterations {out[0][1]} / {out[1][1]}")
    assert out[0][0] and out[1][0] and out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2]) and np.array_equal(out[0][3], out[1][3])


def test_mixed_dirk_from_graphs():
    """The DIRK benchmark's mixed configuration -- FP32-MFMA condensation with FP64
    refinement and FP32 BILU0 factors -- from graphs and kernel by kernel: same
    Newton iterations, the same final state within 1e-10, and the plain FP64
    path's iteration count."""
    from hfem_amd.lib import HDGContext, FP64, MIXED
    pr = make_problem("hex", 3, 2, seed=0, state="taylor_green", amp=1e-3)
    kw = dict(newton_tol=1e-10, newton_maxiter=8, restart=30, gmres_tol=1e-12)
    out = []
    for use, prec, pp in ((False, MIXED, 0), (True, MIXED, 0), (True, FP64, 1)):
        ctx = HDGContext(pr)
        ctx.set_linearization_graph(use)
        conv, nit = ctx.dirk_step(0.1, 3, 3, precision=prec, prec_precision=pp, **kw)
        out.append((conv, nit) + tuple(ctx.get_solution()))
        ctx.close()
    say(f"  mixed dirk hex3 p2: Newton iterations {out[0][1]} (kernel by kernel) / {out[1][1]} (graphs) / "
        f"{out[2][1]} (FP64)")
    assert all(o[0] for o in out) and out[0][1] == out[1][1] == out[2][1]
    rg = np.linalg.norm(out[1][2] - out[0][2]) / np.linalg.norm(out[0][2])
    rel = np.linalg.norm(out[1][2] - out[2][2]) / np.linalg.norm(out[2][2])
    say(f"  mixed state, graphs vs kernel by kernel: {rg:.2e}; mixed vs FP64: {rel:.2e}")
    assert rg < 1e-10 and rel < 1e-9
