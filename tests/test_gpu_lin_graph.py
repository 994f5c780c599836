"""Linearisation graph (hfem_set_linearization_graph, GPU): the FP64 device work of
assemble_linear_system -- fluxes, element assembly, the residual norm's sums,
condensation, BSR scatter -- captured once and replayed.

* Replays give bit for bit what the kernel-by-kernel launches give, over several
  replays on changed states with eager element work between them, and after a
  physics change (which recaptures).
* A Newton solve and a DIRK(3,3) step on the periodic p=2 hexes (the DIRK
  benchmark's element type) run from the graph with the same iterations and the
  same final state as without it.
"""
import numpy as np
import pytest

from hfem_amd.problem import make_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("products", ["hand-written", "rocblas", "fused"])
def test_replays_equal_kernel_by_kernel_launches(products):
    from hfem_amd.lib import HDGContext
    pr = make_problem("tet", 2, 2, seed=4)
    say(f"  tet2 p2: ne={pr.ne}, products {products}")
    a = HDGContext(pr)
    b = HDGContext(pr)
    # condensation choices {inverse FP64, FP32, X, Schur, mixed ...}: -1 rocBLAS, 64 the fused solve
    x, schur = {"hand-written": (0, 0), "rocblas": (-1, -1), "fused": (64, 4)}[products]
    for c in (a, b):
        c.set_condensation_choice([1, 1, x, schur, 0, 0, 0])
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
        ra = a.assemble_linear_system()
        rb = b.assemble_linear_system()
        Hga, Rga = a.read_system()
        Hgb, Rgb = b.read_system()
        say(f"  replay {it}: norm {ra:.6e} / {rb:.6e}")
        assert ra == rb
        assert np.array_equal(Hga, Hgb) and np.array_equal(Rga, Rgb)
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
