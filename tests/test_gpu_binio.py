"""A reference sol.bin loaded into the library, and the solution written back in
the writeSol2File format (GPU)."""
import numpy as np
import pytest

from hfem_amd.binio import load_sol_bin, read_solution_file, write_sol_bin, write_solution_file
from hfem_amd.problem import make_problem

pytestmark = pytest.mark.gpu


def test_sol_bin_into_a_context_and_back(tmp_path):
    from hfem_amd.lib import HDGContext
    prob = make_problem("tet", 1, 2, seed=4)
    rng = np.random.default_rng(1)
    UDG = prob.UDG + 1e-3 * rng.standard_normal(prob.UDG.shape)
    UH = prob.UH + 1e-3 * rng.standard_normal(prob.UH.shape)
    write_sol_bin(tmp_path / "sol.bin", UDG, UH)
    ctx = HDGContext(prob)
    load_sol_bin(ctx, tmp_path / "sol.bin")
    U, H = ctx.get_solution()
    assert np.array_equal(U, UDG) and np.array_equal(H, UH)
    write_solution_file(tmp_path / "out.bin", U, H, prob.ncu)
    ne, nc, npv = U.shape
    Uu, Hr, with_q = read_solution_file(tmp_path / "out.bin", ne, npv, nc, prob.ncu)
    assert not with_q and np.array_equal(Uu, UDG[:, :prob.ncu]) and np.array_equal(Hr, UH.ravel())
    ctx.close()


def test_mesh_bin_assembles_like_the_original(tmp_path):
    """A problem on the mesh read back from a serial mesh.bin assembles and
    condenses to the same blocks, bit for bit."""
    import dataclasses
    import sys, os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_binio import write_mesh_file
    from hfem_amd.binio import mesh_from_bin, read_mesh_bin
    from hfem_amd.lib import HDGContext, FP64
    prob = make_problem("tet", 1, 3, seed=5)
    d = prob.dims
    write_mesh_file(tmp_path / "mesh.bin", prob.mesh, d["npv"], d["ncu"], d["nch"], d["nc"])
    msh = mesh_from_bin(read_mesh_bin(tmp_path / "mesh.bin"), prob.mesh.nd, prob.mesh.elemtype, prob.mesh.porder)
    prob2 = dataclasses.replace(prob, mesh=msh)
    out = []
    for pr in (prob, prob2):
        ctx = HDGContext(pr)
        ctx.assemble_element_matrix_vector()
        ctx.schur_element_matrix_vector(precision=FP64)
        out.append(ctx.read_condensed())
        ctx.close()
    for k in ("AH", "FH", "DinvF", "DinvRu"):
        assert np.array_equal(out[0][k], out[1][k]), k


def test_master_bin_tables_drive_the_library(tmp_path):
    """A context given the tables read from a master.bin condenses to the same
    blocks as one built on the problem's own master element, bit for bit."""
    import sys, os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_binio import write_master_file
    from hfem_amd.binio import master_tables, read_master_bin
    from hfem_amd.lib import HDGContext, FP64
    prob = make_problem("tri", 2, 3, seed=6)
    write_master_file(tmp_path / "master.bin", prob.master)
    tables = master_tables(read_master_bin(tmp_path / "master.bin")[0])
    out = []
    for mt in (None, tables):
        ctx = HDGContext(prob, master_tables=mt)
        ctx.assemble_element_matrix_vector()
        ctx.schur_element_matrix_vector(precision=FP64)
        out.append(ctx.read_condensed())
        ctx.close()
    for k in ("AH", "FH", "DinvF", "DinvRu"):
        assert np.array_equal(out[0][k], out[1][k]), k


def test_reference_case_runs_like_the_original(tmp_path):
    """A problem read back from the four files of a reference case condenses and
    takes Newton steps exactly as the problem the files were written from, with
    the solver parameters the app file carries."""
    import sys, os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_binio import write_case
    from hfem_amd.binio import problem_from_case
    from hfem_amd.lib import HDGContext, FP64
    pr = make_problem("tri", 3, 2, seed=5, angle=0.3)
    write_case(str(tmp_path / "c_"), pr, dict(newton_maxiter=4, newton_tol=1e-10, gmres_tol=1e-12))
    p2, tables, params = problem_from_case(str(tmp_path / "c_"))
    res = []
    for p, mt in ((pr, None), (p2, tables)):
        ctx = HDGContext(p, master_tables=mt)
        ctx.assemble_element_matrix_vector()
        ctx.schur_element_matrix_vector(precision=FP64)
        c = ctx.read_condensed()
        conv, nit, hist = ctx.solve_steady(**params)
        U, UH = ctx.get_solution()
        res.append((c, nit, hist, U, UH))
        ctx.close()
    (c0, n0, h0, U0, H0), (c1, n1, h1, U1, H1) = res
    for k in ("AH", "FH"):
        assert np.array_equal(c0[k], c1[k]), k
    assert n0 == n1 and np.array_equal(h0[:, :3], h1[:, :3])
    assert np.array_equal(U0, U1) and np.array_equal(H0, H1)
