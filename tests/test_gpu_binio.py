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
