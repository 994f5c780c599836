"""The reference's binary solution files (CPU): sol.bin as the preprocessor
writes it (preprocessing/digasopre.m:126-129, :283-297) and readsolstruct reads
it (preprocessing/readbinaryfiles.cpp:624-684), and the writeSol2File output
(preprocessing/ioutilities.cpp:256-300)."""
import numpy as np
import pytest

from hfem_amd.binio import (read_sol_bin, read_solution_file, write_sol_bin, write_solution_file)
from hfem_amd.problem import make_problem


@pytest.fixture(scope="module")
def prob():
    return make_problem("tri", 2, 2, seed=3)


@pytest.mark.parametrize("endian", ["<", ">"])
def test_sol_bin_layout_and_round_trip(tmp_path, prob, endian):
    path = tmp_path / "sol.bin"
    PDG = np.arange(7.0)
    write_sol_bin(path, prob.UDG, prob.UH, PDG, endian=endian)
    raw = np.fromfile(path, dtype=endian + "f8")
    # ten header doubles [numel(UDG), numel(UH), numel(PDG), 0 ...], then the arrays
    assert raw.size == 10 + prob.UDG.size + prob.UH.size + 7
    assert list(raw[:10]) == [prob.UDG.size, prob.UH.size, 7, 0, 0, 0, 0, 0, 0, 0]
    # UDG in the reference layout: element e, component c, node n at n + npv*(c + nc*e)
    ne, nc, npv = prob.UDG.shape
    e, c, n = ne - 1, nc - 2, npv - 1
    assert raw[10 + n + npv * (c + nc * e)] == prob.UDG[e, c, n]
    s = read_sol_bin(path, endian=endian)
    assert np.array_equal(s["UDG"], prob.UDG.ravel())
    assert np.array_equal(s["UH"], prob.UH.ravel())
    assert np.array_equal(s["PDG"], PDG)


def test_sol_bin_without_pdg_and_checks(tmp_path, prob):
    path = tmp_path / "sol.bin"
    write_sol_bin(path, prob.UDG, prob.UH)
    s = read_sol_bin(path)
    assert s["PDG"].size == 0 and s["ndims"][2] == 0
    # the wrong byte order is caught by the header, a short file by the counts
    with pytest.raises(ValueError):
        read_sol_bin(path, endian=">")
    raw = np.fromfile(path, dtype="<f8")
    raw[:-3].tofile(tmp_path / "short.bin")
    with pytest.raises(ValueError, match="do not match"):
        read_sol_bin(tmp_path / "short.bin")


@pytest.mark.parametrize("write_q", [False, True])
def test_solution_output_layout(tmp_path, prob, write_q):
    """writeSol2File: [len, n], then per element the first npv*ncu values of its
    npv*nc block (DCOPY of UDG[i*npv*nc]) or every block whole, then UH."""
    path = tmp_path / "out.bin"
    ne, nc, npv = prob.UDG.shape
    ncu = prob.ncu
    write_solution_file(path, prob.UDG, prob.UH, ncu, write_q=write_q)
    raw = np.fromfile(path, dtype="<f8")
    per = npv * (nc if write_q else ncu)
    assert raw[0] == ne * per and raw[1] == prob.UH.size
    for e in (0, ne - 1):
        assert np.array_equal(raw[2 + e * per:2 + (e + 1) * per], prob.UDG[e].ravel()[:per])
    assert np.array_equal(raw[2 + ne * per:], prob.UH.ravel())
    U, UH, with_q = read_solution_file(path, ne, npv, nc, ncu)
    assert with_q == write_q
    assert np.array_equal(U, prob.UDG if write_q else prob.UDG[:, :ncu])
    assert np.array_equal(UH, prob.UH.ravel())


def test_solution_output_rejects_other_sizes(tmp_path, prob):
    path = tmp_path / "out.bin"
    ne, nc, npv = prob.UDG.shape
    write_solution_file(path, prob.UDG, prob.UH, prob.ncu)
    with pytest.raises(ValueError):
        read_solution_file(path, ne + 1, npv, nc, prob.ncu)
