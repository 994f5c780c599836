"""hfem_amd/binio.py pinned against the reference's own binary readers and
writers (preprocessing/readbinaryfiles.cpp, preprocessing/ioutilities.cpp),
compiled from the reference sources by oracle/ref.mk into
oracle/_ref/libhfemrefio.so (oracle/ref_io.cpp); skipped where that library is
not built (no reference tree).  CPU only.

* A serial case written in the layouts binio reads is read by the reference's
  readappstruct / readmasterstruct / readmeshstruct / readsolstruct and written
  back by its writers byte for byte, and binio reads the reference-written files
  to the same values: both sides agree on every field's position and count.
* solstruct::writeSol2File's output is read by binio (u part and full UDG) and
  binio's writer produces the same bytes.
* One rank's record of a parallel case (several ranks per mesh<k>.bin /
  sol<k>.bin, sliced by app.meshsize / app.solsize) reads the same through the
  reference's parallel branches and through binio.
"""
import ctypes
import os

import numpy as np
import pytest

from hfem_amd import binio
from hfem_amd.problem import make_problem
from test_binio import write_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_ref", "libhfemrefio.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="reference I/O harness not built")


@pytest.fixture(scope="module")
def ref():
    L = ctypes.CDLL(LIB)
    L.refio_roundtrip.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    L.refio_read_rank.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p]
    L.refio_write_sol.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 5 + [ctypes.c_void_p, ctypes.c_void_p,
                                                                            ctypes.c_int]
    return L


@pytest.mark.parametrize("kind,nc,p", [("tri", 2, 3), ("tet", 1, 2), ("hex", 3, 2)])
def test_reference_reads_and_rewrites_a_serial_case(tmp_path, ref, kind, nc, p):
    pr = make_problem(kind, nc, p, seed=4)
    src, dst = str(tmp_path / "a_"), str(tmp_path / "b_")
    write_case(src, pr, dict(newton_maxiter=5, newton_tol=1e-9, gmres_tol=1e-10))
    assert ref.refio_roundtrip(src.encode(), dst.encode()) == 0
    for f in ("app.bin", "master.bin", "mesh.bin", "sol.bin"):
        a = np.fromfile(src + f, dtype="<f8")
        b = np.fromfile(dst + f, dtype="<f8")
        assert a.size == b.size and np.array_equal(a, b), f
    # binio on the reference-written files
    app = binio.read_app_bin(dst + "app.bin")
    assert np.array_equal(app["physicsparam"], pr.param) and np.array_equal(app["bcs"], pr.ui)
    m = binio.read_mesh_bin(dst + "mesh.bin")
    back = binio.mesh_from_bin(m, pr.mesh.nd, pr.mesh.elemtype, pr.mesh.porder)
    assert np.array_equal(back.dgnodes, pr.mesh.dgnodes) and np.array_equal(back.elcon, pr.mesh.elcon)
    (mm,) = binio.read_master_bin(dst + "master.bin")
    t = binio.master_tables(mm)
    for k in ("shapvt", "shapvg", "shapft", "shapfg"):
        assert np.array_equal(t[k], pr.master.flat(k)), k
    s = binio.read_sol_bin(dst + "sol.bin")
    assert np.array_equal(s["UDG"], pr.UDG.ravel()) and np.array_equal(s["UH"], pr.UH.ravel())


@pytest.mark.parametrize("write_q", [0, 1])
def test_solution_output_of_write_sol2file(tmp_path, ref, write_q):
    pr = make_problem("tet", 1, 2, seed=6)
    d = pr.dims
    UDG = np.ascontiguousarray(pr.UDG)
    UH = np.ascontiguousarray(pr.UH).ravel()
    path = str(tmp_path / "ref_sol.bin")
    assert ref.refio_write_sol(path.encode(), pr.ne, d["npv"], d["nc"], d["ncu"], UH.size,
                               UDG.ctypes.data_as(ctypes.c_void_p), UH.ctypes.data_as(ctypes.c_void_p),
                               write_q) == 0
    U, uh, with_q = binio.read_solution_file(path, pr.ne, d["npv"], d["nc"], d["ncu"])
    assert with_q == bool(write_q)
    assert np.array_equal(U, UDG if write_q else UDG[:, :d["ncu"], :]) and np.array_equal(uh, UH)
    mine = str(tmp_path / "mine.bin")
    binio.write_solution_file(mine, UDG, pr.UH, d["ncu"], write_q=bool(write_q))
    assert open(mine, "rb").read() == open(path, "rb").read()


def _parallel_case(prefix, nproc, nfile, rng, hybrid=0):
    """A parallel case of nproc ranks in nfile files: app.bin with the
    meshsize / solsize offsets and mesh<k>.bin / sol<k>.bin holding one record per
    rank in the parallel layout (arbitrary content of consistent sizes: the
    readers check sizes, not meaning).  Returns the records written per rank."""
    per = nproc // nfile
    records = {}
    meshsize, solsize = [], []
    for k in range(1, nfile + 1):
        mfile, sfile, moff, soff = [], [], [0], [0]
        for fr in range(per):
            rank = (k - 1) * per + fr
            ne, nf, nfe, nve, nvf, npe, npf, ncd, nnbsd, ndh = (int(rng.integers(3, 7)), int(rng.integers(5, 9)), 4,
                                                               4, 3, 10, 6, 3, int(rng.integers(1, 4)),
                                                               int(rng.integers(8, 12)))
            head = np.zeros(30)
            head[:24] = [3, ncd, nfe, nve, nvf, ne, nf, 20, ndh, npe, npf, 20, 5, 15, 0, 5, nf, ne - 1, nf - 2, nf - 2,
                         nnbsd, 1, 2, 2]
            head[28], head[29] = hybrid, nproc
            ints = lambda *sh: rng.integers(-1, 50, size=sh).astype(float)
            parts = [head, ints(ne), ints(ne, nve), rng.standard_normal((ne, ncd, npe)), ints(ne, npf * nfe),
                     ints(ne, nfe), ints(ne, nfe), ints(ne, nfe), ints(nf, nvf + 3), ints(nf), ints(nnbsd), ints(ne),
                     ints(ne), ints(nf), ints(nf)]
            if hybrid:
                parts += [ints(ndh), ints(ndh)]
            rec = np.concatenate([np.asarray(q, float).ravel() for q in parts])
            nudg, nuh = int(rng.integers(10, 20)), int(rng.integers(5, 10))
            srec = np.concatenate([[nudg, nuh, 0] + [0] * 7, rng.standard_normal(nudg + nuh)])
            mfile.append(rec)
            sfile.append(srec)
            moff.append(moff[-1] + rec.size)
            soff.append(soff[-1] + srec.size)
            records[rank] = (rec, srec)
        np.concatenate(mfile).tofile(f"{prefix}mesh{k}.bin")
        np.concatenate(sfile).tofile(f"{prefix}sol{k}.bin")
        meshsize += moff
        solsize += soff
    arrays = [meshsize, solsize, [2], [0], [0], [4], [4], [0], [1], [1.0] * 5, [1], [0.0], [0.0], [0] * 24,
              [0.0, 1.0, 0.0], [0] * 16, [1.4, 0.0, 1000.0, 0.72, 0.3, 2.0], [1e-8, 1e-9]]
    head = np.zeros(40)
    head[:18] = [len(a) for a in arrays]
    head[18:28] = [nproc, nfile, 3, 3, 20, 5, 5, 15, 0, 0]
    np.concatenate([head] + [np.asarray(a, float) for a in arrays]).tofile(prefix + "app.bin")
    return records


@pytest.mark.parametrize("nproc,nfile,hybrid", [(2, 1, 0), (4, 2, 0), (3, 3, 1)])
def test_parallel_rank_records(tmp_path, ref, nproc, nfile, hybrid):
    rng = np.random.default_rng(nproc * 10 + nfile)
    prefix = str(tmp_path / "p_")
    _parallel_case(prefix, nproc, nfile, rng, hybrid)
    app = binio.read_app_bin(prefix + "app.bin")
    for rank in range(nproc):
        out = str(tmp_path / f"r{rank}_")
        assert ref.refio_read_rank(prefix.encode(), rank, out.encode()) == 0
        raw = lambda name: np.fromfile(out + name + ".raw", dtype="<f8")
        m = binio.read_mesh_bin_rank(prefix, app, rank)
        s = binio.read_sol_bin_rank(prefix, app, rank)
        d = m["ndims"]
        assert raw("mesh_ndims")[5] == d["ne"] and raw("mesh_ndims")[20] == d["nnbsd"]
        for name in ("elementtype", "t", "elcon", "bf", "t2f", "t2t", "nbsd", "extintelem", "elem2cpu", "extintface",
                     "face2cpu"):
            assert np.array_equal(raw(name), m[name].ravel()), (rank, name)
        assert np.array_equal(raw("dgnodes"), m["dgnodes"].ravel())
        assert np.array_equal(raw("UDG"), s["UDG"]) and np.array_equal(raw("UH"), s["UH"])
    with pytest.raises(ValueError):
        binio.read_mesh_bin_rank(prefix, app, nproc)
