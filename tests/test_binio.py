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


def write_mesh_file(path, msh, npe, ncu, nch, nc):
    """A serial mesh.bin in readmeshstruct's order (the preprocessor's writer,
    preprocessing/digasopre.m:266-278), from one of the build's meshes; f and
    isEDGface are written as zeros (the library does not read them)."""
    ne, nve = msh.t.shape
    nfe = msh.t2f.shape[1]
    npf = msh.elcon.shape[2]
    nvf = 2 if msh.nd == 2 else (3 if msh.elemtype == 0 else 4)
    nf = msh.nf
    head = np.zeros(30)
    head[:16] = [msh.nd, msh.nd, nfe, nve, nvf, ne, nf, msh.p.shape[0], nf * npf, npe, npf, nc, ncu, nc - ncu, 0,
                 nch]
    head[28], head[29] = 0, 1                                   # HDG, one rank
    t2t = -np.ones((ne, nfe))
    for e1, a1, e2, a2 in msh.f2e:
        if e2 >= 0:
            t2t[e1, a1], t2t[e2, a2] = e2, e1
    parts = [head, np.zeros(ne), msh.t.ravel(), np.transpose(msh.dgnodes, (0, 2, 1)).ravel(), msh.elcon.ravel(),
             msh.bf.ravel(), msh.t2f.ravel(), t2t.ravel(), np.zeros(nf * (nvf + 3)), np.zeros(nf)]
    np.concatenate([np.asarray(p, dtype=np.float64).ravel() for p in parts]).tofile(path)
    # the preprocessor's size formula for the serial file
    return 30 + (1 + nve + npe * msh.nd + npf * nfe + nfe + nfe + nfe) * ne + (nvf + 3) * nf + nf


@pytest.mark.parametrize("kind,nc_,p", [("tri", 2, 2), ("tet", 1, 3), ("hex", 3, 2)])
def test_mesh_bin_to_library_mesh(tmp_path, kind, nc_, p):
    from hfem_amd.binio import mesh_from_bin, read_mesh_bin
    pr = make_problem(kind, nc_, p, seed=2)
    msh = pr.mesh
    d = pr.dims
    size = write_mesh_file(tmp_path / "mesh.bin", msh, d["npv"], d["ncu"], d["nch"], d["nc"])
    assert np.fromfile(tmp_path / "mesh.bin", dtype="<f8").size == size
    m = read_mesh_bin(tmp_path / "mesh.bin")
    assert m["ndims"]["ne"] == pr.ne and m["ndims"]["npf"] == d["npf"]
    back = mesh_from_bin(m, msh.nd, msh.elemtype, msh.porder)
    assert np.array_equal(back.dgnodes, msh.dgnodes)
    for k in ("elcon", "bf", "t2f", "f2e", "ent2entStart", "ent2ent", "t"):
        assert np.array_equal(getattr(back, k), getattr(msh, k)), k
    # a file whose dimensions do not account for its length is refused
    np.fromfile(tmp_path / "mesh.bin", dtype="<f8")[:-1].tofile(tmp_path / "short.bin")
    with pytest.raises(ValueError, match="need"):
        read_mesh_bin(tmp_path / "short.bin")


def write_master_file(path, m):
    """One element type of master.bin in readmasterstruct's order (the
    preprocessor's writemaster), from the build's master element; the tables the
    library does not use (shapmv, the R variants' own rules, permnode, permedge)
    are zeros or copies of the standard ones, of the sizes the reader expects."""
    nd, npv, ngv, npf, ngf, nfe = m.nd, m.npv, m.ngv, m.npf, m.ngf, m.nfe
    nvf = nd if m.elemtype == 0 else 2 ** (nd - 1)
    nve = nd + 1 if m.elemtype == 0 else 2 ** nd
    head = np.zeros(20)
    head[:11] = [nd, m.elemtype, 0, npv, npv, ngv, ngv, nfe, nve, 0, 0]
    F = lambda name: m.flat(name)
    parts = [head, [npf] * nfe, [npf] * nfe, [ngf] * nfe, [ngf] * nfe, [nvf] * nfe,
             F("plocvl"), F("gpvl"), F("gwvl"), F("shapvt"), F("shapvg"), np.zeros(npv * ngv * (nd + 1)),
             F("shapvgdotshapvl"), F("gpvl"), F("gwvl"), F("shapvt"), F("shapvg"), np.zeros(npv * ngv * (nd + 1)),
             np.zeros(nve)]
    for j in range(nfe):
        parts += [np.zeros(npf * (nd - 1)), F("gpfc"), F("gwfc"), F("shapft"), F("shapfg"), np.zeros(npf * ngf * nd),
                  F("shapfgdotshapfc"), F("gpfc"), F("gwfc"), F("shapft"), F("shapfg"), np.zeros(npf * ngf * nd),
                  m.perm[:, j], m.perm[:, j], np.arange(nvf)]
    np.concatenate([np.asarray(p, dtype=np.float64).ravel() for p in parts]).tofile(path)


@pytest.mark.parametrize("kind,nc_,p", [("tri", 2, 3), ("tet", 1, 2), ("hex", 3, 2)])
def test_master_bin_tables(tmp_path, kind, nc_, p):
    from hfem_amd.binio import master_tables, read_master_bin
    m = make_problem(kind, nc_, p, seed=0).master
    write_master_file(tmp_path / "master.bin", m)
    (mm,) = read_master_bin(tmp_path / "master.bin")
    assert (mm["npe"], mm["nge"], mm["nfe"], int(mm["npf"][0]), int(mm["ngf"][0])) == (m.npv, m.ngv, m.nfe, m.npf,
                                                                                          m.ngf)
    t = master_tables(mm)
    for k in ("shapvt", "shapvg", "shapft", "shapfg"):
        assert np.array_equal(t[k], m.flat(k)), k
    assert np.array_equal(t["perm"], m.perm.reshape(-1, order="F"))
    raw = np.fromfile(tmp_path / "master.bin", dtype="<f8")
    raw[:-1].tofile(tmp_path / "short.bin")
    with pytest.raises(ValueError):
        read_master_bin(tmp_path / "short.bin")


def test_app_bin_and_its_settings(tmp_path):
    """app.bin in readappstruct's order (as writeappstruct writes it), and the
    reference's meaning of its flags (initializestructs.cpp) as solver
    parameters the library takes."""
    from hfem_amd.binio import app_settings, read_app_bin
    from hfem_amd.lib import hfem_solver_params
    flag = np.zeros(24)
    flag[0], flag[10], flag[15], flag[16], flag[17], flag[18], flag[19], flag[20] = 1, 1, 1, 2, 0, 0, 0, 1
    problem = np.zeros(16)
    problem[3], problem[4], problem[13] = 3, 3, 7
    arrays = [[0, 123], [0, 45], [3], [0], [0], [6], [6], [0], [1, 2], [0.5, 0.25], [1], [1.0, 0.0, 0.0, 0.0, 2.5],
              [0.01, 0.02], flag, [4.0, 1.0, 0.0], problem, [1.4, 0.0, 1000.0, 0.72, 0.3, 2.0], [1e-9, 1e-11]]
    head = np.zeros(40)
    head[:18] = [len(a) for a in arrays]
    head[18:28] = [1, 1, 3, 3, 20, 5, 5, 15, 0, 0]
    np.concatenate([head] + [np.asarray(a, dtype=np.float64) for a in arrays]).tofile(tmp_path / "app.bin")
    app = read_app_bin(tmp_path / "app.bin")
    assert app["nd"] == 3 and app["ncu"] == 5 and list(app["porder"]) == [3] and list(app["bcm"]) == [1, 2]
    assert np.array_equal(app["physicsparam"], [1.4, 0.0, 1000.0, 0.72, 0.3, 2.0])
    st = app_settings(app)
    assert st["fc_u"] == 4.0 and st["fc_q"] == 1.0 and st["tdep"] == 1
    assert st["dirk_order"] == 3 and st["dirk_stages"] == 3 and np.array_equal(st["dt"], [0.01, 0.02])
    p = st["solver_params"]
    assert (p["newton_tol"], p["gmres_tol"], p["newton_maxiter"], p["precond"], p["orthog"], p["ordering"]) == \
        (1e-9, 1e-11, 7, 2, 0, 1)
    assert (p["quasi_newton"], p["qn_accuracy"], p["reuse_ordering"]) == (1, 0, 1)
    hfem_solver_params.make(**p)                   # the library's parameter block takes them as they are
    np.fromfile(tmp_path / "app.bin", dtype="<f8")[:-1].tofile(tmp_path / "short.bin")
    with pytest.raises(ValueError, match="need"):
        read_app_bin(tmp_path / "short.bin")


def write_case(prefix, pr, params):
    """The four files of a serial reference case for one of the build's problems."""
    from hfem_amd.binio import write_sol_bin
    d = pr.dims
    write_mesh_file(prefix + "mesh.bin", pr.mesh, d["npv"], d["ncu"], d["nch"], d["nc"])
    write_master_file(prefix + "master.bin", pr.master)
    write_sol_bin(prefix + "sol.bin", pr.UDG, pr.UH)
    flag = np.zeros(24)
    flag[0], flag[10], flag[19], flag[20] = pr.tdep, 1, 1, 1
    problem = np.zeros(16)
    problem[1] = 1                                  # appname: Navier-Stokes
    problem[13] = params["newton_maxiter"]
    for (arr, i), v in params.get("override", {}).items():
        if arr == "problem":
            problem[i] = v
        elif arr == "flag":
            flag[i] = v
    bcm = params.get("bcm", [1])
    arrays = [[0], [0], [pr.mesh.porder], [pr.mesh.elemtype], [0], [2 * pr.mesh.porder], [2 * pr.mesh.porder], [0],
              bcm, list(pr.ui), [1], [0.0], [0.0], flag, [pr.fc_u, pr.fc_q, 0.0], problem,
              list(params.get("param", pr.param)), [params["newton_tol"], params["gmres_tol"]]]
    head = np.zeros(40)
    head[:18] = [len(a) for a in arrays]
    head[18:28] = [1, 1, pr.nd, pr.nd, pr.nc, pr.ncu, pr.nch, pr.nc - pr.ncu, 0, 0]
    np.concatenate([head] + [np.asarray(a, dtype=np.float64) for a in arrays]).tofile(prefix + "app.bin")


def test_problem_from_case(tmp_path):
    from hfem_amd.binio import problem_from_case
    pr = make_problem("tet", 1, 2, seed=3)
    write_case(str(tmp_path / "c_"), pr, dict(newton_maxiter=9, newton_tol=1e-8, gmres_tol=1e-9))
    p2, tables, params = problem_from_case(str(tmp_path / "c_"))
    assert np.array_equal(p2.UDG, pr.UDG) and np.array_equal(p2.UH, pr.UH)
    assert np.array_equal(p2.param, pr.param) and np.array_equal(p2.ui, pr.ui)
    assert (p2.fc_u, p2.fc_q, p2.tdep) == (pr.fc_u, pr.fc_q, pr.tdep)
    assert np.array_equal(p2.mesh.elcon, pr.mesh.elcon) and np.array_equal(p2.mesh.dgnodes, pr.mesh.dgnodes)
    assert p2.dims == pr.dims
    assert np.array_equal(tables["shapvt"], pr.master.flat("shapvt"))
    assert (params["newton_maxiter"], params["newton_tol"], params["gmres_tol"]) == (9, 1e-8, 1e-9)


@pytest.mark.parametrize("arr,i,v", [("problem", 0, 1), ("problem", 1, 0), ("problem", 5, 1), ("problem", 6, 2),
                                     ("problem", 7, 1), ("problem", 8, 1), ("problem", 9, 1), ("problem", 10, 1),
                                     ("problem", 11, 1), ("flag", 1, 1)])
def test_problem_from_case_refuses_other_physics(tmp_path, arr, i, v):
    """A case that selects physics or a scheme the element path does not have
    (another appname, LF/Roe tau, a viscosity model, SGS/ALE/AV, the wave form,
    EDG) is refused instead of being solved as constant-tau far-field NS."""
    from hfem_amd.binio import problem_from_case
    pr = make_problem("tet", 1, 2, seed=3)
    write_case(str(tmp_path / "c_"), pr, dict(newton_maxiter=9, newton_tol=1e-8, gmres_tol=1e-9,
                                              override={(arr, i): v}))
    with pytest.raises(ValueError, match=rf"{arr}\[{i}\]"):
        problem_from_case(str(tmp_path / "c_"))


def test_problem_from_case_refuses_wall_and_short_parameters(tmp_path):
    from hfem_amd.binio import problem_from_case
    pr = make_problem("tet", 1, 2, seed=3)
    base = dict(newton_maxiter=9, newton_tol=1e-8, gmres_tol=1e-9)
    write_case(str(tmp_path / "w_"), pr, dict(base, bcm=[2]))          # a wall boundary
    with pytest.raises(ValueError, match="far field"):
        problem_from_case(str(tmp_path / "w_"))
    write_case(str(tmp_path / "s_"), pr, dict(base, param=list(pr.param[:5])))
    with pytest.raises(ValueError, match="physicsparam"):
        problem_from_case(str(tmp_path / "s_"))


def test_mesh_bin_with_extra_coordinate_components(tmp_path):
    """ncd > nd (extra dgnodes components, digasopre.m:97): only the first nd
    coordinates are the geometry; fewer than nd are refused."""
    from hfem_amd.binio import mesh_from_bin, read_mesh_bin
    pr = make_problem("tri", 2, 2, seed=2)
    msh, d = pr.mesh, pr.dims
    write_mesh_file(tmp_path / "mesh.bin", msh, d["npv"], d["ncu"], d["nch"], d["nc"])
    raw = np.fromfile(tmp_path / "mesh.bin", dtype="<f8")
    ne, npv, nve = pr.ne, d["npv"], msh.t.shape[1]
    o = 30 + ne + ne * nve
    dg = raw[o:o + ne * 2 * npv].reshape(ne, 2, npv)
    extra = np.concatenate([dg, np.full((ne, 1, npv), 7.5)], axis=1)   # a third, unrelated component
    head = raw[:30].copy()
    head[1] = 3
    np.concatenate([head, raw[30:o], extra.ravel(), raw[o + dg.size:]]).tofile(tmp_path / "mesh3.bin")
    m = read_mesh_bin(tmp_path / "mesh3.bin")
    back = mesh_from_bin(m, 2, msh.elemtype, msh.porder)
    assert np.array_equal(back.dgnodes, msh.dgnodes)
    m["ndims"]["ncd"] = 1
    with pytest.raises(ValueError, match="ncd"):
        mesh_from_bin(m, 2, msh.elemtype, msh.porder)


def test_sol_bin_pdg_count_without_payload(tmp_path, prob):
    """The preprocessor writes numel(PDG) into the header but the PDG values only
    for the wave formulation (digasopre.m:129, :292-297): such a file is read
    with an empty PDG."""
    from hfem_amd.binio import read_sol_bin, write_sol_bin
    path = tmp_path / "sol.bin"
    write_sol_bin(path, prob.UDG, prob.UH)
    raw = np.fromfile(path, dtype="<f8")
    raw[2] = 17
    raw.tofile(path)
    s = read_sol_bin(path)
    assert s["PDG"].size == 0 and np.array_equal(s["UH"], prob.UH.ravel())
