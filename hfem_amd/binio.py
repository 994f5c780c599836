"""The reference's binary files, read (and the solution files written) with numpy.

* ``sol.bin`` -- the initial state the MATLAB preprocessor writes for a serial
  run (preprocessing/digasopre.m:126-129 and :283-297) and the C++ reads
  (readsolstruct, preprocessing/readbinaryfiles.cpp:624-684): ten doubles of
  dimensions ``[numel(UDG), numel(UH), numel(PDG), 0, ...]``, then UDG, UH and
  PDG, every value a double in the byte order of the writer (native unless the
  preprocessor was told otherwise).
* the solution output of ``writeSol2File`` (preprocessing/ioutilities.cpp:256-300):
  two doubles ``[len, n]``, then per element the first ``npv*ncu`` values of its
  UDG block (the u part; the whole ``npv*nc`` block of every element when q is
  written, writeQflag 1), then the ``n`` values of UH.
* a serial ``mesh.bin``, ``master.bin`` and ``app.bin`` (read only), in the
  order readmeshstruct, readmasterstruct and readappstruct read them;
* one rank's record of a parallel case (``mesh<k>.bin`` / ``sol<k>.bin`` holding
  several ranks each, sliced by ``app.meshsize`` / ``app.solsize``), as the
  parallel branches of readmeshstruct and readsolstruct read it.

Array layouts are the reference's, which the library uses as they are: UDG
element-major ``(ne, nc, npv)`` (MATLAB ``npv x nc x ne``), UH ``(nnodes, nch)``
(MATLAB ``nch x nnodes``).
"""
import numpy as np

SOL_HEADER = 10


def _dt(endian):
    if endian not in ("<", ">"):
        raise ValueError("endian must be '<' (little) or '>' (big)")
    return np.dtype(endian + "f8")


def write_sol_bin(path, UDG, UH, PDG=None, endian="<"):
    """Write ``sol.bin``: the header [numel(UDG), numel(UH), numel(PDG), 0 x 7],
    then UDG, UH and (when given) PDG, flattened in the reference layout."""
    parts = [np.asarray(a, dtype=np.float64).ravel() for a in (UDG, UH)]
    parts.append(np.zeros(0) if PDG is None else np.asarray(PDG, dtype=np.float64).ravel())
    head = np.zeros(SOL_HEADER)
    head[:3] = [p.size for p in parts]
    np.concatenate([head] + parts).astype(_dt(endian)).tofile(path)


def read_sol_bin(path, endian="<"):
    """Read ``sol.bin``; returns dict(ndims, UDG, UH, PDG) with flat arrays.
    The dimensions must account for the whole file, as readsolstruct assumes."""
    raw = np.fromfile(path, dtype=_dt(endian)).astype(np.float64)
    if raw.size < SOL_HEADER:
        raise ValueError(f"{path}: {raw.size} values, shorter than the {SOL_HEADER}-value header")
    nd = raw[:SOL_HEADER]
    if not np.all(nd == np.round(nd)) or np.any(nd < 0):
        raise ValueError(f"{path}: header {nd[:3]} is not a list of counts (wrong byte order?)")
    nudg, nuh, npdg = (int(v) for v in nd[:3])
    if SOL_HEADER + nudg + nuh == raw.size and npdg > 0:
        # the preprocessor writes numel(PDG) into the header but the PDG values only
        # for the wave formulation (digasopre.m:129, :292-297): no payload, no PDG
        npdg = 0
    if SOL_HEADER + nudg + nuh + npdg != raw.size:
        raise ValueError(f"{path}: header counts {nudg}+{nuh}+{npdg} do not match {raw.size - SOL_HEADER} values")
    o = SOL_HEADER
    UDG = raw[o:o + nudg]
    UH = raw[o + nudg:o + nudg + nuh]
    PDG = raw[o + nudg + nuh:]
    return dict(ndims=nd.astype(np.int64), UDG=UDG, UH=UH, PDG=PDG)


def write_solution_file(path, UDG, UH, ncu, write_q=False, endian="<"):
    """Write the ``writeSol2File`` output for UDG (ne, nc, npv) and UH."""
    UDG = np.asarray(UDG, dtype=np.float64)
    ne, nc, npv = UDG.shape
    uh = np.asarray(UH, dtype=np.float64).ravel()
    body = UDG if write_q else UDG[:, :ncu, :]
    head = np.array([float(body.size), float(uh.size)])
    np.concatenate([head, body.ravel(), uh]).astype(_dt(endian)).tofile(path)


def read_solution_file(path, ne, npv, nc, ncu, endian="<"):
    """Read a ``writeSol2File`` output; returns (U, UH, with_q): U is
    (ne, nc, npv) when q was written, else (ne, ncu, npv)."""
    raw = np.fromfile(path, dtype=_dt(endian)).astype(np.float64)
    if raw.size < 2:
        raise ValueError(f"{path}: no header")
    ln, n = int(raw[0]), int(raw[1])
    if raw.size != 2 + ln + n:
        raise ValueError(f"{path}: header [{ln}, {n}] does not match {raw.size - 2} values")
    if ln == ne * npv * nc:
        ncomp, with_q = nc, True
    elif ln == ne * npv * ncu:
        ncomp, with_q = ncu, False
    else:
        raise ValueError(f"{path}: {ln} values are neither ne*npv*nc nor ne*npv*ncu")
    U = raw[2:2 + ln].reshape(ne, ncomp, npv)
    return U, raw[2 + ln:], with_q


def load_sol_bin(ctx, path, endian="<"):
    """Set a context's state from a reference ``sol.bin``: UDG and UH as read,
    SH zero (the reference starts a steady solve with zero sources)."""
    prob = ctx.prob
    d = prob.dims
    s = read_sol_bin(path, endian)
    ne, nc, npv, nch = prob.ne, d["nc"], d["npv"], d["nch"]
    if s["UDG"].size != ne * nc * npv:
        raise ValueError(f"{path}: UDG has {s['UDG'].size} values, the mesh needs {ne * nc * npv}")
    nuh = prob.UH.size
    if s["UH"].size != nuh:
        raise ValueError(f"{path}: UH has {s['UH'].size} values, the mesh needs {nuh}")
    UDG = s["UDG"].reshape(ne, nc, npv)
    UH = s["UH"].reshape(prob.UH.shape)
    ctx.set_solution(UDG, UH, np.zeros_like(UDG))
    return UDG, UH


MESH_HEADER = 30


def read_mesh_bin(path, endian="<"):
    """Read a serial ``mesh.bin`` in the order readmeshstruct reads it
    (preprocessing/readbinaryfiles.cpp, serial branch): 30 dimensions, then
    elementtype (ne), t (nve x ne), dgnodes (npe x ncd x ne), elcon
    (npf*nfe x ne), bf, t2f, t2t (nfe x ne each), f ((nvf+3) x nf) and isEDGface
    (nf), all doubles, indices 0-based.  Returns the dimensions by name and the
    arrays in their MATLAB shapes (column-major: the last index is the element
    or face); the file must hold exactly these, as the preprocessor's size
    formula (digasopre.m:278) gives it."""
    raw = np.fromfile(path, dtype=_dt(endian)).astype(np.float64)
    if raw.size < MESH_HEADER:
        raise ValueError(f"{path}: shorter than the {MESH_HEADER}-value header")
    nd = raw[:MESH_HEADER]
    if not np.all(nd == np.round(nd)):
        raise ValueError(f"{path}: header is not a list of integers (wrong byte order?)")
    nd = nd.astype(np.int64)
    names = ["dim", "ncd", "nfe", "nve", "nvf", "ne", "nf", "nv", "ndh", "npe", "npf", "nc", "ncu", "ncq", "ncp",
             "nch"]
    d = {k: int(nd[i]) for i, k in enumerate(names)}
    d["hybrid"], d["nproc"] = int(nd[-2]), int(nd[-1])
    if d["nproc"] != 1:
        raise ValueError(f"{path}: a file of {d['nproc']} ranks; only serial meshes are read")
    ne, nf, nfe, nve, nvf, npe, npf, ncd = (d[k] for k in ("ne", "nf", "nfe", "nve", "nvf", "npe", "npf", "ncd"))
    shapes = [("elementtype", (ne,)), ("t", (ne, nve)), ("dgnodes", (ne, ncd, npe)), ("elcon", (ne, npf * nfe)),
              ("bf", (ne, nfe)), ("t2f", (ne, nfe)), ("t2t", (ne, nfe)), ("f", (nf, nvf + 3)), ("isEDGface", (nf,))]
    need = MESH_HEADER + sum(int(np.prod(s)) for _, s in shapes)
    if need != raw.size:
        raise ValueError(f"{path}: the dimensions need {need} values, the file holds {raw.size}")
    out = {"ndims": d}
    o = MESH_HEADER
    for name, shp in shapes:
        n = int(np.prod(shp))
        a = raw[o:o + n].reshape(shp)       # the element (face) index slowest, as MATLAB writes
        o += n
        out[name] = a if name == "dgnodes" else a.astype(np.int64)
    return out


def rank_record(app, rank, kind):
    """Where rank ``rank``'s record of a parallel case lies (the parallel branches
    of readmeshstruct / readsolstruct, preprocessing/readbinaryfiles.cpp): the
    file number k of ``mesh<k>.bin`` / ``sol<k>.bin``, the record's offset and
    its length in doubles, from app.nproc, app.nfile and app.meshsize (kind
    'mesh') or app.solsize (kind 'sol'), which hold nproc/nfile + 1 offsets per
    file."""
    nproc, nfile = int(app["nproc"]), int(app["nfile"])
    if nproc < 2:
        raise ValueError("a serial case (app.nproc = 1) has no rank records")
    if nfile < 1 or nproc % nfile:
        raise ValueError(f"nproc {nproc} is not a multiple of nfile {nfile}")
    if not 0 <= rank < nproc:
        raise ValueError(f"rank {rank} outside [0, {nproc})")
    per = nproc // nfile
    filenumber, file_rank = rank // per + 1, rank % per
    sizes = np.asarray(app["meshsize" if kind == "mesh" else "solsize"], dtype=np.int64)
    pos = (per + 1) * (filenumber - 1)
    if sizes.size < pos + per + 1:
        raise ValueError(f"app.{kind}size has {sizes.size} offsets, {nfile} files of {per} ranks need "
                         f"{nfile * (per + 1)}")
    start, end = int(sizes[pos + file_rank]), int(sizes[pos + file_rank + 1])
    return filenumber, start, end - start


def _record(path, start, length, endian):
    raw = np.fromfile(path, dtype=_dt(endian)).astype(np.float64)
    if start < 0 or length < 0 or start + length > raw.size:
        raise ValueError(f"{path}: record [{start}, {start + length}) outside the file's {raw.size} values")
    return raw[start:start + length]


def read_mesh_bin_rank(prefix, app, rank, endian="<"):
    """Rank ``rank``'s record of a parallel ``mesh<k>.bin`` (``prefix`` +
    'mesh<k>.bin'), in the order the parallel branch of readmeshstruct reads it:
    30 dimensions (nnbsd = ndims[20]), then elementtype, t, dgnodes, elcon, bf,
    t2f, t2t, f, isEDGface as in a serial file, then nbsd (nnbsd), extintelem,
    elem2cpu (ne each), extintface, face2cpu (nf each) and, for hybrid != 0,
    extintent and ent2cpu (ndh each).  The record's length must be what the
    dimensions give, as readmeshstruct checks."""
    k, start, length = rank_record(app, rank, "mesh")
    rec = _record(f"{prefix}mesh{k}.bin", start, length, endian)
    if rec.size < MESH_HEADER:
        raise ValueError(f"rank {rank}: record shorter than the {MESH_HEADER}-value header")
    nd = rec[:MESH_HEADER]
    if not np.all(nd == np.round(nd)):
        raise ValueError(f"rank {rank}: header is not a list of integers (wrong byte order?)")
    nd = nd.astype(np.int64)
    names = ["dim", "ncd", "nfe", "nve", "nvf", "ne", "nf", "nv", "ndh", "npe", "npf", "nc", "ncu", "ncq", "ncp",
             "nch", "nn", "nie", "nif", "nin", "nnbsd", "nxe", "nxf", "nxn"]
    d = {key: int(nd[i]) for i, key in enumerate(names)}
    d["hybrid"], d["nproc"] = int(nd[-2]), int(nd[-1])
    if d["nproc"] != int(app["nproc"]):
        raise ValueError(f"rank {rank}: the record is for {d['nproc']} ranks, app.bin for {app['nproc']}")
    ne, nf, nfe, nve, nvf, npe, npf, ncd, ndh = (d[x] for x in ("ne", "nf", "nfe", "nve", "nvf", "npe", "npf", "ncd",
                                                                 "ndh"))
    shapes = [("elementtype", (ne,)), ("t", (ne, nve)), ("dgnodes", (ne, ncd, npe)), ("elcon", (ne, npf * nfe)),
              ("bf", (ne, nfe)), ("t2f", (ne, nfe)), ("t2t", (ne, nfe)), ("f", (nf, nvf + 3)), ("isEDGface", (nf,)),
              ("nbsd", (d["nnbsd"],)), ("extintelem", (ne,)), ("elem2cpu", (ne,)), ("extintface", (nf,)),
              ("face2cpu", (nf,))]
    if d["hybrid"] != 0:
        shapes += [("extintent", (ndh,)), ("ent2cpu", (ndh,))]
    need = MESH_HEADER + sum(int(np.prod(sh)) for _, sh in shapes)
    if need != rec.size:
        raise ValueError(f"rank {rank}: the dimensions need {need} values, the record holds {rec.size}")
    out = {"ndims": d}
    o = MESH_HEADER
    for name, shp in shapes:
        n = int(np.prod(shp))
        a = rec[o:o + n].reshape(shp)
        o += n
        out[name] = a if name == "dgnodes" else a.astype(np.int64)
    return out


def read_sol_bin_rank(prefix, app, rank, endian="<"):
    """Rank ``rank``'s record of a parallel ``sol<k>.bin``, as the parallel branch
    of readsolstruct reads it: the ten-value header, then UDG, UH and PDG; the
    counts must account for the record."""
    k, start, length = rank_record(app, rank, "sol")
    rec = _record(f"{prefix}sol{k}.bin", start, length, endian)
    if rec.size < SOL_HEADER:
        raise ValueError(f"rank {rank}: record shorter than the {SOL_HEADER}-value header")
    nd = rec[:SOL_HEADER]
    nudg, nuh, npdg = (int(v) for v in nd[:3])
    if SOL_HEADER + nudg + nuh + npdg != rec.size:
        raise ValueError(f"rank {rank}: header counts {nudg}+{nuh}+{npdg} do not match {rec.size - SOL_HEADER}")
    o = SOL_HEADER
    return dict(ndims=nd.astype(np.int64), UDG=rec[o:o + nudg], UH=rec[o + nudg:o + nudg + nuh],
                PDG=rec[o + nudg + nuh:])


def mesh_from_bin(m, nd, elemtype, porder):
    """The library's Mesh from the arrays of ``read_mesh_bin`` (HDG, one element
    type): dgnodes, elcon, bf and t2f as the reference uses them, the face ->
    element map and the trace graph (faces sharing an element) derived from t2f."""
    from .mesh import Mesh
    d = m["ndims"]
    ne, nfe, npf, nf = d["ne"], d["nfe"], d["npf"], d["nf"]
    if d["hybrid"] != 0:
        raise ValueError("only HDG meshes (hybrid 0) are read")
    if d["ncd"] < nd:
        raise ValueError(f"mesh.bin has ncd={d['ncd']} coordinate components, the geometry needs nd={nd}")
    t2f = m["t2f"]
    f2e = -np.ones((nf, 4), dtype=np.int64)
    for ie in range(ne):
        for a in range(nfe):
            fid = t2f[ie, a]
            if f2e[fid, 0] < 0:
                f2e[fid, 0], f2e[fid, 1] = ie, a
            else:
                f2e[fid, 2], f2e[fid, 3] = ie, a
    nbr = [set() for _ in range(nf)]
    for ie in range(ne):
        fs = set(int(v) for v in t2f[ie])
        for fi in fs:
            nbr[fi] |= fs
    start, cols = [0], []
    for fi in range(nf):
        cols.extend([fi] + sorted(nbr[fi] - {fi}))
        start.append(len(cols))
    return Mesh(nd=nd, elemtype=elemtype, porder=porder, p=np.zeros((d["nv"], nd)), t=m["t"],
                dgnodes=np.transpose(m["dgnodes"][:, :nd, :], (0, 2, 1)).copy(), t2f=t2f, bf=m["bf"],
                elcon=m["elcon"].reshape(ne, nfe, npf), f2e=f2e, nf=nf,
                ent2entStart=np.array(start, dtype=np.int64), ent2ent=np.array(cols, dtype=np.int64))


MASTER_HEADER = 20


def read_master_bin(path, nelem=1, endian="<"):
    """Read ``master.bin`` in the order readmasterstruct reads it
    (preprocessing/readbinaryfiles.cpp): per element type 20 dimensions (dim,
    elemtype, nodetype, npe, nme, nge, ngeR, nfe, nve, nle, nplmax, ...), the
    per-face counts npf, nmf, ngf, ngfR, nvf (nfe each) and npl (nle), the
    volume tables (plocvl, gpvl, gwvl, shapvt, shapvg, shapmv, shapvgdotshapvl
    and their R variants, permnode, permedge), then per face plocfc, gpfc, gwfc,
    shapft, shapfg, shapmf, shapfgdotshapfc, the R variants, perm, permgeom and
    face.  Arrays come back flat, column-major as written; index arrays as ints.
    Returns a list with one dict per element type."""
    raw = np.fromfile(path, dtype=_dt(endian)).astype(np.float64)
    o = 0
    out = []

    def take(n, integer=False):
        nonlocal o
        if o + n > raw.size:
            raise ValueError(f"{path}: ends inside the master data")
        a = raw[o:o + n]
        o += n
        return a.astype(np.int64) if integer else a.copy()

    for _ in range(nelem):
        nd = take(MASTER_HEADER)
        if not np.all(nd == np.round(nd)):
            raise ValueError(f"{path}: header is not a list of integers (wrong byte order?)")
        dim, elemtype, nodetype, npe, nme, nge, ngeR, nfe, nve, nle, nplmax = (int(v) for v in nd[:11])
        m = dict(dim=dim, elemtype=elemtype, nodetype=nodetype, npe=npe, nme=nme, nge=nge, ngeR=ngeR, nfe=nfe,
                 nve=nve, nle=nle, nplmax=nplmax)
        for k in ("npf", "nmf", "ngf", "ngfR", "nvf"):
            m[k] = take(nfe, True)
        m["npl"] = take(nle, True)
        d1 = dim + 1
        for k, n in (("plocvl", npe * dim), ("gpvl", nge * dim), ("gwvl", nge), ("shapvt", npe * nge * d1),
                     ("shapvg", npe * nge * d1), ("shapmv", nme * nge * d1), ("shapvgdotshapvl", npe * npe * nge * d1),
                     ("gpvlR", ngeR * dim), ("gwvlR", ngeR), ("shapvtR", npe * ngeR * d1), ("shapvgR", npe * ngeR * d1),
                     ("shapmvR", nme * ngeR * d1)):
            m[k] = take(n)
        m["permnode"] = take(nve, True)
        m["permedge"] = take(nle * nplmax, True)
        faces = []
        for j in range(nfe):
            npf, nmf, ngf, ngfR, nvf = (int(m[k][j]) for k in ("npf", "nmf", "ngf", "ngfR", "nvf"))
            f = {}
            for k, n in (("plocfc", npf * (dim - 1)), ("gpfc", ngf * (dim - 1)), ("gwfc", ngf),
                         ("shapft", npf * ngf * dim), ("shapfg", npf * ngf * dim), ("shapmf", nmf * ngf * dim),
                         ("shapfgdotshapfc", npf * npf * ngf * dim), ("gpfcR", ngfR * (dim - 1)), ("gwfcR", ngfR),
                         ("shapftR", npf * ngfR * dim), ("shapfgR", npf * ngfR * dim), ("shapmfR", nmf * ngfR * dim)):
                f[k] = take(n)
            f["perm"] = take(npf, True)
            f["permgeom"] = take(nmf, True)
            f["face"] = take(nvf, True)
            faces.append(f)
        m["faces"] = faces
        out.append(m)
    if o != raw.size:
        raise ValueError(f"{path}: {raw.size - o} values left after {nelem} element type(s)")
    return out


def master_tables(m):
    """The tables hfem_set_master takes, from one element type of
    ``read_master_bin``: shapvt, shapvg, face 0's shapft and shapfg (the C++
    element path uses face 0's tables for every face,
    assembly/elementassembly.cpp:161,593) and perm (npf x nfe, column-major)."""
    return dict(shapvt=m["shapvt"], shapvg=m["shapvg"], shapft=m["faces"][0]["shapft"],
                shapfg=m["faces"][0]["shapfg"], perm=np.concatenate([f["perm"] for f in m["faces"]]))


APP_HEADER = 40
_APP_ARRAYS = [("meshsize", True), ("solsize", True), ("porder", True), ("elemtype", True), ("nodetype", True),
               ("pgauss", True), ("pgaussR", True), ("quadtype", True), ("bcm", True), ("bcs", False), ("bcd", True),
               ("bcv", False), ("dt", False), ("flag", True), ("factor", False), ("problem", True),
               ("physicsparam", False), ("solversparam", False)]


def read_app_bin(path, endian="<"):
    """Read ``app.bin`` in the order readappstruct reads it
    (preprocessing/readbinaryfiles.cpp): 40 dimensions (the lengths of the 18
    arrays, then nproc, nfile, nd, ncd, nc, ncu, nch, ncq, ncp, nco), then
    meshsize, solsize, porder, elemtype, nodetype, pgauss, pgaussR, quadtype,
    bcm, bcs, bcd, bcv, dt, flag, factor, problem, physicsparam, solversparam."""
    raw = np.fromfile(path, dtype=_dt(endian)).astype(np.float64)
    if raw.size < APP_HEADER:
        raise ValueError(f"{path}: shorter than the {APP_HEADER}-value header")
    nd = raw[:APP_HEADER]
    if not np.all(nd == np.round(nd)) or np.any(nd < 0):
        raise ValueError(f"{path}: header is not a list of counts (wrong byte order?)")
    nd = nd.astype(np.int64)
    need = APP_HEADER + int(nd[:18].sum())
    if need != raw.size:
        raise ValueError(f"{path}: the dimensions need {need} values, the file holds {raw.size}")
    app = {"ndims": nd}
    o = APP_HEADER
    for i, (name, integer) in enumerate(_APP_ARRAYS):
        a = raw[o:o + nd[i]]
        o += int(nd[i])
        app[name] = a.astype(np.int64) if integer else a.copy()
    for i, k in enumerate(("nproc", "nfile", "nd", "ncd", "nc", "ncu", "nch", "ncq", "ncp", "nco")):
        app[k] = int(nd[18 + i])
    return app


def app_settings(app):
    """The reference's meaning of the app arrays (initializestructs.cpp:91-143,
    :185-201) as the library takes them: the physics (physicsparam, fc_u =
    factor[0], fc_q = factor[1], tdep = flag[0]), DIRK stages and order
    (problem[4], problem[3]), the time steps, and the solver parameters
    (hfem_solver_params fields): Newton tolerance solversparam[0], GMRES
    tolerance solversparam[1], Newton iterations problem[13], GMRES(200) with at
    most 1000 iterations as the C++ fixes them, preconditioner flag[16] (0 RAS and
    1 subdomain block Jacobi are BILU0 on one rank, 2 entity-wise block Jacobi) or
    none when flag[17] is -1, orthogonalisation flag[19], ordering flag[20] (MDF
    when nonzero), ordering reuse flag[10], quasi-Newton flag[15] with storage
    accuracy flag[18], preconditioner precision flag[24] (precPrecision, :201)."""
    flag, fac, prob, sp = app["flag"], app["factor"], app["problem"], app["solversparam"]

    def at(a, i, default):
        return a[i] if len(a) > i else default
    prec = int(at(flag, 16, 0))
    params = dict(newton_tol=float(at(sp, 0, 1e-7)), gmres_tol=float(at(sp, 1, 1e-6)),
                  newton_maxiter=int(at(prob, 13, 10)), true_newton_maxiter=int(at(prob, 13, 10)),
                  restart=200, gmres_maxit=1000,
                  precond=-1 if int(at(flag, 17, 0)) == -1 else (2 if prec == 2 else 0),
                  orthog=int(at(flag, 19, 1)), ordering=1 if int(at(flag, 20, 1)) != 0 else 0,
                  reuse_ordering=int(at(flag, 10, 1)), quasi_newton=int(at(flag, 15, 0)),
                  qn_accuracy=int(at(flag, 18, 1)), prec_precision=1 if int(at(flag, 24, 1)) != 0 else 0)
    return dict(physicsparam=app["physicsparam"], fc_u=float(at(fac, 0, 0.0)), fc_q=float(at(fac, 1, 1.0)),
                tdep=int(at(flag, 0, 0)), dirk_order=int(at(prob, 3, 0)), dirk_stages=int(at(prob, 4, 0)),
                dt=app["dt"], solver_params=params)


# app.bin settings the library implements, as (array, index, required value,
# meaning): anything else selects physics or a scheme the element path does not
# have (the reference dispatches on them: boundary type bn = bcm[ib],
# assembly/elementassembly.cpp:1102-1104, application/FM/fbou.cpp types 1-11;
# stabilisation, application/FM/stabilizationTensor.cpp:3494-3528; the
# appname, viscosity, SGS, ALE and AV switches, preprocessing/initializestructs.cpp:91-143)
SUPPORTED_APP = [("problem", 0, 0, "hybrid (0 = HDG)"), ("problem", 1, 1, "appname (1 = Navier-Stokes)"),
                 ("problem", 5, 0, "convStabMethod (0 = constant tau)"),
                 ("problem", 6, 0, "diffStabMethod (0 = constant tau)"),
                 ("problem", 7, 0, "rotatingFrame"), ("problem", 8, 0, "viscosityModel (0 = constant)"),
                 ("problem", 9, 0, "SGSmodel"), ("problem", 10, 0, "ALE"), ("problem", 11, 0, "AV"),
                 ("flag", 1, 0, "wave")]


def check_supported(app, has_boundary=True):
    """Raise ValueError unless the app settings are those the library implements:
    HDG, Navier-Stokes, constant tau, constant viscosity, no SGS/ALE/AV, no wave
    formulation, at least the six physics parameters {gamma, eps, Re, Pr, Minf,
    tau} and, on a mesh with boundary faces, a far-field first boundary (bcm[0]
    = 1) with a state of nch values in bcs."""
    for arr, i, want, what in SUPPORTED_APP:
        a = app[arr]
        have = int(a[i]) if len(a) > i else 0
        if have != want:
            raise ValueError(f"app.bin {arr}[{i}] = {have}: {what} must be {want} for this library")
    if len(app["physicsparam"]) < 6:
        raise ValueError(f"app.bin physicsparam has {len(app['physicsparam'])} values, the NS flux needs 6")
    if has_boundary:
        bcm = app["bcm"]
        if len(bcm) < 1 or int(bcm[0]) != 1:
            raise ValueError(f"app.bin bcm = {list(bcm)}: boundary 0 must be far field (type 1) for this library")
        if len(app["bcs"]) < app["nch"]:
            raise ValueError(f"app.bin bcs has {len(app['bcs'])} values, the far-field state needs nch={app['nch']}")


def problem_from_case(prefix, endian="<"):
    """A problem from a reference case: the files ``prefix + 'app.bin'``,
    ``'master.bin'``, ``'mesh.bin'`` and ``'sol.bin'`` the preprocessor writes
    for a serial run (digasopre.m).  Returns (problem, master tables, solver
    parameters): pass the tables as ``HDGContext(problem, master_tables=...)``
    and the parameters to ``solve_steady`` / ``dirk_step``.  The library's
    boundary flux is the far-field one with one state for every boundary face,
    taken from the first boundary's entries of app.bcs (``bcs[ib*nch + k]``,
    assembly/elementassembly.cpp:693); a mesh with more than one boundary, and
    app settings that select other physics or schemes (``check_supported``), are
    refused."""
    from .problem import Problem
    app = read_app_bin(prefix + "app.bin", endian)
    check_supported(app, has_boundary=False)
    nd, ncu, nch, nc = app["nd"], app["ncu"], app["nch"], app["nc"]
    if len(app["elemtype"]) != 1:
        raise ValueError("meshes of one element type only")
    elemtype, porder = int(app["elemtype"][0]), int(app["porder"][0])
    msh = mesh_from_bin(read_mesh_bin(prefix + "mesh.bin", endian), nd, elemtype, porder)
    ids = set(int(v) for v in msh.bf[msh.bf < 0].ravel())
    if ids - {-1}:
        raise ValueError(f"boundary ids {sorted(ids)}: the library has one far-field boundary")
    check_supported(app, has_boundary=bool(ids))
    masters = read_master_bin(prefix + "master.bin", 1, endian)
    tables = master_tables(masters[0])
    s = read_sol_bin(prefix + "sol.bin", endian)
    ne, npv = msh.ne, masters[0]["npe"]
    npf = msh.elcon.shape[2]
    UDG = s["UDG"].reshape(ne, nc, npv)
    UH = s["UH"].reshape(msh.nf * npf, nch)
    st = app_settings(app)
    # a mesh without boundary faces never reads the far-field state
    ui = np.asarray(app["bcs"][:nch]) if ids else np.zeros(nch)
    prob = Problem(name=prefix, mesh=msh, nd=nd, ncu=ncu, nch=nch, nc=nc, param=np.asarray(st["physicsparam"]),
                   ui=ui, UDG=UDG, UH=UH, SH=np.zeros_like(UDG), fc_u=st["fc_u"], fc_q=st["fc_q"], tdep=st["tdep"])
    return prob, tables, st["solver_params"]
