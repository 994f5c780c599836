"""The reference's binary solution files, read and written with numpy.

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

Array layouts are the reference's, which the library uses as they are: UDG
element-major ``(ne, nc, npv)`` (MATLAB ``npv x nc x ne``), UH ``(nnodes, nch)``
(MATLAB ``nch x nnodes``).  Parallel files (several ranks in one file, sliced
by ``app.solsize``) are not read.
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
