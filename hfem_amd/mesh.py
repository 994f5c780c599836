"""Structured test meshes and the connectivity arrays the element path reads.

Generates, for jittered unit squares/cubes, the arrays of ``meshstruct``
(preprocessing/digasopre.h) consumed by the element path:

* ``dgnodes[ie, n, x]`` (reference layout ``dgnodes[ie*npv*ncd + x*npv + n]``,
  Appendix A of the survey; read at assembly/elementassembly.cpp:39),
* ``t2f[ie, a]`` element-to-face, ``bf[ie, a]`` (< 0 on boundary faces; boundary
  id ``-bf-1``, assembly/elementassembly.cpp:1078,1102-1103),
* ``elcon[ie, a, m]`` global trace node of local face ``a`` node ``m``
  (assembly/elementassembly.cpp:605-609), numbered ``face*npf + k`` with the
  canonical node order of the face's first element, like
  preprocessing/mkelconmatlab.m:101-135,
* the block-sparse trace graph ``ent2entStart/ent2ent`` with the self block
  first and neighbours ascending (preprocessing/elcon2entcon.m:39-59).

Element patterns follow mesh/squaremesh.m:29-50 (SW-NE diagonals, Cartesian
element order) and mesh/cubemesh.m:30-46 (6 tets per cube with the ``t0`` split;
hexes).  Interior vertices are jittered uniformly by +-``jitter*h`` with a seeded
numpy PCG64 generator; periodic meshes jitter every lattice vertex.
"""
from dataclasses import dataclass, field

import numpy as np

from .master import make_master

__all__ = ["Mesh", "square_tri_mesh", "cube_tet_mesh", "cube_hex_mesh", "element_faces"]

# local faces as vertex lists (master/mkmasternodes.m; preprocessing/mkt2fmatlab.m)
_FACES = {
    (2, 0): [[1, 2], [2, 0], [0, 1]],
    (3, 0): [[1, 2, 3], [0, 3, 2], [0, 1, 3], [0, 2, 1]],
    (3, 1): [[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [2, 3, 7, 6], [1, 2, 6, 5], [3, 0, 4, 7]],
}


def element_faces(nd, elemtype):
    return _FACES[(nd, elemtype)]


@dataclass
class Mesh:
    nd: int
    elemtype: int
    porder: int
    p: np.ndarray            # (nv, nd) vertex coordinates
    t: np.ndarray            # (ne, nve) vertices, 0-based
    dgnodes: np.ndarray      # (ne, npv, nd)
    t2f: np.ndarray          # (ne, nfe)
    bf: np.ndarray           # (ne, nfe)
    elcon: np.ndarray        # (ne, nfe, npf)
    f2e: np.ndarray          # (nf, 4): e1, a1, e2, a2 (e2=-1 on boundary)
    nf: int
    ent2entStart: np.ndarray = field(default=None)
    ent2ent: np.ndarray = field(default=None)
    periodic: bool = False

    @property
    def ne(self):
        return self.t.shape[0]

    @property
    def master(self):
        return make_master(self.nd, self.elemtype, self.porder)

    def involution_ok(self):
        """True when every element's face-node map is its own inverse.

        The reference scatter (assembly/globalassembly.cpp:54-60,77-85) indexes as
        if it were; the build's scatter uses the true inverse and agrees with the
        reference exactly when this holds.
        """
        npf = self.elcon.shape[2]
        ind = self.elcon - self.t2f[:, :, None] * npf
        back = np.take_along_axis(ind, ind, axis=2)
        return bool(np.all(back == np.arange(npf)[None, None, :]))


def _vertex_basis(nd, elemtype, xi):
    """Linear (simplex) / multilinear (hex) vertex basis at master points xi (n, nd)."""
    if elemtype == 0:
        return np.concatenate([1.0 - xi.sum(1, keepdims=True), xi], axis=1)
    x, y, z = xi[:, 0], xi[:, 1], xi[:, 2]
    cols = []
    for (a, b, c) in [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]:
        cols.append((x if a else 1 - x) * (y if b else 1 - y) * (z if c else 1 - z))
    return np.stack(cols, axis=1)


def _finish(nd, elemtype, porder, p, t, vpos, match_pos, periodic=False):
    """Faces, elcon, bf and trace graph from vertices ``t`` (ne, nve).

    ``vpos``: per element vertex coordinates used for dgnodes (ne, nve, nd);
    ``match_pos``: per element vertex coordinates used to match face nodes
    between the two elements of a face (unjittered, periodic-reduced).
    """
    m = make_master(nd, elemtype, porder)
    ne = t.shape[0]
    faces = element_faces(nd, elemtype)
    nfe = len(faces)
    phi = _vertex_basis(nd, elemtype, m.plocvl)                 # (npv, nve)
    dgnodes = np.einsum("nv,evx->enx", phi, vpos)
    mpos = np.einsum("nv,evx->enx", phi, match_pos)

    key2face = {}
    f2e = []
    t2f = np.zeros((ne, nfe), dtype=np.int64)
    for ie in range(ne):
        for a, fv in enumerate(faces):
            key = tuple(sorted(int(t[ie, v]) for v in fv))
            fid = key2face.get(key)
            if fid is None:
                fid = len(f2e)
                key2face[key] = fid
                f2e.append([ie, a, -1, -1])
            else:
                assert f2e[fid][2] < 0, "face shared by more than two elements"
                f2e[fid][2] = ie
                f2e[fid][3] = a
            t2f[ie, a] = fid
    f2e = np.array(f2e, dtype=np.int64)
    nf = len(f2e)
    npf = m.npf
    perm = m.perm                                                # (npf, nfe)
    elcon = np.zeros((ne, nfe, npf), dtype=np.int64)
    bf = np.zeros((ne, nfe), dtype=np.int64)
    for fid, (e1, a1, e2, a2) in enumerate(f2e):
        elcon[e1, a1, :] = fid * npf + np.arange(npf)
        if e2 < 0:
            bf[e1, a1] = -1                                      # boundary id 0
            continue
        bf[e1, a1] = e2
        bf[e2, a2] = e1
        x1 = mpos[e1, perm[:, a1], :]
        x2 = mpos[e2, perm[:, a2], :]
        if periodic:
            d = x2[:, None, :] - x1[None, :, :]
            d = d - np.round(d)
        else:
            d = x2[:, None, :] - x1[None, :, :]
        dist = np.sqrt((d ** 2).sum(-1))
        k = dist.argmin(1)
        assert np.all(dist[np.arange(npf), k] < 1e-8), "face node matching failed"
        assert len(set(k.tolist())) == npf
        elcon[e2, a2, :] = fid * npf + k

    # trace graph: rows = faces, self first then neighbours ascending
    nbr = [set() for _ in range(nf)]
    for ie in range(ne):
        fs = set(int(v) for v in t2f[ie])
        for fi in fs:
            nbr[fi] |= fs
    start = [0]
    cols = []
    for fi in range(nf):
        others = sorted(nbr[fi] - {fi})
        cols.extend([fi] + others)
        start.append(len(cols))
    return Mesh(nd=nd, elemtype=elemtype, porder=porder, p=p, t=t, dgnodes=dgnodes, t2f=t2f, bf=bf,
                elcon=elcon, f2e=f2e, nf=nf, ent2entStart=np.array(start, dtype=np.int64),
                ent2ent=np.array(cols, dtype=np.int64), periodic=periodic)


def _jitter(p, h, jitter, rng, interior):
    q = p.copy()
    delta = rng.uniform(-jitter * h, jitter * h, size=p.shape)
    q[interior] += delta[interior]
    return q


def square_tri_mesh(ncell, porder, seed=0, jitter=0.05):
    """Unit square, ncell x ncell cells, 2 triangles per cell (squaremesh parity 0)."""
    m = n = ncell + 1
    x, y = np.meshgrid(np.arange(m) / (m - 1), np.arange(n) / (n - 1), indexing="ij")
    p0 = np.stack([x.reshape(-1, order="F"), y.reshape(-1, order="F")], axis=1)
    tris = []
    for jy in range(n - 1):
        for ix in range(m - 1):
            v = ix + jy * m
            tris.append([v, v + m + 1, v + m])      # [1, m+2, m+1]
            tris.append([v, v + 1, v + m + 1])      # [1, 2, m+2]
    t = np.array(tris, dtype=np.int64)
    h = 1.0 / ncell
    interior = (p0[:, 0] > 1e-12) & (p0[:, 0] < 1 - 1e-12) & (p0[:, 1] > 1e-12) & (p0[:, 1] < 1 - 1e-12)
    rng = np.random.default_rng(seed)
    p = _jitter(p0, h, jitter, rng, interior)
    return _finish(2, 0, porder, p, t, p[t], p0[t])


_T0 = np.array([[5, 6, 7, 3], [5, 6, 2, 3], [5, 1, 2, 3], [6, 8, 7, 4], [6, 7, 4, 3], [6, 2, 4, 3]]) - 1


def cube_tet_mesh(ncell, porder, seed=0, jitter=0.05):
    """Unit cube, ncell^3 cubes x 6 tets (cubemesh t0 split), oriented positively."""
    m = n = o = ncell + 1
    g = np.meshgrid(*([np.arange(m) / (m - 1)] * 3), indexing="ij")
    p0 = np.stack([c.reshape(-1, order="F") for c in g], axis=1)
    cmap = np.array([0, 1, m, m + 1, m * n, m * n + 1, m * n + m, m * n + m + 1])
    base = cmap[_T0]                                          # (6, 4)
    tets = []
    for k in range(6):
        for ix in range(m - 1):
            for iy in range(n - 1):
                for iz in range(o - 1):
                    tets.append(base[k] + ix + iy * m + iz * m * n)
    t = np.array(tets, dtype=np.int64)
    h = 1.0 / ncell
    interior = np.all((p0 > 1e-12) & (p0 < 1 - 1e-12), axis=1)
    rng = np.random.default_rng(seed)
    p = _jitter(p0, h, jitter, rng, interior)
    # fix orientation as fixmesh does (flip the first two vertices of negative-volume tets)
    a = p[t]
    vol = np.einsum("ij,ij->i", np.cross(a[:, 1] - a[:, 0], a[:, 2] - a[:, 0]), a[:, 3] - a[:, 0])
    flip = vol < 0
    t[flip, 0], t[flip, 1] = t[flip, 1].copy(), t[flip, 0].copy()
    return _finish(3, 0, porder, p, t, p[t], p0[t])


def cube_hex_mesh(ncell, porder, seed=0, jitter=0.05, periodic=True):
    """Unit cube of ncell^3 hexes (cubemesh elemtype 1 vertex order), x fastest."""
    N = ncell
    h = 1.0 / N
    rng = np.random.default_rng(seed)
    if periodic:
        lat = np.meshgrid(*([np.arange(N) / N] * 3), indexing="ij")
        p0 = np.stack([c.reshape(-1, order="F") for c in lat], axis=1)
        p = p0 + rng.uniform(-jitter * h, jitter * h, size=p0.shape)

        def vid(i, j, k):
            return (i % N) + (j % N) * N + (k % N) * N * N
    else:
        m = N + 1
        lat = np.meshgrid(*([np.arange(m) / N] * 3), indexing="ij")
        p0 = np.stack([c.reshape(-1, order="F") for c in lat], axis=1)
        interior = np.all((p0 > 1e-12) & (p0 < 1 - 1e-12), axis=1)
        p = _jitter(p0, h, jitter, rng, interior)

        def vid(i, j, k):
            return i + j * m + k * m * m
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    t = []
    vpos = []
    mpos = []
    for kz in range(N):
        for jy in range(N):
            for ix in range(N):
                ids = [vid(ix + a, jy + b, kz + c) for (a, b, c) in corners]
                t.append(ids)
                shift = np.array([[(ix + a) // N if periodic else 0, (jy + b) // N if periodic else 0,
                                   (kz + c) // N if periodic else 0] for (a, b, c) in corners], dtype=float)
                vpos.append(p[ids] + shift)
                mpos.append(np.array([[(ix + a) * h, (jy + b) * h, (kz + c) * h] for (a, b, c) in corners]))
    t = np.array(t, dtype=np.int64)
    return _finish(3, 1, porder, p, t, np.array(vpos), np.array(mpos), periodic=periodic)
