"""Master-element tables for the HDG element path.

Builds the same arrays the reference reads into ``masterstruct``
(preprocessing/readbinaryfiles.cpp:95-223, written by preprocessing/digasopre.m:569-631
from master/mkmasterelement.m:1-145):

* nodes and face permutations follow master/mkmasternodes.m (uniform nodes,
  nodetype 0): trinodes2d (:76-135), tetnodes3d (:137-223), hexnodes3d (:496-602),
  faceindex (:621-633);
* shape functions are the nodal (Lagrange) basis of those nodes.  The reference
  obtains them as ``nf/A`` with a Koornwinder Vandermonde (master/mkshape.m:1-55);
  here the same polynomials are built exactly in rational arithmetic from the
  monomial basis and rounded once to float64, which differs from the
  reference's tables only by the reference's own rounding;
* derived products follow master/mkmasterelement.m:69-80 (volume) and
  :103-121 (faces).

Arrays are kept in the reference's column-major shapes (numpy arrays indexed
like the MATLAB arrays, flattened with ``order='F'`` when handed to C).
"""
from fractions import Fraction
from functools import lru_cache
import itertools

import numpy as np

from .quadrature import quad_rule

__all__ = ["Master", "make_master", "element_dims"]


# --------------------------------------------------------------------------
# nodes (mkmasternodes.m)
# --------------------------------------------------------------------------
def _nodes1d(p):
    return [Fraction(i, p) for i in range(p + 1)]


def _tri_nodes(p):
    """trinodes2d, nodetype 0: ndgrid order (x fastest), keep x+y<=1."""
    pts = []
    for j in range(p + 1):
        for i in range(p + 1):
            if i + j <= p:
                pts.append((Fraction(i, p), Fraction(j, p)))
    return pts


def _tri_perm(p):
    """permedge of trinodes2d (1-based there, 0-based here): faces [2 3],[3 1],[1 2]."""
    f1 = [p + 1]
    f2 = [(p + 1) * (p + 2) // 2]
    f3 = [1]
    for ii in range(2, p + 2):
        f1.append(f1[-1] + p + 2 - ii)
        f2.append(f2[-1] - ii)
        f3.append(ii)
    return np.array([f1, f2, f3], dtype=np.int64).T - 1  # (npf, nfe)


def _faceindex(p):
    n = p + 1
    a = np.zeros((n, n), dtype=np.int64)
    k = 1
    for j in range(n):
        for i in range(n - j):
            a[i, j] = k
            k += 1
    # b = a'; fi = b(b>0) in column-major order of b == row-major order of a
    return np.array([a[i, j] for i in range(n) for j in range(n) if a[i, j] > 0], dtype=np.int64) - 1


def _tet_nodes(p):
    """tetnodes3d, nodetype 0: layers in z, each a scaled trinodes2d."""
    xi = _nodes1d(p)
    pts = [(x, y, Fraction(0)) for (x, y) in _tri_nodes(p)]
    for ii in range(p - 1, 0, -1):
        for (x, y) in _tri_nodes(ii):
            s = xi[ii]
            pts.append((s * x, s * y, xi[p - ii]))
    pts.append((Fraction(0), Fraction(0), Fraction(1)))
    return pts


def _tet_perm(p, pts):
    """permface of tetnodes3d: faces [2 3 4],[1 4 3],[1 2 4],[1 3 2]."""
    f1 = [n for n, q in enumerate(pts) if q[0] + q[1] + q[2] == 1]
    x0 = [n for n, q in enumerate(pts) if q[0] == 0]
    y0 = [n for n, q in enumerate(pts) if q[1] == 0]
    z0 = [n for n, q in enumerate(pts) if q[2] == 0]
    fi = _faceindex(p)
    f2 = [x0[i] for i in fi]
    f4 = [z0[i] for i in fi]
    return np.array([f1, f2, y0, f4], dtype=np.int64).T


def _hex_nodes(p):
    xi = _nodes1d(p)
    return [(xi[i], xi[j], xi[k]) for k in range(p + 1) for j in range(p + 1) for i in range(p + 1)]


def _krongrid(a, b):
    # c(i,j) = a(i)+b(j); c(:) column-major -> a index fastest
    return [ai + bj for bj in b for ai in a]


def _hex_perm(p):
    m = n = k = p + 1
    f1 = _krongrid(range(1, m * (n - 1) + 2, m), range(0, m))
    f2 = _krongrid(range(m * n * (k - 1) + 1, m * n * (k - 1) + m + 1), range(0, m * (n - 1) + 1, m))
    f3 = _krongrid(range(1, m + 1), range(0, m * n * (k - 1) + 1, m * n))
    f4 = _krongrid(range(m * n, m * n - m, -1), range(0, m * n * (k - 1) + 1, m * n))
    f5 = _krongrid(range(m, m * n + 1, m), range(0, m * n * (k - 1) + 1, m * n))
    f6 = _krongrid(range(m * n - m + 1, 0, -m), range(0, m * n * (k - 1) + 1, m * n))
    return np.array([f1, f2, f3, f4, f5, f6], dtype=np.int64).T - 1


def _hex_face_nodes(p):
    """plocfc of hexnodes3d for faces 1..6 (all faces share the same 2D grid shape)."""
    xi = _nodes1d(p)
    # face 1: ndgrid(eta, xi) -> first coordinate fastest
    g = [(xi[i], xi[j]) for j in range(p + 1) for i in range(p + 1)]
    return g


# --------------------------------------------------------------------------
# exact Lagrange basis
# --------------------------------------------------------------------------
def _exps(p, dim, elemtype):
    if elemtype == 0:
        return [e for e in itertools.product(range(p + 1), repeat=dim) if sum(e) <= p]
    return list(itertools.product(range(p + 1), repeat=dim))


def _exact_inverse(A):
    n = len(A)
    M = [list(row) + [Fraction(int(i == j)) for j in range(n)] for i, row in enumerate(A)]
    for c in range(n):
        piv = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[piv] = M[piv], M[c]
        inv = 1 / M[c][c]
        M[c] = [v * inv for v in M[c]]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c]
                M[r] = [a - f * b for a, b in zip(M[r], M[c])]
    return [row[n:] for row in M]


def _mono(x, e):
    v = Fraction(1)
    for xi, ei in zip(x, e):
        if ei:
            v *= xi ** ei
    return v


def _dmono(x, e, d):
    if e[d] == 0:
        return Fraction(0)
    v = Fraction(e[d])
    for k, (xi, ei) in enumerate(zip(x, e)):
        ee = ei - 1 if k == d else ei
        if ee:
            v *= xi ** ee
    return v


def _shape(p, nodes, pts, elemtype):
    """mkshape: nodal basis and derivatives, returned as (np, npoints, dim+1)."""
    dim = len(nodes[0]) if nodes else 0
    if dim == 0:
        return np.ones((1, len(pts), 1))
    exps = _exps(p, dim, elemtype)
    assert len(exps) == len(nodes)
    V = [[_mono(x, e) for e in exps] for x in nodes]
    C = _exact_inverse(V)  # phi_n(x) = sum_k m_k(x) C[k][n]
    npn = len(nodes)
    out = np.zeros((npn, len(pts), dim + 1))
    for g, x in enumerate(pts):
        xf = [Fraction(float(v)) for v in x]
        vals = [[_mono(xf, e) for e in exps]] + [[_dmono(xf, e, d) for e in exps] for d in range(dim)]
        for d in range(dim + 1):
            row = vals[d]
            for n in range(npn):
                s = Fraction(0)
                for k in range(npn):
                    ck = C[k][n]
                    if ck and row[k]:
                        s += row[k] * ck
                out[n, g, d] = float(s)
    return out


# --------------------------------------------------------------------------
# master structure
# --------------------------------------------------------------------------
def element_dims(nd, elemtype, p):
    """(npv, npf, nfe) for the element family."""
    if nd == 2 and elemtype == 0:
        return (p + 1) * (p + 2) // 2, p + 1, 3
    if nd == 3 and elemtype == 0:
        return (p + 1) * (p + 2) * (p + 3) // 6, (p + 1) * (p + 2) // 2, 4
    if nd == 2 and elemtype == 1:
        return (p + 1) ** 2, p + 1, 4
    if nd == 3 and elemtype == 1:
        return (p + 1) ** 3, (p + 1) ** 2, 6
    raise ValueError((nd, elemtype))


class Master:
    """Subset of ``masterstruct`` (preprocessing/digasopre.h) used on the element path."""

    def __init__(self, nd, elemtype, porder, pgauss):
        self.nd, self.elemtype, self.porder, self.pgauss = nd, elemtype, porder, pgauss
        p = porder
        if nd == 2 and elemtype == 0:
            nodes = _tri_nodes(p)
            perm = _tri_perm(p)
            fnodes = [(x,) for x in _nodes1d(p)]
        elif nd == 3 and elemtype == 0:
            nodes = _tet_nodes(p)
            perm = _tet_perm(p, nodes)
            fnodes = _tri_nodes(p)
        elif nd == 3 and elemtype == 1:
            nodes = _hex_nodes(p)
            perm = _hex_perm(p)
            fnodes = _hex_face_nodes(p)
        else:
            raise ValueError("unsupported element family")
        self.plocvl = np.array([[float(v) for v in q] for q in nodes])
        self.plocfc = np.array([[float(v) for v in q] for q in fnodes])
        self.perm = np.ascontiguousarray(perm)          # (npf, nfe), 0-based element nodes
        self.permgeom = self.perm.copy()               # morder == porder
        self.npv = len(nodes)
        self.npf, self.nfe = self.perm.shape
        self.ndf = self.npf * self.nfe

        gp, gw = quad_rule(pgauss, nd, elemtype)
        self.gpvl, self.gwvl = gp, gw
        self.ngv = len(gw)
        shapvl = _shape(p, nodes, [tuple(r) for r in gp], elemtype)   # (npv, ngv, nd+1)
        self.shapvl = shapvl
        self.shapvt = np.transpose(shapvl, (1, 0, 2)).copy()           # (ngv, npv, nd+1)
        self.shapvg = shapvl * gw[None, :, None]                       # (npv, ngv, nd+1)
        # shapvgdotshapvl((ii-1)*npv+jj, g, d) = shapvg(jj,g,d) * shapvl(ii,g,1)
        sgl = self.shapvg[None, :, :, :] * shapvl[:, None, :, 0:1]     # [ii, jj, g, d]
        self.shapvgdotshapvl = sgl.reshape(self.npv * self.npv, self.ngv, nd + 1, order="C")
        # row index (ii-1)*npv + jj with jj fastest: C-order reshape of [ii, jj] gives ii*npv+jj

        fp, fw = quad_rule(pgauss, nd - 1, elemtype)
        self.gpfc, self.gwfc = fp, fw
        self.ngf = len(fw)
        shapfc = _shape(p, fnodes, [tuple(r) for r in fp], elemtype)   # (npf, ngf, nd)
        self.shapfc = shapfc
        self.shapft = np.transpose(shapfc, (1, 0, 2)).copy()           # (ngf, npf, nd)
        self.shapfg = shapfc * fw[None, :, None]                       # (npf, ngf, nd)
        sfc = self.shapfg[None, :, :, :] * shapfc[:, None, :, 0:1]
        self.shapfgdotshapfc = sfc.reshape(self.npf * self.npf, self.ngf, nd, order="C")

    def flat(self, name):
        """Column-major flattening, as written to master.bin and read by the C++."""
        return np.ascontiguousarray(np.asarray(getattr(self, name)).reshape(-1, order="F"))


@lru_cache(maxsize=None)
def make_master(nd, elemtype, porder, pgauss=None):
    if pgauss is None:
        pgauss = max(2 * porder, 1)
    return Master(nd, elemtype, porder, pgauss)
