"""Master tables, quadrature and mesh connectivity against what mathematics fixes."""
from math import factorial

import numpy as np
import pytest

from hfem_amd.master import make_master, _shape, _tri_nodes, _tet_nodes, _hex_nodes
from hfem_amd.quadrature import quad_rule
from hfem_amd.mesh import square_tri_mesh, cube_tet_mesh, cube_hex_mesh


def simplex_monomial(exps):
    """int over the unit simplex of prod x_i^a_i = prod a_i! / (sum a + d)!"""
    d = len(exps)
    return np.prod([factorial(a) for a in exps]) / factorial(sum(exps) + d)


@pytest.mark.parametrize("dim,pg", [(2, 2), (2, 4), (2, 6), (2, 8), (2, 9), (2, 10), (3, 2), (3, 4), (3, 6), (3, 8)])
def test_simplex_rules_exact(dim, pg):
    x, w = quad_rule(pg, dim, 0)
    # the tabulated weights/points carry 14-16 significant digits
    tol = 2e-13 if dim == 3 else 4e-15
    import itertools
    for e in itertools.product(range(pg + 1), repeat=dim):
        if sum(e) > pg:
            continue
        approx = np.sum(w * np.prod(x ** np.array(e)[None, :], axis=1))
        assert abs(approx - simplex_monomial(e)) < tol, (e, approx, simplex_monomial(e))


@pytest.mark.parametrize("pg", [2, 4, 6])
def test_tensor_and_line_rules_exact(pg):
    x, w = quad_rule(pg, 1, 0)
    for a in range(pg + 1):
        assert abs(np.sum(w * x[:, 0] ** a) - 1.0 / (a + 1)) < 1e-15
    x, w = quad_rule(pg, 3, 1)
    for a in range(pg + 1):
        assert abs(np.sum(w * x[:, 0] ** a * x[:, 2] ** (pg - a)) - 1.0 / ((a + 1) * (pg - a + 1))) < 1e-14


@pytest.mark.parametrize("nd,et,p", [(2, 0, 1), (2, 0, 2), (2, 0, 4), (3, 0, 1), (3, 0, 3), (3, 0, 4), (3, 1, 2)])
def test_nodal_basis(nd, et, p):
    m = make_master(nd, et, p)
    nodes = {(2, 0): _tri_nodes, (3, 0): _tet_nodes, (3, 1): _hex_nodes}[(nd, et)](p)
    # Kronecker property at the nodes
    at_nodes = _shape(p, nodes, nodes, et)
    assert np.allclose(at_nodes[:, :, 0], np.eye(m.npv), atol=1e-15)
    # partition of unity and zero derivative sums at Gauss points
    assert np.allclose(m.shapvt[:, :, 0].sum(1), 1.0, atol=1e-14)
    assert np.allclose(m.shapvt[:, :, 1:].sum(1), 0.0, atol=1e-12)
    # shapvg = shapvl diag(w)  (master/mkmasterelement.m:72)
    assert np.allclose(m.shapvg, m.shapvl * m.gwvl[None, :, None], rtol=0, atol=0)
    # exact integration of the mass matrix row sums: sum_j int phi_i phi_j = int phi_i
    mass = np.einsum("ig,jg,g->ij", m.shapvl[:, :, 0], m.shapvl[:, :, 0], m.gwvl)
    assert np.allclose(mass.sum(1), m.shapvg[:, :, 0].sum(1), atol=1e-15)
    # shapvgdotshapvl((ii-1)*npv+jj, g, d) = shapvg(jj,g,d) shapvl(ii,g,1)
    ii, jj = 1, m.npv - 1
    assert np.allclose(m.shapvgdotshapvl[ii * m.npv + jj], m.shapvg[jj] * m.shapvl[ii, :, 0:1])


@pytest.mark.parametrize("nd,et,p", [(2, 0, 2), (3, 0, 3), (3, 1, 2)])
def test_face_permutations_lie_on_faces(nd, et, p):
    m = make_master(nd, et, p)
    x = m.plocvl
    for a in range(m.nfe):
        pts = x[m.perm[:, a]]
        # all face nodes are coplanar and the face shape at face nodes reproduces them
        centred = pts - pts.mean(0)
        assert np.linalg.matrix_rank(centred, tol=1e-12) == nd - 1
        assert len(set(m.perm[:, a].tolist())) == m.npf


@pytest.mark.parametrize("maker,args", [(square_tri_mesh, (3, 2)), (cube_tet_mesh, (2, 3)), (cube_hex_mesh, (3, 2))])
def test_mesh_connectivity(maker, args):
    msh = maker(*args, seed=3)
    m = msh.master
    assert msh.involution_ok()
    # every interior face seen from both sides maps to the same physical nodes
    for fid, (e1, a1, e2, a2) in enumerate(msh.f2e):
        if e2 < 0:
            continue
        g1 = msh.elcon[e1, a1]
        g2 = msh.elcon[e2, a2]
        assert sorted(g1.tolist()) == sorted(g2.tolist())
        x1 = msh.dgnodes[e1, m.perm[:, a1]][np.argsort(g1)]
        x2 = msh.dgnodes[e2, m.perm[:, a2]][np.argsort(g2)]
        d = x1 - x2
        if msh.periodic:
            d -= np.round(d)
        assert np.abs(d).max() < 1e-12
    # trace graph: self block first, neighbours ascending, pattern symmetric
    rs, cj = msh.ent2entStart, msh.ent2ent
    pairs = set()
    for i in range(msh.nf):
        row = cj[rs[i]:rs[i + 1]]
        assert row[0] == i and np.all(np.diff(row[1:]) > 0)
        pairs |= {(i, int(j)) for j in row}
    assert all((j, i) in pairs for (i, j) in pairs)


def test_positive_jacobians():
    msh = cube_tet_mesh(3, 1, seed=5)
    a = msh.dgnodes
    vol = np.einsum("ij,ij->i", np.cross(a[:, 1] - a[:, 0], a[:, 2] - a[:, 0]), a[:, 3] - a[:, 0]) / 6
    assert np.all(vol > 0)
    assert abs(vol.sum() - 1.0) < 1e-12
