"""Pins the CPU oracle: golden vectors from the reference's own element path,
plus checks fixed by the mathematics (exact derivatives, freestream
preservation, dense condensation, assembly-free BSR products)."""
import glob
import os

import numpy as np
import pytest

from hfem_amd.problem import make_problem, freestream_state

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {"tri2_p2": ("tri", 2, 2), "tri2_p4": ("tri", 2, 4), "tet1_p2": ("tet", 1, 2),
         "tet1_p3": ("tet", 1, 3), "hex3_p2": ("hex", 3, 2), "tet1_p4": ("tet", 1, 4),
         "tri2_p5": ("tri", 2, 5)}

# Matrices and condensed trace blocks agree to a few ulps of their norm; the
# element residual Ru (and inv(D) Ru) is a small difference of O(1) volume and
# face terms, so its relative error is amplified by that cancellation (the
# reference's own impl 0 and impl 1 differ by ~1e-12 on it).
RTOL_MAT = 1e-13
RTOL_RES = 1e-10


def relerr(a, b):
    a = np.asarray(a).reshape(len(b), -1)
    b = np.asarray(b).reshape(len(b), -1)
    return float((np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)).max())


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference_golden(oracle_mod, name):
    O = oracle_mod
    g = np.load(os.path.join(GOLD, f"elem_{name}.npz"))
    kind, nc, p = CASES[name]
    pr = make_problem(kind, nc, p, seed=1)
    # the fixture's inputs are the ones this generator produces
    assert np.array_equal(pr.UDG, g["UDG"]) and np.array_equal(pr.UH, g["UH"])
    assert np.array_equal(pr.dgnodes_ref(), g["dgnodes"]) and np.array_equal(pr.mesh.elcon, g["elcon"])
    keep = g["keep"]
    orc = O.Oracle(pr)
    out = {k: v[keep] for k, v in orc.assemble().items()}
    for k in ("D", "F", "K", "H", "Rh"):
        assert relerr(out[k], g[k]) < RTOL_MAT, (k, relerr(out[k], g[k]))
    assert relerr(out["Ru"], g["Ru"]) < RTOL_RES
    d = pr.dims
    for impl in (0, 1):
        if f"AH_impl{impl}" not in g:
            continue
        b = {k: v.copy() for k, v in out.items()}
        O.schur(impl, d["npv"], d["ncu"], b)
        assert relerr(b["H"], g[f"AH_impl{impl}"]) < RTOL_MAT
        assert relerr(b["Rh"], g[f"FH_impl{impl}"]) < RTOL_RES
        assert relerr(b["Ru"], g[f"DinvRu_impl{impl}"]) < RTOL_RES
        if f"DinvF_impl{impl}" in g:
            assert relerr(b["F"], g[f"DinvF_impl{impl}"]) < RTOL_MAT


def test_reference_live_when_built(oracle_mod):
    O = oracle_mod
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built")
    for kind, nc, p, seed in [("tri", 3, 3, 11), ("tet", 1, 3, 12)]:
        pr = make_problem(kind, nc, p, seed=seed)
        o = O.Oracle(pr).assemble()
        r = O.Reference(pr).element(impl=0)
        b = {k: v.copy() for k, v in o.items()}
        O.schur(0, pr.dims["npv"], pr.dims["ncu"], b)
        assert relerr(b["H"], r["AH"]) < RTOL_MAT
        assert relerr(b["F"], r["DinvF"]) < RTOL_MAT
        q_o = O.Oracle(pr).getq(pr.UDG)
        q_r = O.Reference(pr).getq(pr.UDG)
        assert np.abs(q_o - q_r).max() < 1e-12 * np.abs(q_r).max()


@pytest.mark.parametrize("kind,nd", [("tri", 2), ("tet", 3)])
def test_flux_jacobian_exact(oracle_mod, kind, nd):
    O = oracle_mod
    pr = make_problem(kind, 1, 1, seed=2)
    orc = O.Oracle(pr)
    rng = np.random.default_rng(4)
    u0 = freestream_state(nd, 1.4, 0.3)
    npt = 5
    udg = np.concatenate([u0[None, :] + 0.05 * rng.standard_normal((npt, nd + 2)),
                          0.1 * rng.standard_normal((npt, (nd + 2) * nd))], axis=1)
    f, fu = orc.flux(udg)
    # inviscid closed form at q = 0
    z = udg.copy()
    z[:, nd + 2:] = 0
    f0, _ = orc.flux(z)
    r, m, rE = z[:, 0], z[:, 1:nd + 1], z[:, nd + 1]
    p = 0.4 * (rE - 0.5 * (m ** 2).sum(1) / r)
    for d in range(nd):
        assert np.allclose(f0[:, 0, d], m[:, d], rtol=1e-15)
        for i in range(nd):
            assert np.allclose(f0[:, 1 + i, d], m[:, i] * m[:, d] / r + (i == d) * p, rtol=1e-14)
        assert np.allclose(f0[:, nd + 1, d], m[:, d] / r * (rE + p), rtol=1e-14)
    # Jacobian against central differences
    h = 1e-6
    for k in range(udg.shape[1]):
        e = np.zeros_like(udg)
        e[:, k] = h
        fd = (orc.flux(udg + e)[0] - orc.flux(udg - e)[0]) / (2 * h)
        assert np.allclose(fu[..., k], fd, atol=1e-7, rtol=1e-6), k


def test_fbou_jacobian(oracle_mod):
    O = oracle_mod
    pr = make_problem("tet", 1, 1, seed=2)
    orc = O.Oracle(pr)
    rng = np.random.default_rng(9)
    u0 = freestream_state(3, 1.4, 0.3)
    npt = 4
    udg = np.concatenate([u0[None] + 0.02 * rng.standard_normal((npt, 5)), rng.standard_normal((npt, 15))], 1)
    uh = u0[None] + 0.02 * rng.standard_normal((npt, 5))
    nl = rng.standard_normal((npt, 3))
    nl /= np.linalg.norm(nl, axis=1, keepdims=True)
    fb, fbu, fbuh = orc.fbou(udg, uh, nl)
    # boundary state equal to the far field makes the flux vanish
    z = np.tile(u0, (npt, 1))
    assert np.abs(orc.fbou(np.concatenate([z, udg[:, 5:]], 1), z, nl)[0]).max() < 1e-13
    h = 1e-6
    for k in range(5):
        e = np.zeros_like(uh)
        e[:, k] = h
        fd = (orc.fbou(udg, uh + e, nl)[0] - orc.fbou(udg, uh - e, nl)[0]) / (2 * h)
        assert np.allclose(fbuh[:, :, k], fd, atol=1e-7), k
        ev = np.zeros_like(udg)
        ev[:, k] = h
        fd = (orc.fbou(udg + ev, uh, nl)[0] - orc.fbou(udg - ev, uh, nl)[0]) / (2 * h)
        assert np.allclose(fbu[:, :, k], fd, atol=1e-7), k
    assert np.abs(fbu[:, :, 5:]).max() == 0.0


def test_element_residual_jacobian(oracle_mod):
    """D, F, K, H are minus the derivatives of Ru, Rh w.r.t. u and uh with q = q(u, uh)."""
    O = oracle_mod
    pr = make_problem("tri", 1, 2, seed=6)
    orc = O.Oracle(pr)
    d = pr.dims
    ie = 0
    U0 = orc.getq(pr.UDG)             # a state whose q satisfies the local q-equation
    base = orc.assemble(ie, 1, UDG=U0)
    D = O.mat(base["D"], d["na"], d["na"])[0]
    K = O.mat(base["K"], d["nb"], d["na"])[0]
    npv, ncu = d["npv"], d["ncu"]
    h = 1e-6
    for col in [0, npv + 2, 3 * npv + 1]:
        c, n = divmod(col, npv)
        vals = []
        for s in (h, -h):
            U = U0.copy()
            U[ie, c, n] += s
            U = orc.getq(U)           # q follows u through the local q-equation
            r = orc.assemble(ie, 1, UDG=U)
            vals.append((r["Ru"][0], r["Rh"][0]))
        dRu = (vals[0][0] - vals[1][0]) / (2 * h)
        dRh = (vals[0][1] - vals[1][1]) / (2 * h)
        assert np.allclose(-dRu, D[:, col], atol=2e-7), col
        assert np.allclose(-dRh, K[:, col], atol=2e-7), col


@pytest.mark.parametrize("kind,nc,p", [("tri", 2, 2), ("tet", 1, 3)])
def test_condensation_dense(oracle_mod, kind, nc, p):
    O = oracle_mod
    pr = make_problem(kind, nc, p, seed=8)
    d = pr.dims
    na, nb = d["na"], d["nb"]
    out = O.Oracle(pr).assemble(0, 2)
    b = {k: v.copy() for k, v in out.items()}
    O.schur(0, d["npv"], d["ncu"], b)
    for t in range(2):
        D = O.mat(out["D"], na, na)[t]
        F = O.mat(out["F"], na, nb)[t]
        K = O.mat(out["K"], nb, na)[t]
        H = O.mat(out["H"], nb, nb)[t]
        AH = H - K @ np.linalg.solve(D, F)
        FH = out["Rh"][t] - K @ np.linalg.solve(D, out["Ru"][t])
        assert relerr(O.mat(b["H"], nb, nb)[t][None], AH[None]) < 1e-13
        assert relerr(b["Rh"][t][None], FH[None]) < 1e-12
        # recovery identity for any trace state uh: u = inv(D)Ru - inv(D)F uh solves the
        # first block row, and the second block row then reads K u + H uh = Rh - FH + AH uh
        uh = np.random.default_rng(t).standard_normal(nb)
        u = b["Ru"][t] - O.mat(b["F"], na, nb)[t] @ uh
        assert np.allclose(D @ u + F @ uh, out["Ru"][t], rtol=0, atol=1e-11 * np.abs(F @ uh).max())
        lhs = K @ u + H @ uh - O.mat(b["H"], nb, nb)[t] @ uh
        assert np.allclose(lhs, out["Rh"][t] - b["Rh"][t], rtol=0, atol=1e-11 * np.abs(H @ uh).max())


@pytest.mark.parametrize("kind,nc,p", [("tri", 3, 2), ("tet", 1, 2)])
def test_freestream_preserved(oracle_mod, kind, nc, p):
    O = oracle_mod
    pr = make_problem(kind, nc, p, seed=0)
    u0 = freestream_state(pr.nd, 1.4, 0.3)
    pr.UDG[:] = 0
    pr.UDG[:, :pr.ncu, :] = u0[None, :, None]
    pr.UH[:] = u0[None, :]
    out = O.Oracle(pr).assemble()
    assert np.abs(out["Ru"]).max() < 1e-12
    _, Rg = O.scatter(pr, np.zeros((pr.ne, pr.dims["nb"] ** 2)), out["Rh"])
    assert np.abs(Rg).max() < 1e-12


def test_bsr_assembly_free_product(oracle_mod):
    O = oracle_mod
    pr = make_problem("tri", 3, 2, seed=5)
    d = pr.dims
    out = O.Oracle(pr).assemble()
    O.schur(0, d["npv"], d["ncu"], out)
    Hg, Rg = O.scatter(pr, out["H"], out["Rh"])
    x = np.random.default_rng(1).standard_normal(pr.mesh.nf * d["npf"] * d["nch"])
    y = O.bsr_matvec(pr, Hg, x)
    yref = np.zeros_like(x)
    AH = O.mat(out["H"], d["nb"], d["nb"])
    for ie in range(pr.ne):
        g = pr.mesh.elcon[ie].reshape(-1)
        idx = (g[:, None] * d["nch"] + np.arange(d["nch"])[None, :]).reshape(-1)
        yref[idx] += AH[ie] @ x[idx]
    assert np.allclose(y, yref, rtol=1e-13, atol=1e-13)
    fref = np.zeros_like(x)
    for ie in range(pr.ne):
        g = pr.mesh.elcon[ie].reshape(-1)
        idx = (g[:, None] * d["nch"] + np.arange(d["nch"])[None, :]).reshape(-1)
        fref[idx] += out["Rh"][ie]
    assert np.allclose(Rg, fref, rtol=1e-14, atol=1e-14)
