"""The CPU oracle of the trace-system solver pinned by properties the mathematics fixes.

The reference's C++ binds GMRES, BILU0 and the MDF ordering from closed
libraries (solvers/dgsolvers.cpp:16-20), so no reference output exists for
them: the oracle follows the MATLAB specifications (kernel/gmres_bilu0.m,
computeBILU0.m, applyBILU0.m, computeMDFordering.m, computeExactMDFordering.m,
mkBlockJacobi.m) and is pinned here by
* BILU0: the ILU(0) identity (L U)_ij = A_ij on the block pattern, and
  exactness on a block path graph (no fill);
* MDF: the exact weights equal the Frobenius norm of the fill that eliminating
  a row first would discard (dense elimination), and a path graph is ordered
  0, 1, 2, ... (no discarded fill, ties to the first index as MATLAB's min);
* GMRES: the solution against a dense direct solve, monotone residual
  histories, and the left-preconditioned stopping test;
* getU: the recovered element update solves the element's first block row
  D du + F duh = Ru.
"""
import numpy as np
import pytest

from hfem_amd.problem import make_problem


def path_bsr(O, n, bsz, seed=0):
    """Block tridiagonal matrix on a path graph 0-1-...-(n-1), diagonal block first."""
    rng = np.random.default_rng(seed)
    rowpts, cols, blocks = [0], [], []
    for i in range(n):
        nb = [i] + [j for j in (i - 1, i + 1) if 0 <= j < n]
        for j in nb:
            B = rng.standard_normal((bsz, bsz)) * 0.3
            if i == j:
                B += 4 * np.eye(bsz)
            cols.append(j)
            blocks.append(B.T.reshape(-1))      # column-major
        rowpts.append(len(cols))
    return O.BSR(rowpts, cols, np.concatenate(blocks), bsz)


@pytest.fixture(scope="module")
def tri_system(oracle_mod):
    prob = make_problem("tri", 2, 2, seed=5)
    M, Rg, rnorm, blocks = oracle_mod.trace_system(prob)
    return prob, M, Rg, rnorm, blocks


def lu_from_bilu0(M, o2u, u2o, Mg):
    """Dense L (unit) and U in the original row numbering from the factored blocks."""
    n, b = M.nrows, M.bsz
    L = np.eye(n * b)
    U = np.zeros((n * b, n * b))
    blk = Mg.reshape(-1, b, b).transpose(0, 2, 1)
    for i in range(n):
        for k in range(M.rowpts[i], M.rowpts[i + 1]):
            j = M.colind[k]
            S = (slice(i * b, (i + 1) * b), slice(j * b, (j + 1) * b))
            if j == i:
                U[S] = np.linalg.inv(blk[k])
            elif u2o[j] < u2o[i]:
                L[S] = blk[k]
            else:
                U[S] = blk[k]
    return L, U


def test_bilu0_ilu0_identity(oracle_mod, tri_system):
    O = oracle_mod
    prob, M, *_ = tri_system
    o2u, u2o = O.mdf_order(M)
    assert sorted(o2u) == list(range(M.nrows))
    assert np.all(u2o[o2u] == np.arange(M.nrows))
    Mg, info = O.bilu0_factor(M, o2u, u2o)
    assert info == 0
    L, U = lu_from_bilu0(M, o2u, u2o, Mg)
    A = M.dense()
    LU = L @ U
    b = M.bsz
    err = 0.0
    for i in range(M.nrows):
        for k in range(M.rowpts[i], M.rowpts[i + 1]):
            j = M.colind[k]
            S = (slice(i * b, (i + 1) * b), slice(j * b, (j + 1) * b))
            err = max(err, np.abs(LU[S] - A[S]).max() / np.abs(A).max())
    assert err < 1e-12
    # apply = (L U)^{-1} x
    x = np.random.default_rng(1).standard_normal(M.nrows * b)
    v = O.bilu0_apply(M, o2u, u2o, Mg, x)
    # backward error of the two block triangular solves
    assert np.linalg.norm(LU @ v - x) < 1e-13 * np.linalg.norm(L, 2) * np.linalg.norm(U, 2) * np.linalg.norm(v)


def test_bilu0_and_mdf_exact_on_path(oracle_mod):
    O = oracle_mod
    M = path_bsr(O, 9, 4)
    for exact in (False, True):
        o2u, u2o = O.mdf_order(M, exact=exact)
        assert list(o2u) == list(range(9))
    Mg, info = O.bilu0_factor(M, o2u, u2o)
    x = np.random.default_rng(2).standard_normal(9 * 4)
    v = O.bilu0_apply(M, o2u, u2o, Mg, x)
    np.testing.assert_allclose(v, np.linalg.solve(M.dense(), x), rtol=0, atol=1e-12)


def test_mdf_exact_weight_is_discarded_fill(oracle_mod, tri_system):
    """The first exact-MDF pivot minimises the norm of the discarded fill of a
    dense block elimination step (computeExactMDFordering.m:23-45, :93)."""
    O = oracle_mod
    _, M, *_ = tri_system
    A = M.dense()
    b = M.bsz
    pattern = np.zeros((M.nrows, M.nrows), bool)
    for i in range(M.nrows):
        pattern[i, M.colind[M.rowpts[i]:M.rowpts[i + 1]]] = True
    w = np.zeros(M.nrows)
    for k in range(M.nrows):
        K = slice(k * b, (k + 1) * b)
        upd = A[:, K] @ np.linalg.solve(A[K, K], A[K, :])
        fill = 0.0
        for i in range(M.nrows):
            for j in range(M.nrows):
                if i != k and j != k and i != j and pattern[i, k] and pattern[k, j] and not pattern[i, j]:
                    fill += np.sum(upd[i * b:(i + 1) * b, j * b:(j + 1) * b] ** 2)
        w[k] = np.sqrt(fill)
    o2u, _ = O.mdf_order(M, exact=True)
    assert w[o2u[0]] <= w.min() * (1 + 1e-12) + 1e-300
    assert o2u[0] == int(np.flatnonzero(w <= w.min() * (1 + 1e-12))[0])


@pytest.mark.parametrize("prec", ["bilu0", "bj", "none"])
@pytest.mark.parametrize("orthog", [0, 1])
def test_gmres_solves_trace_system(oracle_mod, tri_system, prec, orthog):
    O = oracle_mod
    _, M, Rg, *_ = tri_system
    kw = {}
    if prec == "bilu0":
        o2u, u2o = O.mdf_order(M)
        Mg, _ = O.bilu0_factor(M, o2u, u2o)
        kw = dict(prec=O.PREC_BILU0, Mg=Mg, o2u=o2u, u2o=u2o)
    elif prec == "bj":
        Dinv, _ = O.block_jacobi_factor(M)
        kw = dict(prec=O.PREC_BJ, Mg=Dinv)
    else:
        kw = dict(prec=O.PREC_NONE)
    # unpreconditioned GMRES(20) stagnates on this system (condition ~1e7): give it the full space
    restart = 20 if prec != "none" else len(Rg)
    x, flag, iters, rev = O.gmres(M, Rg, restart=restart, tol=1e-12, maxit=2000, orthog=orthog, **kw)
    assert flag == 0
    xd = np.linalg.solve(M.dense(), Rg)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8
    assert rev[-1] <= 1e-12
    # residual norms within a restart cycle never increase
    assert np.all(np.diff(rev[:min(len(rev), restart)]) <= 1e-14)
    # the stopping test is on the preconditioned residual: check it independently
    if prec == "bj":
        Dinv = kw["Mg"].reshape(-1, M.bsz, M.bsz).transpose(0, 2, 1)
        pr = lambda v: np.concatenate([Dinv[i] @ v[i * M.bsz:(i + 1) * M.bsz] for i in range(M.nrows)])
        x8, flag8, _, rev8 = O.gmres(M, Rg, restart=20, tol=1e-8, maxit=2000, orthog=orthog, **kw)
        ratio = np.linalg.norm(pr(Rg - M.matvec(x8))) / np.linalg.norm(pr(Rg))
        assert flag8 == 0 and ratio <= 1.001e-8 and abs(ratio - rev8[-1]) < 1e-12


def test_gmres_unpreconditioned_full_krylov_is_exact(oracle_mod):
    O = oracle_mod
    M = path_bsr(O, 5, 3, seed=3)
    b = np.random.default_rng(4).standard_normal(15)
    x, flag, iters, _ = O.gmres(M, b, prec=O.PREC_NONE, restart=15, tol=1e-14, maxit=15)
    assert iters <= 15
    np.testing.assert_allclose(x, np.linalg.solve(M.dense(), b), rtol=1e-10, atol=1e-12)


def test_update_u_solves_first_block_row(oracle_mod, tri_system):
    O = oracle_mod
    prob, M, Rg, rnorm, blocks = tri_system
    d = prob.dims
    orc = O.Oracle(prob)
    raw = orc.assemble()
    duh = np.random.default_rng(6).standard_normal(prob.UH.size)
    UDG0 = np.zeros_like(prob.UDG)
    U1 = O.update_u(prob, blocks["F"], blocks["Ru"], duh, 1.0, UDG0)
    na, nb, ndf, nch, npv, nc = d["na"], d["nb"], d["ndf"], d["nch"], d["npv"], d["nc"]
    for e in range(prob.ne):
        du = U1[e].reshape(nc, npv)[: d["ncu"]].reshape(-1)
        el = prob.mesh.elcon.reshape(prob.ne, ndf)[e]
        de = duh.reshape(-1, nch)[el].reshape(-1)
        D = raw["D"][e].reshape(na, na).T
        F = raw["F"][e].reshape(nb, na).T
        np.testing.assert_allclose(D @ du + F @ de, raw["Ru"][e], rtol=0, atol=1e-10 * np.abs(raw["Ru"][e]).max())


def test_newton_converges_quadratically(oracle_mod):
    """solveNonlinearProblem on a far-field box flow at 0.3 rad (the normal velocity
    does not vanish on any boundary face, so the residual is smooth)."""
    import newton_oracle as N
    prob = make_problem("tri", 3, 2, seed=5, angle=0.3)
    UDG, UH, ok, hist = N.newton_solve(prob, prob.UDG, prob.UH, prob.SH, newton_tol=1e-11, maxiter=6,
                                       restart=30, gmres_tol=1e-12)
    assert ok
    r = [h["old"] for h in hist] + [hist[-1]["new"]]
    assert len(hist) <= 4
    # quadratic: r_{k+1} <= C r_k^2 once in the asymptotic range
    assert r[-1] < 1e-11 and r[-2] < 1e-4 and r[-1] < 10 * r[-2] ** 2
    assert all(h["alphas"] == [1.0] for h in hist)


def test_dirk_coefficients_from_butcher_tables():
    """d = inv(A), c = b inv(A), t of kernel/hdg_solve_dirk.m:218-227 against the
    constants of solvers/DIRKcoeff.cpp."""
    import newton_oracle as N
    a1 = 0.4358665215
    t1 = (1 + a1) / 2
    b1 = -(6 * a1 ** 2 - 16 * a1 + 1) / 4
    b2 = (6 * a1 ** 2 - 20 * a1 + 5) / 4
    A = np.array([[a1, 0, 0], [t1 - a1, a1, 0], [b1, b2, a1]])
    b = np.array([b1, b2, a1])
    c, d, t = N.dirk_coefficients(3, 3)
    np.testing.assert_allclose(d.reshape(3, 3, order="F"), np.linalg.inv(A), rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(c, b @ np.linalg.inv(A), rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(t, [a1, t1, 1.0], rtol=1e-9)
    s2 = np.sqrt(2.0)
    A2 = np.array([[1 - s2 / 2, 0], [s2 / 2, 1 - s2 / 2]])
    c, d, t = N.dirk_coefficients(2, 2)
    np.testing.assert_allclose(d.reshape(2, 2, order="F"), np.linalg.inv(A2), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(c, np.array([s2 / 2, 1 - s2 / 2]) @ np.linalg.inv(A2), atol=1e-12)


def test_dirk_step_keeps_a_uniform_state(oracle_mod):
    """Every stage source reproduces a constant state exactly (consistency of the
    stage formulas of solvers/dgsolvers.cpp:316-341)."""
    import newton_oracle as N
    prob = make_problem("tri", 2, 2, seed=1, angle=0.3, amp=0.0, qamp=0.0)
    UDG, UH, SH, hist = N.dirk_step(prob, prob.UDG, prob.UH, prob.SH, dt=0.1, newton_tol=1e-12, restart=30,
                                    gmres_tol=1e-12)
    np.testing.assert_allclose(UDG, prob.UDG, rtol=0, atol=1e-12)
    np.testing.assert_allclose(UH, prob.UH, rtol=0, atol=1e-12)
    assert all(len(h) == 1 and h[0]["old"] < 1e-10 for h in hist)


def scalar_block_graph(O, tri_system, seed=7):
    """The trace graph of a small triangle mesh (every row couples to 4 faces,
    pairs of neighbours that are not neighbours of each other, so eliminating a
    row discards fill) with blocks h_ij * I, h_ij random and not symmetric, and
    every diagonal block sqrt(bsz) * I."""
    _, M, *_ = tri_system
    rng = np.random.default_rng(seed)
    b = M.bsz
    blocks = []
    for i in range(M.nrows):
        for k in range(M.rowpts[i], M.rowpts[i + 1]):
            h = np.sqrt(b) if M.colind[k] == i else rng.uniform(0.05, 2.0)
            blocks.append((h * np.eye(b)).reshape(-1))
    return O.BSR(M.rowpts, M.colind, np.concatenate(blocks), b)


def test_approximate_mdf_equals_exact_on_scalar_blocks(oracle_mod, tri_system):
    """With blocks h_ij I and diagonal blocks sqrt(bsz) I, the approximate weight
    C_ik C_kj = |h_ik| |h_kj| of kernel/computeMDFordering.m:24-47 equals the exact
    discarded-fill weight ||H_ik inv(H_kk) H_kj||_F of computeExactMDFordering.m,
    so the two greedy orderings coincide -- and both equal a brute-force greedy
    that, at every step, eliminates densely each unnumbered row k of the original
    matrix and measures the fill it would discard among the unnumbered rows
    (pairs i, j of neighbours of k that are not neighbours of each other), taking
    the smallest (lowest row on ties).  A weight that counts adjacent pairs,
    numbered rows or the diagonal block, or a wrong update set, fails it."""
    O = oracle_mod
    M = scalar_block_graph(O, tri_system)
    o2u_a, _ = O.mdf_order(M, exact=False)
    o2u_e, _ = O.mdf_order(M, exact=True)
    A, b, n = M.dense(), M.bsz, M.nrows
    pattern = np.zeros((n, n), bool)
    for i in range(n):
        pattern[i, M.colind[M.rowpts[i]:M.rowpts[i + 1]]] = True
    numbered = np.zeros(n, bool)
    brute = []
    for _ in range(n):
        w = np.full(n, np.inf)
        for k in np.flatnonzero(~numbered):
            K = slice(k * b, (k + 1) * b)
            upd = A[:, K] @ np.linalg.solve(A[K, K], A[K, :])
            live = [i for i in range(n) if not numbered[i] and i != k and pattern[i, k]]
            w[k] = np.sqrt(sum(np.sum(upd[i * b:(i + 1) * b, j * b:(j + 1) * b] ** 2)
                               for i in live for j in live if i != j and not pattern[i, j]))
        p = int(np.flatnonzero(w <= w.min() * (1 + 1e-12))[0])
        brute.append(p)
        numbered[p] = True
    assert list(o2u_e) == brute
    assert list(o2u_a) == brute


def test_mdf_exact_first_pivot_on_scalar_blocks(oracle_mod, tri_system):
    """On the scalar-block graph the first pivot of both orderings is the row whose
    dense block elimination discards the least fill."""
    O = oracle_mod
    M = scalar_block_graph(O, tri_system)
    A, b = M.dense(), M.bsz
    pattern = np.zeros((M.nrows, M.nrows), bool)
    for i in range(M.nrows):
        pattern[i, M.colind[M.rowpts[i]:M.rowpts[i + 1]]] = True
    w = np.zeros(M.nrows)
    for k in range(M.nrows):
        K = slice(k * b, (k + 1) * b)
        upd = A[:, K] @ np.linalg.solve(A[K, K], A[K, :])
        w[k] = np.sqrt(sum(np.sum(upd[i * b:(i + 1) * b, j * b:(j + 1) * b] ** 2)
                           for i in range(M.nrows) for j in range(M.nrows)
                           if len({i, j, k}) == 3 and pattern[i, k] and pattern[k, j] and not pattern[i, j]))
    for exact in (False, True):
        assert O.mdf_order(M, exact=exact)[0][0] == int(np.argmin(w))


def _ls_run(N, r_of_net, old, tol=1e-12, robust=None, qn=False):
    net = [0.0]

    def apply_step(a):
        net[0] += a

    alphas, r, failed = N.line_search(apply_step, lambda: r_of_net(net[0]), old, tol, 1e-4, robust, qn)
    return alphas, net[0], r, failed


def test_line_search_known_answers():
    """solveNonlinearProblem's line search (solvers/dgsolvers.cpp:152-187) on
    residuals of the net step s, against sequences derived by hand from that code:
    a full step that overshoots is retracted by half (1, -1/2); two retractions
    give 1, -1/2, -1/4; a step that never helps is halved 14 times until alpha
    = 2^-14 < 1e-4, then the last applied step is undone twice over (-2^-13)
    back to the previous iterate -- the first time in a solve only (robustMode);
    a second failure in the same solve stays at net 2^-13; with quasi-Newton it
    always goes back."""
    import newton_oracle as N
    a, net, r, f = _ls_run(N, lambda s: abs(s - 0.45) + 0.01, old=0.46)
    assert a == [1.0, -0.5] and net == 0.5 and not f
    a, net, r, f = _ls_run(N, lambda s: abs(s - 0.2) + 0.01, old=0.21)
    assert a == [1.0, -0.5, -0.25] and net == 0.25 and not f
    halvings = [-(2.0 ** -k) for k in range(1, 14)]
    robust = [0]
    a, net, r, f = _ls_run(N, lambda s: 1.0 + s, old=1.0, robust=robust)
    assert f and a == [1.0] + halvings + [-(2.0 ** -13)] and net == 0.0 and robust == [1]
    a, net, r, f = _ls_run(N, lambda s: 1.0 + s, old=1.0, robust=robust)
    assert f and a == [1.0] + halvings and net == 2.0 ** -13
    a, net, r, f = _ls_run(N, lambda s: 1.0 + s, old=1.0, robust=[1], qn=True)
    assert f and a[-1] == -(2.0 ** -13) and net == 0.0
    # NaN counts as a failed step; a residual below the Newton tolerance is accepted
    a, net, r, f = _ls_run(N, lambda s: float("nan") if s > 0.6 else 0.5, old=0.4)
    assert a[:2] == [1.0, -0.5] and net == 0.5 and not f and r == 0.5 or (f and net == 0.0)
    a, net, r, f = _ls_run(N, lambda s: 1e-13 + s * 0, old=1e-14, tol=1e-12)
    assert a == [1.0] and not f


@pytest.mark.parametrize("nstage,order", [(1, 1), (2, 2), (2, 3), (3, 3), (3, 4)])
def test_dirk_stage_sources_reproduce_butcher_dirk(nstage, order):
    """The stage sources of solveUnsteadyProblem (solvers/dgsolvers.cpp:316-341)
    with fc_u = d_ii/dt, on the linear test equation u' = -lam u solved exactly
    per stage ((d_ii/dt + lam) U_i = SH_i), give the stage values of the DIRK
    defined by its Butcher matrix A = inv(d): (I + dt lam A) U = u_n 1.  The
    Butcher matrices of (3,3) and (2,2) are kernel/hdg_solve_dirk.m:218-227's; the
    others are inv(d) of the coefficient table (pinned for (3,3) and (2,2) by
    test_dirk_coefficients_from_butcher_tables).  A transposed d index, a dropped
    stage term or the wrong previous-stage source fails it."""
    import newton_oracle as N
    c, d, t = N.dirk_coefficients(nstage, order)
    dm = d.reshape(nstage, nstage, order="F")
    if (nstage, order) == (3, 3):
        a1 = 0.4358665215
        t1 = (1 + a1) / 2
        A = np.array([[a1, 0, 0], [t1 - a1, a1, 0],
                      [-(6 * a1 ** 2 - 16 * a1 + 1) / 4, (6 * a1 ** 2 - 20 * a1 + 5) / 4, a1]])
        tol = 1e-9
    elif (nstage, order) == (2, 2):
        s2 = np.sqrt(2.0)
        A = np.array([[1 - s2 / 2, 0], [s2 / 2, 1 - s2 / 2]])
        tol = 1e-12
    else:
        A = np.linalg.inv(dm)
        tol = 1e-12
    rng = np.random.default_rng(nstage * 10 + order)
    dt = 0.3
    un = rng.uniform(0.5, 2.0, size=(4, 3, 5))
    lam = rng.uniform(0.1, 5.0, size=un.shape)
    U, SH = un.copy(), np.zeros_like(un)
    stages = []
    for s in range(nstage):
        SH = N.stage_source(s, nstage, d, dt, un, U, SH)
        U = SH / (dm[s, s] / dt + lam)
        stages.append(U)
    for idx in np.ndindex(un.shape):
        direct = np.linalg.solve(np.eye(nstage) + dt * lam[idx] * A, un[idx] * np.ones(nstage))
        got = np.array([st[idx] for st in stages])
        np.testing.assert_allclose(got, direct, rtol=tol, atol=0)
