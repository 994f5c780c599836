"""Newton and DIRK drivers of the CPU oracle (test infrastructure only).

Composes the oracle's element path, condensation, scatter, ordering,
preconditioner and GMRES exactly as the reference's drivers call them:

* ``newton_solve``: solveNonlinearProblem, solvers/dgsolvers.cpp:142-205 --
  assemble (rNorm of the current state), MDF ordering, preconditioner, GMRES,
  then the line search that applies alpha = 1 and, while the residual grows,
  retracts by half the last step (sol.alpha = -|alpha/2|) down to 1e-4, where it
  returns to the previous iterate (robust mode).
* ``update_solution``: updateSol, solvers/solutionupdate.cpp:69-98 -- UH += alpha
  duh, getU on every element, then getQ (assembly/elementassembly.cpp:1414-1479).
* ``residual_norm``: computeResidualNorm, assembly/globalassembly.cpp:618-678.
* ``dirk_coefficients`` / ``dirk_step``: DIRKcoeff (solvers/DIRKcoeff.cpp:6-118)
  and the stage loop of solveUnsteadyProblem (solvers/dgsolvers.cpp:221-444).
* ``trace_system_qn`` / ``assemble_rhs`` / ``newton_solve_qn``: the quasi-Newton
  path -- updateQuasiNewtonStructures (assembly/globalassembly.cpp:327-389),
  assembleRHS (:559-616) with schurElementVector (assembly/elementschur.cpp:345-458)
  and the Jacobian reuse of solveLinearProblem / solveNonlinearProblem
  (solvers/dgsolvers.cpp:40-205).  The stored operand is inv(D) (the explicit
  inverse this build's condensation forms) rather than the reference's LU factors;
  in exact arithmetic both give the same inv(D) Ru.
"""
import dataclasses

import numpy as np

import oracle as O


def residual_norm(prob, UDG, UH, SH):
    """sqrt(sum_e |Ru_e|^2 + |r|^2), r = the assembled uncondensed Rh."""
    orc = O.Oracle(prob)
    b = orc.assemble(0, prob.ne, UDG=UDG, UH=UH, SH=SH)
    d = prob.dims
    _, r = O.scatter(prob, np.zeros((prob.ne, d["nb"] * d["nb"])), b["Rh"])
    return float(np.sqrt((b["Ru"] ** 2).sum() + (r ** 2).sum()))


def update_solution(prob, UDG, UH, SH, blocks, duh, alpha):
    """updateSol: returns the new (UDG, UH)."""
    UH = UH + alpha * duh.reshape(UH.shape)
    UDG = O.update_u(prob, blocks["F"], blocks["Ru"], duh, alpha, UDG)
    if prob.fc_q != 0:
        UDG = O.Oracle(prob).getq(UDG, UH, SH)
    return UDG, UH


def linear_solve(prob, M, Rg, prec="bilu0", restart=30, gmres_tol=1e-10, maxit=1000, orthog=1, exact_mdf=False,
                 order=None):
    """computeOrdering + computePreconditioner + gmres (solvers/dgsolvers.cpp:79-105).
    ``order`` is a dict caching the ordering (reuseOrdering): computed when empty."""
    if prec == "bilu0":
        if order is None:
            order = {}
        if "o2u" not in order:
            order["o2u"], order["u2o"] = O.mdf_order(M, exact=exact_mdf)
        o2u, u2o = order["o2u"], order["u2o"]
        Mg, _ = O.bilu0_factor(M, o2u, u2o)
        return O.gmres(M, Rg, O.PREC_BILU0, Mg, o2u, u2o, restart=restart, tol=gmres_tol, maxit=maxit, orthog=orthog)
    if prec == "bj":
        Dinv, _ = O.block_jacobi_factor(M)
        return O.gmres(M, Rg, O.PREC_BJ, Dinv, restart=restart, tol=gmres_tol, maxit=maxit, orthog=orthog)
    return O.gmres(M, Rg, O.PREC_NONE, restart=restart, tol=gmres_tol, maxit=maxit, orthog=orthog)


def line_search(apply_step, residual, old, newton_tol, alpha_min=1e-4, robust=None, quasi_newton=False):
    """The line search of solveNonlinearProblem (solvers/dgsolvers.cpp:152-187):
    apply alpha = 1; while the new residual is NaN/Inf or above both the old one
    and the tolerance, retract by half of the last step (sol.alpha = -|alpha/2|,
    applied on top of what is there); once the step falls below alpha_min, go back
    to the previous iterate (-2|alpha|) -- but only in the first such failure of a
    nonlinear solve or with quasi-Newton (robustMode, reset to 0 at the start of
    the solve, :114, set at :177).  apply_step(alpha) applies a step,
    residual() returns the residual norm after it.  ``robust`` is a one-element
    list holding robustMode across the Newton iterations of one solve.  Returns
    (alphas applied, final residual, whether the line search failed)."""
    robust = [0] if robust is None else robust
    alphas = []
    alpha = 1.0
    while True:
        apply_step(alpha)
        alphas.append(alpha)
        rnorm = residual()
        if not np.isfinite(rnorm) or (rnorm > old and rnorm > newton_tol):
            a = abs(alpha / 2)
            alpha = -a
            if a < alpha_min:
                if robust[0] == 0 or quasi_newton:
                    robust[0] = 1
                    alpha = -2.0 * abs(alpha)
                    apply_step(alpha)
                    alphas.append(alpha)
                return alphas, rnorm, True
        else:
            return alphas, rnorm, False


def newton_solve(prob, UDG, UH, SH, newton_tol=1e-7, maxiter=10, alpha_min=1e-4, order=None, **lin):
    """Returns (UDG, UH, converged, history); history holds one dict per Newton iteration.
    The ordering is computed at the first iteration and reused (reuseOrdering,
    solvers/dgsolvers.cpp:87-92); pass ``order`` to share it between solves."""
    order = {} if order is None else order
    UDG = np.array(UDG, dtype=float)
    UH = np.array(UH, dtype=float)
    rnorm = newton_tol + 1.0
    history = []
    it = 0
    robust = [0]
    while rnorm > newton_tol and it < maxiter:
        it += 1
        M, Rg, old, blocks = O.trace_system(prob, UDG=UDG, UH=UH, SH=SH)
        if not np.isfinite(old):
            raise FloatingPointError("residual norm is NaN or Inf")
        x, flag, iters, rev = linear_solve(prob, M, Rg, order=order, **lin)
        rec = dict(old=old, gmres_iters=iters, gmres_flag=flag, alphas=[])
        history.append(rec)
        if it == 1 and old < newton_tol:
            rnorm = old
            rec["new"] = rnorm
            break
        state = [UDG, UH]

        def apply_step(alpha):
            state[0], state[1] = update_solution(prob, state[0], state[1], SH, blocks, x, alpha)

        rec["alphas"], rnorm, _ = line_search(apply_step, lambda: residual_norm(prob, state[0], state[1], SH),
                                              old, newton_tol, alpha_min, robust)
        UDG, UH = state
        rec["new"] = rnorm
    return UDG, UH, rnorm < newton_tol, history


def dirk_coefficients(nstage, order):
    """(c, d, t) of solvers/DIRKcoeff.cpp; d column-major nstage x nstage."""
    table = {
        (1, 1): ([1.0], [1.0], [1.0]),
        (1, 2): ([2.0], [2.0], [0.5]),
        (2, 2): ([0.0, 1.0], [3.414213562373096, -8.242640687119289, 0.0, 3.414213562373096],
                 [0.292893218813452, 1.0]),
        (2, 3): ([1.098076211353316, 0.633974596215561],
                 [1.267949192431123, 0.928203230275509, 0.0, 1.267949192431123],
                 [0.788675134594813, 0.211324865405187]),
        (3, 3): ([0.0, 0.0, 1.0],
                 [2.294280360323567, -1.484721005721436, -8.556127801733982, 0.0, 2.294280360323568,
                  3.391748836909791, 0.0, 0.0, 2.294280360323567],
                 [0.435866521500000, 0.717933260750000, 1.0]),
        (3, 4): ([0.445622407287714, 1.064177772475912, 0.120614758428183],
                 [0.935822227524088, 0.497940606760015, -0.345865915800969, 0.0, 0.935822227524088,
                  2.867525668568206, 0.0, 0.0, 0.935822227524088],
                 [1.068579021301629, 0.5, -0.068579021301629]),
    }
    c, d, t = table[(nstage, order)]
    return np.array(c), np.array(d), np.array(t)


def stage_source(istage, nstage, d, dt, UDGi, UDG, SH):
    """SH of DIRK stage istage (solvers/dgsolvers.cpp:316-341)."""
    D = lambda i, j: d[i + nstage * j]
    if istage == 0:
        return (D(0, 0) / dt) * UDGi
    if istage == 1:
        return ((D(1, 0) + D(1, 1)) / dt) * UDGi - UDG * D(1, 0) / dt
    if istage == 2:
        return (((D(2, 0) + D(2, 1) + D(2, 2)) / dt) * UDGi
                - (D(2, 0) / D(1, 0)) * ((D(1, 0) + D(1, 1)) * UDGi / dt - SH)
                - UDG * D(2, 1) / dt)
    raise ValueError("DIRK scheme not implemented for more than 3 stages")


def dirk_step(prob, UDG, UH, SH, dt, nstage=3, order=3, **newton):
    """One time step of solveUnsteadyProblem; returns (UDG, UH, SH, per-stage histories)."""
    c, d, t = dirk_coefficients(nstage, order)
    ncu = prob.ncu
    UDGi = np.array(UDG, dtype=float)
    hist = []
    ordering = newton.pop("ordering_cache", {})
    for s in range(nstage):
        fc_u = d[s * nstage + s] / dt
        SH = stage_source(s, nstage, d, dt, UDGi, UDG, SH)
        SH = np.array(SH)
        SH[:, ncu:, :] = 0.0                       # no q source for non-wave problems (:344-349)
        ps = dataclasses.replace(prob, fc_u=fc_u, fc_q=1.0, tdep=1)
        UDG, UH, ok, h = newton_solve(ps, UDG, UH, SH, order=ordering, **newton)
        if not ok:
            raise RuntimeError(f"DIRK stage {s + 1} did not converge")
        hist.append(h)
    return UDG, UH, SH, hist


def quasi_newton_store(blocks, accuracy=1):
    """inv(D) and K of every element from uncondensed blocks (element-major flat,
    column-major); rounded to float and widened back for accuracy 0 (DinvFloat,
    Kfloat), as schurElementVector reads them."""
    n, na = blocks["Ru"].shape
    Dinv = np.empty_like(blocks["D"])
    for e in range(n):
        A, info = O.dense_inverse(blocks["D"][e].reshape(na, na, order="F"))
        if info:
            raise FloatingPointError(f"singular D in element {e}")
        Dinv[e] = A.reshape(-1, order="F")
    K = blocks["K"].copy()
    if accuracy == 0:
        Dinv = Dinv.astype(np.float32).astype(np.float64)
        K = K.astype(np.float32).astype(np.float64)
    return dict(Dinv=Dinv, K=K)


def trace_system_qn(prob, UDG, UH, SH, accuracy=1):
    """assembleLinearSystem with quasiNewton = 1: trace_system's outputs plus the
    stored inv(D) and K of this Jacobian."""
    orc = O.Oracle(prob)
    raw = orc.assemble(0, prob.ne, UDG=UDG, UH=UH, SH=SH)
    stored = quasi_newton_store(raw, accuracy)
    M, Rg, rnorm, blocks = O.trace_system(prob, UDG=UDG, UH=UH, SH=SH)
    return M, Rg, rnorm, blocks, stored


def assemble_rhs(prob, UDG, UH, SH, stored):
    """assembleRHS: element vectors at (UDG, UH, SH) condensed with the stored
    operands, y = inv(D) Ru and FH = Rh - K y, scattered into Rg; rNorm from the
    uncondensed Ru and the assembled uncondensed Rh.  Returns (Rg, rNorm, y)."""
    d = prob.dims
    na, nb, ne = d["na"], d["nb"], prob.ne
    b = O.Oracle(prob).assemble(0, ne, UDG=UDG, UH=UH, SH=SH)
    zeros = np.zeros((ne, nb * nb))
    _, r = O.scatter(prob, zeros, b["Rh"])
    rnorm = float(np.sqrt((b["Ru"] ** 2).sum() + (r ** 2).sum()))
    y = np.empty((ne, na))
    FH = np.empty((ne, nb))
    for e in range(ne):
        Dinv = stored["Dinv"][e].reshape(na, na, order="F")
        K = stored["K"][e].reshape(nb, na, order="F")
        y[e] = Dinv @ b["Ru"][e]
        FH[e] = b["Rh"][e] - K @ y[e]
    _, Rg = O.scatter(prob, zeros, FH)
    return Rg, rnorm, y


def newton_solve_qn(prob, UDG, UH, SH, newton_tol=1e-7, maxiter=10, alpha_min=1e-4, accuracy=1, order=None,
                    **lin):
    """solveNonlinearProblem with quasiNewton = 1 where the time test of
    solvers/dgsolvers.cpp:137-141 always passes: after an iteration the Jacobian
    (Hg, its preconditioner, the stored inv(D), K and inv(D) F) is reused until a
    damped step recomputes it (:171).  history records ``reused`` per iteration."""
    order = {} if order is None else order
    UDG = np.array(UDG, dtype=float)
    UH = np.array(UH, dtype=float)
    rnorm = newton_tol + 1.0
    history = []
    it = 0
    reuse = False
    M = stored = DinvF = None
    while rnorm > newton_tol and it < maxiter:
        it += 1
        reused = reuse
        if not reused:
            M, Rg, old, blocks, stored = trace_system_qn(prob, UDG, UH, SH, accuracy)
            DinvF = blocks["F"]
            upd = blocks
        else:
            Rg, old, y = assemble_rhs(prob, UDG, UH, SH, stored)
            upd = dict(F=DinvF, Ru=y)
        if not np.isfinite(old):
            raise FloatingPointError("residual norm is NaN or Inf")
        x, flag, iters, rev = linear_solve(prob, M, Rg, order=order, **lin)
        rec = dict(old=old, gmres_iters=iters, gmres_flag=flag, alphas=[], reused=reused)
        history.append(rec)
        reuse = True
        if it == 1 and old < newton_tol:
            rnorm = old
            rec["new"] = rnorm
            break
        state = [UDG, UH]

        def apply_step(alpha):
            state[0], state[1] = update_solution(prob, state[0], state[1], SH, upd, x, alpha)

        rec["alphas"], rnorm, _ = line_search(apply_step, lambda: residual_norm(prob, state[0], state[1], SH),
                                              old, newton_tol, alpha_min, quasi_newton=True)
        UDG, UH = state
        if len(rec["alphas"]) > 1:
            reuse = False                      # a damped step recomputes the Jacobian (:171)
        rec["new"] = rnorm
    return UDG, UH, rnorm < newton_tol, history
