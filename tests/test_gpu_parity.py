"""HIP element path vs the CPU oracle and the reference's golden vectors (GPU).

Tolerances (relative Frobenius norm, max over elements):
* FP64 path: element matrices and condensed blocks <= 1e-13; the element
  residual Ru and inv(D) Ru <= 1e-10 (cancellation between O(1) volume and face
  terms; the reference's own impl 0 and impl 1 differ by ~1e-12 on it).
  Ru is the difference of volume and face integrals of the size of Rh; at p=1
  it nearly cancels (|Ru| ~ 1e-5 |Rh| on the seeded tets), so a rounding of
  1e-16 |Rh| is 1e-11 of |Ru|.  Residual errors are therefore measured against
  max(|Ru_e|, 1e-4 |Rh_e|) per element, and inv(D) Ru against that floor
  carried through inv(D)'s gain on the element (|inv(D)Ru_e| / |Ru_e|).
* Mixed path (FP32 factorisation + FP64 refinement): <= 1e-10 on AH, FH,
  inv(D)F, inv(D)Ru.
"""
import os

import numpy as np
import pytest

from hfem_amd.problem import make_problem

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RTOL_MAT = 1e-13
RTOL_RES = 5e-12
RTOL_MIXED = 1e-10


def relerr(a, b):
    a = np.asarray(a).reshape(len(b), -1)
    b = np.asarray(b).reshape(len(b), -1)
    return float((np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)).max())


def res_floor(Ru, Rh):
    """Per-element floor 1e-4 |Rh_e| for the residual's relative error."""
    n = len(Ru)
    return 1e-4 * np.linalg.norm(np.asarray(Rh).reshape(n, -1), axis=1)


def relerr_res(a, b, floor):
    a = np.asarray(a).reshape(len(b), -1)
    b = np.asarray(b).reshape(len(b), -1)
    return float((np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), floor)).max())


def say(*a):
    print(*a, flush=True)


CASES = [("tri", 2, 2), ("tri", 2, 4), ("tet", 1, 2), ("tet", 1, 3), ("hex", 3, 2), ("tet", 2, 3), ("tet", 1, 4),
         ("tri", 2, 1), ("tri", 2, 3), ("tet", 1, 1), ("hex", 3, 1), ("tri", 2, 5)]


@pytest.mark.parametrize("kind,nc,p", CASES)
def test_assembly_and_condensation_vs_oracle(oracle_mod, kind, nc, p):
    from hfem_amd.lib import HDGContext, FP64, MIXED
    O = oracle_mod
    pr = make_problem(kind, nc, p, seed=3)
    d = pr.dims
    say(f"case {kind}{nc} p{p}: ne={pr.ne} na={d['na']} nb={d['nb']}")
    ref = O.Oracle(pr).assemble()
    ctx = HDGContext(pr)
    ctx.assemble_element_matrix_vector()
    g = ctx.read_elements()
    for k in ("D", "F", "K", "H", "Rh"):
        e = relerr(g[k], ref[k])
        say(f"  assemble {k}: {e:.2e}")
        assert e < RTOL_MAT, k
    fl_ru = res_floor(ref["Ru"], ref["Rh"])
    e = relerr_res(g["Ru"], ref["Ru"], fl_ru)
    say(f"  assemble Ru: {e:.2e} (relative {relerr(g['Ru'], ref['Ru']):.2e})")
    assert e < RTOL_RES, "Ru"
    b = {k: v.copy() for k, v in ref.items()}
    O.schur(0, d["npv"], d["ncu"], b)
    nr = np.linalg.norm(ref["Ru"].reshape(pr.ne, -1), axis=1)
    fl_y = fl_ru * np.linalg.norm(b["Ru"].reshape(pr.ne, -1), axis=1) / np.maximum(nr, 1e-300)
    floors = {"DinvRu": fl_y, "FH": 0.0}
    ctx.schur_element_matrix_vector(precision=FP64)
    c = ctx.read_condensed()
    assert np.all(c["info"] == 0)
    for name, got, want, tol in [("DinvF", c["DinvF"], b["F"], RTOL_MAT), ("AH", c["AH"], b["H"], RTOL_MAT),
                                 ("DinvRu", c["DinvRu"], b["Ru"], RTOL_RES), ("FH", c["FH"], b["Rh"], RTOL_MAT)]:
        e = relerr_res(got, want, floors[name]) if name in floors else relerr(got, want)
        say(f"  fp64 condense {name}: {e:.2e}")
        assert e < tol, name
    # mixed precision on freshly assembled blocks
    ctx.assemble_element_matrix_vector()
    ctx.schur_element_matrix_vector(precision=MIXED)
    c = ctx.read_condensed()
    for name, got, want in [("DinvF", c["DinvF"], b["F"]), ("AH", c["AH"], b["H"]), ("DinvRu", c["DinvRu"], b["Ru"]),
                            ("FH", c["FH"], b["Rh"])]:
        e = relerr_res(got, want, floors[name]) if name in floors else relerr(got, want)
        say(f"  mixed condense {name}: {e:.2e}")
        assert e < RTOL_MIXED, name
    ctx.close()


@pytest.mark.parametrize("name,kind,nc,p", [("tri2_p2", "tri", 2, 2), ("tri2_p4", "tri", 2, 4),
                                            ("tet1_p2", "tet", 1, 2), ("tet1_p3", "tet", 1, 3),
                                            ("hex3_p2", "hex", 3, 2), ("tet1_p4", "tet", 1, 4),
                                            ("tri2_p5", "tri", 2, 5)])
def test_against_reference_golden(name, kind, nc, p):
    """Directly against the reference element path's outputs, both of its Schur
    implementations where the fixture holds them (0: LU with partial pivoting;
    1, its default, preprocessing/digasopre.m:466-468: block LU, valid on
    elements without interior nodes)."""
    from hfem_amd.lib import HDGContext, FP64
    gold = np.load(os.path.join(GOLD, f"elem_{name}.npz"))
    pr = make_problem(kind, nc, p, seed=1)
    say(f"golden {name}")
    keep = gold["keep"]
    ctx = HDGContext(pr)
    ctx.assemble_element_matrix_vector()
    g = ctx.read_elements()
    for k in ("D", "F", "K", "H", "Rh"):
        e = relerr(g[k][keep], gold[k])
        say(f"  {k}: {e:.2e}")
        assert e < RTOL_MAT, k
    fl_ru = res_floor(gold["Ru"], gold["Rh"])
    e = relerr_res(g["Ru"][keep], gold["Ru"], fl_ru)
    say(f"  Ru: {e:.2e}")
    assert e < RTOL_RES
    ctx.schur_element_matrix_vector(precision=FP64)
    c = ctx.read_condensed()
    nr = np.linalg.norm(gold["Ru"], axis=1)
    for impl in (0, 1):
        if f"AH_impl{impl}" not in gold:
            continue
        fl_y = fl_ru * np.linalg.norm(gold[f"DinvRu_impl{impl}"], axis=1) / np.maximum(nr, 1e-300)
        errs = {"AH": relerr(c["AH"][keep], gold[f"AH_impl{impl}"]),
                "FH": relerr(c["FH"][keep], gold[f"FH_impl{impl}"]),
                "DinvRu": relerr_res(c["DinvRu"][keep], gold[f"DinvRu_impl{impl}"], fl_y)}
        if f"DinvF_impl{impl}" in gold:
            errs["DinvF"] = relerr(c["DinvF"][keep], gold[f"DinvF_impl{impl}"])
        say(f"  impl {impl}: " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()))
        for k, v in errs.items():
            assert v < (RTOL_RES if k == "DinvRu" else RTOL_MAT), (impl, k)
    ctx.close()


@pytest.mark.parametrize("kind,nc,p", [("tri", 3, 2), ("tet", 1, 3)])
def test_linear_system_and_matvec(oracle_mod, kind, nc, p):
    from hfem_amd.lib import HDGContext, FP64
    O = oracle_mod
    pr = make_problem(kind, nc, p, seed=4)
    d = pr.dims
    say(f"system {kind}{nc} p{p}")
    ref = O.Oracle(pr).assemble()
    rnorm_ref = np.sqrt((ref["Ru"] ** 2).sum() + (O.scatter(pr, np.zeros((pr.ne, d["nb"] ** 2)), ref["Rh"])[1] ** 2).sum())
    O.schur(0, d["npv"], d["ncu"], ref)
    Hg_ref, Rg_ref = O.scatter(pr, ref["H"], ref["Rh"])
    ctx = HDGContext(pr)
    rnorm = ctx.assemble_linear_system(FP64)
    assert abs(rnorm - rnorm_ref) < 1e-12 * rnorm_ref
    Hg, Rg = ctx.read_system()
    assert np.linalg.norm(Hg - Hg_ref) < RTOL_MAT * np.linalg.norm(Hg_ref)
    assert np.linalg.norm(Rg - Rg_ref) < RTOL_RES * np.linalg.norm(Rg_ref)
    x = np.random.default_rng(0).standard_normal(len(Rg))
    y = ctx.bsr_matvec(x)
    yref = O.bsr_matvec(pr, Hg_ref, x)
    assert np.linalg.norm(y - yref) < 1e-13 * np.linalg.norm(yref)
    ctx.close()


def test_tuned_condensation(oracle_mod):
    """Every launch candidate the tuner can pick gives the same blocks (1e-13
    FP64, 1e-10 mixed; a different summation order is allowed), the tuner's pick
    too, and tuning changes no element block."""
    from hfem_amd.lib import HDGContext, FP64, MIXED
    O = oracle_mod
    pr = make_problem("tet", 2, 3, seed=6)
    d = pr.dims
    say(f"tuned condensation tet2 p3: ne={pr.ne}")
    ref = O.Oracle(pr).assemble()
    O.schur(0, d["npv"], d["ncu"], ref)
    ctx = HDGContext(pr)
    ctx.assemble_element_matrix_vector()
    before = ctx.read_elements()
    choice, ms = ctx.tune_condensation()
    say(f"  choice {choice} ms {ms}")
    after = ctx.read_elements()
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    assert len(choice) == 7 and all(m > 0 for m in ms)

    def check(tag):
        for prec, tol in ((FP64, RTOL_MAT), (MIXED, RTOL_MIXED)):
            ctx.assemble_element_matrix_vector()
            ctx.schur_element_matrix_vector(precision=prec)
            c = ctx.read_condensed()
            for name, want in (("DinvF", ref["F"]), ("AH", ref["H"]), ("FH", ref["Rh"])):
                e = relerr(c[name], want)
                say(f"  {tag} precision {prec} {name}: {e:.2e}")
                assert e < tol, (tag, name)
            if prec == MIXED:
                sweeps, left, fb = ctx.condense_report()
                say(f"  {tag} mixed sweeps {sweeps} unconverged {left} fallback {fb}")
                assert left == 0 and fb == 0 and 1 <= sweeps <= 4

    check("tuned")
    # every hand-written variant of every product, and every inverse launch shape
    from hfem_amd.lib import load
    for v in range(load().hfem_condense_variants()):
        pick = [v % 3, v % 3] + [v] * 5
        if ctx.set_condensation_choice(pick, strict=False):
            check(f"variant {v}")
    # the fused per-element solves (FP64 X and the mixed X0 + sweeps)
    if ctx.set_condensation_choice([0, 0, 64, 0, 64, 0, 0], strict=False):
        check("fused solve")
    # rocBLAS for the FP64 products that have it
    if ctx.set_condensation_choice([0, 0, -1, -1, 0, -1, 0], strict=False):
        check("rocBLAS")
    ctx.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_condensation_by_parts(precision):
    """Condensing [0, k) and then [k, ne) gives the blocks of one condensation of
    [0, ne): the element offsets of the batched steps (inverse, GEMMs,
    refinement residual) are right.  rocBLAS may pick other kernels for the
    smaller batches, so the comparison allows rounding (1e-13)."""
    from hfem_amd.lib import HDGContext, FP64, MIXED
    prec = FP64 if precision == "fp64" else MIXED
    pr = make_problem("tet", 2, 3, seed=8)
    say(f"condensation by parts tet2 p3 {precision}: ne={pr.ne}")
    ctx = HDGContext(pr)
    ctx.assemble_element_matrix_vector()
    ctx.schur_element_matrix_vector(precision=prec)
    whole = ctx.read_condensed()
    k = pr.ne // 3
    ctx.assemble_element_matrix_vector()
    ctx.schur_element_matrix_vector(0, k, prec)
    ctx.schur_element_matrix_vector(k, pr.ne - k, prec)
    parts = ctx.read_condensed()
    for name in ("DinvF", "DinvRu", "AH", "FH"):
        e = relerr(parts[name], whole[name])
        say(f"  {name}: {e:.2e}")
        assert e < RTOL_MAT, name
    ctx.close()


@pytest.mark.parametrize("kind,nc,p", [("tet", 1, 3), ("tri", 2, 2), ("hex", 3, 1)])
def test_assemble_vector_matches_full_assembly(kind, nc, p):
    """assembleElementVector: the residual-only assembly writes the same Ru and Rh
    as the full assembly (bit for bit: the same residual code reads the same flux
    values) and leaves the element matrices alone."""
    from hfem_amd.lib import HDGContext
    pr = make_problem(kind, nc, p, seed=2)
    ctx = HDGContext(pr)
    ctx.assemble_element_matrix_vector()
    full = ctx.read_elements()
    pr2 = make_problem(kind, nc, p, seed=9)          # another state: Ru, Rh change, D.. must not
    ctx.set_solution(pr2.UDG, pr2.UH, pr2.SH)
    ctx.assemble_element_vector()
    vec = ctx.read_elements()
    for k in ("D", "F", "K", "H"):
        assert np.array_equal(vec[k], full[k]), k
    ctx.assemble_element_matrix_vector()
    full2 = ctx.read_elements()
    for k in ("Ru", "Rh"):
        assert np.array_equal(vec[k], full2[k]), k
        assert not np.array_equal(vec[k], full[k]), k
    ctx.close()


def test_errors_are_reported():
    from hfem_amd.lib import HDGContext, HfemError
    pr = make_problem("tri", 1, 1, seed=0)
    ctx = HDGContext(pr)
    with pytest.raises(HfemError, match="outside"):
        ctx.assemble_element_matrix_vector(0, pr.ne + 1)
    with pytest.raises(HfemError, match="precision"):
        ctx.schur_element_matrix_vector(0, 1, precision=7)
    with pytest.raises(HfemError, match="nsample"):
        ctx.tune_condensation(-1)
    ctx.close()
