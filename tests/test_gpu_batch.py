"""Ragged element batches (hfem_batch_*, GPU): the element path of contexts of
different families -- tets of orders 1-4, hexes of orders 1-2 -- in single
persistent launches.

* Every context receives bit for bit what its own hfem_assemble / hfem_condense
  write (the batch kernels run the same per-element bodies; the inverse's
  waves per element change which wave holds a tile, not the arithmetic), with
  the buckets' element counts chosen so that the inverse's groups of 2 and 4
  elements per workgroup end on partial rounds.
* The mixed condensation through the batch equals each context's own mixed
  condensation bit for bit (and the FP64 one within the mixed bound).
* Directly against the oracle on a batch of the per-context parity test's
  small tets of orders 1-3 (its seeds and tolerances).
* After the batch tuner has picked the waves per element, the bucket order and
  the launch forms (one persistent launch, or one launch per bucket), and with
  every launch form forced, the results are unchanged.
"""
import numpy as np
import pytest

from hfem_amd.problem import make_problem

pytestmark = pytest.mark.gpu

FAMILIES = [("tet", 3, 1), ("tet", 2, 2), ("tet", 1, 3), ("tet", 1, 4), ("hex", 3, 1), ("hex", 3, 2)]
ELEM = ("D", "F", "K", "H", "Ru", "Rh")
COND = ("DinvRu", "DinvF", "AH", "FH", "info")


def say(*a):
    print(*a, flush=True)


def per_context(ctxs):
    from hfem_amd.lib import FP64
    out = []
    for c in ctxs:
        c.assemble_element_matrix_vector()
        e = c.read_elements()
        c.schur_element_matrix_vector(precision=FP64)
        out.append((e, c.read_condensed()))
    return out


def batched(batch, ctxs):
    batch.assemble_element_matrix_vector()
    el = [c.read_elements() for c in ctxs]
    batch.schur_element_matrix_vector()
    return [(e, c.read_condensed()) for e, c in zip(el, ctxs)]


def assert_same(a, b, names):
    for (ea, ca), (eb, cb), nm in zip(a, b, names):
        for k in ELEM:
            assert np.array_equal(ea[k], eb[k]), (nm, k)
        for k in COND:
            assert np.array_equal(ca[k], cb[k]), (nm, k)


def test_batch_equals_per_context_calls():
    from hfem_amd.lib import HDGBatch, HDGContext
    probs = [make_problem(k, n, p, seed=5) for k, n, p in FAMILIES]
    names = [f"{k}{n} p{p} ne={pr.ne}" for (k, n, p), pr in zip(FAMILIES, probs)]
    say("batch:", ", ".join(names))
    ctxs = [HDGContext(pr) for pr in probs]
    want = per_context(ctxs)
    batch = HDGBatch(ctxs)
    got = batched(batch, ctxs)
    assert_same(got, want, names)
    say("  default shapes: bit-identical")
    waves, tasm, tinv = batch.tune()
    say(f"  tuned waves per element {waves}, assembly ms {np.round(tasm, 3)}, inverse ms {np.round(tinv, 3)}")
    for (k, n, p), w, pr in zip(FAMILIES, waves, probs):
        nt = (pr.dims["na"] + 15) // 16
        assert w in ((8,) if nt > 8 else (4, 2) if nt <= 4 else (8, 4, 2))
    # the tuner also tuned each context's products: per-context calls again
    want = per_context(ctxs)
    got = batched(batch, ctxs)
    assert_same(got, want, names)
    say("  tuned shapes: bit-identical")
    nts = [(pr.dims["na"] + 15) // 16 for pr in probs]
    for rule in (lambda nt: 8 if nt > 8 else 2, lambda nt: 8 if nt > 8 else 4, lambda nt: 8 if nt > 4 else 4):
        w = [rule(nt) for nt in nts]
        batch.set_inverse_waves(w)
        got = batched(batch, ctxs)
        assert_same(got, want, names)
        say(f"  inverse waves {w}: bit-identical")
    (ma, mi), tm = batch.launch()
    say(f"  tuned launch forms: assembly {ma}, inverse {mi} ({tm})")
    for modes in ((1, 1), (0, 1), (1, 0), (0, 0)):
        batch.set_launch(*modes)
        got = batched(batch, ctxs)
        assert_same(got, want, names)
        say(f"  launch forms {modes}: bit-identical")
    batch.close()
    for c in ctxs:
        c.close()


def test_batch_against_oracle(oracle_mod):
    from hfem_amd.lib import HDGBatch, HDGContext
    from test_gpu_parity import RTOL_MAT, RTOL_RES, relerr, relerr_res, res_floor
    O = oracle_mod
    # the per-context parity cases (same seeds, same tolerances), run as one batch
    fams = [("tet", 1, 1), ("tet", 1, 2), ("tet", 1, 3)]
    probs = [make_problem(k, n, p, seed=3) for k, n, p in fams]
    ctxs = [HDGContext(pr) for pr in probs]
    batch = HDGBatch(ctxs)
    got = batched(batch, ctxs)
    for (k, n, p), pr, (g, c) in zip(fams, probs, got):
        say(f"oracle {k}{n} p{p}: ne={pr.ne}")
        ref = O.Oracle(pr).assemble()
        for key in ("D", "F", "K", "H", "Rh"):
            assert relerr(g[key], ref[key]) < RTOL_MAT, key
        fl_ru = res_floor(ref["Ru"], ref["Rh"])
        assert relerr_res(g["Ru"], ref["Ru"], fl_ru) < RTOL_RES
        d = pr.dims
        b = {kk: v.copy() for kk, v in ref.items()}
        O.schur(0, d["npv"], d["ncu"], b)
        nr = np.linalg.norm(ref["Ru"].reshape(pr.ne, -1), axis=1)
        fl_y = fl_ru * np.linalg.norm(b["Ru"].reshape(pr.ne, -1), axis=1) / np.maximum(nr, 1e-300)
        assert np.all(c["info"] == 0)
        assert relerr(c["DinvF"], b["F"]) < RTOL_MAT
        assert relerr(c["AH"], b["H"]) < RTOL_MAT
        assert relerr(c["FH"], b["Rh"]) < RTOL_MAT
        assert relerr_res(c["DinvRu"], b["Ru"], fl_y) < RTOL_RES
    batch.close()
    for c in ctxs:
        c.close()


def test_batch_mixed_equals_per_context_mixed():
    """The mixed condensation through the batch (each bucket's FP32-MFMA inverse
    and FP64 refinement on its own stream) equals each context's own mixed
    condensation bit for bit, and the FP64 condensation within the mixed bound."""
    from hfem_amd.lib import HDGBatch, HDGContext, FP64, MIXED
    from test_gpu_parity import RTOL_MIXED, relerr
    probs = [make_problem("tet", 2, 1, seed=2), make_problem("tet", 2, 2, seed=2), make_problem("tet", 1, 3, seed=2),
             make_problem("tet", 1, 4, seed=2)]
    ctxs = [HDGContext(pr) for pr in probs]
    want, fp64 = [], []
    for c in ctxs:
        c.assemble_element_matrix_vector()
        c.schur_element_matrix_vector(precision=FP64)
        fp64.append(c.read_condensed())
        c.assemble_element_matrix_vector()
        c.schur_element_matrix_vector(precision=MIXED)
        want.append(c.read_condensed())
    batch = HDGBatch(ctxs)
    batch.assemble_element_matrix_vector()
    batch.schur_element_matrix_vector(MIXED)
    for c, w, f, pr in zip(ctxs, want, fp64, probs):
        got = c.read_condensed()
        for name in ("DinvRu", "DinvF", "AH", "FH", "info"):
            assert np.array_equal(got[name], w[name]), name
        e = max(relerr(got[k], f[k]) for k in ("DinvF", "AH", "FH"))
        say(f"  mixed batch, ne={pr.ne} na={pr.dims['na']}: bit-identical to the context's own; vs FP64 {e:.1e}")
        assert e < RTOL_MIXED
    batch.close()
    for c in ctxs:
        c.close()


def test_batch_with_empty_and_partial_buckets():
    """Buckets of 0 elements and buckets that take only the first elements of
    their context, in every launch form: the taken elements get the per-context
    results, the others keep their blocks."""
    from hfem_amd.lib import HDGBatch, HDGContext, FP64
    probs = [make_problem("tet", 2, 1, seed=3), make_problem("tet", 1, 2, seed=3), make_problem("tet", 1, 3, seed=3)]
    ctxs = [HDGContext(pr) for pr in probs]
    take = [0, 5, 3]
    want = []
    for c, k in zip(ctxs, take):
        c.assemble_element_matrix_vector()
        c.schur_element_matrix_vector(precision=FP64)
        want.append(c.read_condensed())
    batch = HDGBatch(ctxs, take)
    for modes in ((0, 0), (1, 1)):
        batch.set_launch(*modes)
        # the untaken elements keep the per-context results; the taken ones are redone
        batch.assemble_element_matrix_vector()
        batch.schur_element_matrix_vector()
        for c, k, w in zip(ctxs, take, want):
            got = c.read_condensed()
            for name in ("DinvRu", "DinvF", "AH", "FH", "info"):
                assert np.array_equal(got[name], w[name]), (modes, k, name)
        say(f"  launch forms {modes}, elements taken {take}: bit-identical")
    waves, tasm, tinv = batch.tune()
    say(f"  tuned waves {waves}")
    batch.close()
    for c in ctxs:
        c.close()


def test_batch_rejects_what_it_cannot_run():
    from hfem_amd.lib import HDGBatch, HDGContext, HfemError
    a = HDGContext(make_problem("tet", 1, 2, seed=1))
    b = HDGContext(make_problem("tet", 1, 2, seed=2))
    t = HDGContext(make_problem("tri", 2, 2, seed=1))
    with pytest.raises(HfemError, match="same element family"):
        HDGBatch([a, b])
    with pytest.raises(HfemError, match="not a batch family"):
        HDGBatch([a, t])
    with pytest.raises(HfemError, match="given twice"):
        HDGBatch([a, a])
    with pytest.raises(HfemError, match="outside"):
        HDGBatch([])
    bt = HDGBatch([a])
    with pytest.raises(HfemError, match="waves per element not available"):
        bt.set_inverse_waves([8])
    with pytest.raises(HfemError, match="modes must be 0 or 1"):
        bt.set_launch(2, 0)
    with pytest.raises(HfemError, match="outside"):
        HDGBatch([a], [a.prob.ne + 1])
    bt.close()
    for c in (a, b, t):
        c.close()
