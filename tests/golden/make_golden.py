"""Writes the golden element fixtures from the reference's own element path.

Runs oracle/_ref/libhfemref.so (the reference's assembly/elementassembly.cpp and
assembly/elementschur.cpp compiled from their sources by oracle/ref.mk, with MKL
BLAS/LAPACK) on small seeded cases and stores, for a few elements per case, the
inputs and the element blocks before (Ru, D, F, K, H, Rh) and after
condensation (DinvRu, DinvF, AH, FH) for both schurImplementation 0 and 1.
Run from the repository root in the build container:

    make -f oracle/ref.mk && python tests/golden/make_golden.py [case ...]
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from hfem_amd.problem import make_problem  # noqa: E402
import oracle as O  # noqa: E402

# name -> (kind, ncell, porder, elements to keep, impl-1 valid, keep DinvF)
# schurImplementation 1 inverts the npv x npv diagonal blocks of D; the density
# block is singular (zero rows) when the element has interior nodes (triangles
# p>=3, tets p>=4, hexes p>=2), so impl-1 outputs are stored only where valid.
CASES = {
    "tri2_p2": ("tri", 2, 2, list(range(8)), True, True),
    "tri2_p4": ("tri", 2, 4, [0, 3], False, True),
    "tet1_p2": ("tet", 1, 2, [0, 4], True, True),
    "tet1_p3": ("tet", 1, 3, [0], True, True),
    "hex3_p2": ("hex", 3, 2, [13], False, False),
    "tet1_p4": ("tet", 1, 4, [3], False, True),
    "tri2_p5": ("tri", 2, 5, [0, 5], False, True),
}


def main(names=None):
    out_dir = os.path.dirname(os.path.abspath(__file__))
    for name, (kind, ncell, porder, keep, impl1, dinvf) in CASES.items():
        if names and name not in names:
            continue
        pr = make_problem(kind, ncell, porder, seed=1)
        ref = O.Reference(pr)
        res = {}
        for impl in ((0, 1) if impl1 else (0,)):
            r = ref.element(impl=impl)
            for k, v in r.items():
                if impl == 1 and k in ("Ru", "D", "F", "K", "H", "Rh"):
                    continue
                if k == "DinvF" and not dinvf:
                    continue
                res[f"{k}_impl{impl}" if k in ("DinvRu", "DinvF", "AH", "FH") else k] = v[keep]
        np.savez_compressed(os.path.join(out_dir, f"elem_{name}.npz"), keep=np.array(keep),
                            UDG=pr.UDG, UH=pr.UH, dgnodes=pr.dgnodes_ref(), elcon=pr.mesh.elcon, bf=pr.mesh.bf,
                            param=pr.param, ui=pr.ui, **res)
        print(name, pr.dims["na"], pr.dims["nb"], "elements", keep)


if __name__ == "__main__":
    main(sys.argv[1:])
