"""Builds the HIP library hfem_amd/libhfem_amd.so for gfx950 with hipcc.

Run ``python -m hfem_amd.build`` (or ``__graft_entry__.build()``).  The library is
built in-tree so that it travels with the repository snapshot.  Each source is
compiled to its own object in parallel (every kernel is launched from the file
that defines it, so no relocatable device code is needed), then linked.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhfem_amd.so")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = ["api.hip", "assemble.hip", "condense.hip", "bgemm.hip", "xsolve.hip", "trace.hip", "solver.hip", "solve.hip", "noalloc.hip", "batch.hip"]

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17",
    # no packed-FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
# rocBLAS for the plain batched GEMMs kept as a measured comparison
LINK_FLAGS = ["-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed: " + " ".join(cmd[-3:]))


def build(verbose=False, jobs=None):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in srcs]
    cmds = [[hipcc()] + HIPCC_FLAGS + ["-c", "-o", o, os.path.join(CSRC, s)] for s, o in zip(srcs, objs)]
    jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + LINK_FLAGS, verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
