"""The parts of bench.py that run without a GPU (CPU)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("na,nb,mflop", [(100, 200, 12.73), (60, 60, 1.022), (72, 36, None)])
def test_condensation_flop_count(na, nb, mflop):
    """(2/3) na^3 + 2 na^2 nb + 2 nb^2 na + 2 na^2 + 2 nb na: LU of D, inv(D)[F Ru]
    and K X, as counted for the roofline (p=3 tets: 12.73 MFLOP per element;
    p=4 triangles 1.022)."""
    f = bench.condense_flops(na, nb)
    direct = 2 / 3 * na ** 3 + 2 * na * na * nb + 2 * nb * nb * na + 2 * na * na + 2 * nb * na
    assert f == pytest.approx(direct, rel=1e-15)
    if mflop is not None:
        assert f / 1e6 == pytest.approx(mflop, abs=5e-3)


def test_roofline_traffic_lines_are_read_from_the_profile():
    for k in ("assemble_kernel+flux_kernel",):
        t = bench.hbm_traffic(k)
        assert t is not None and 1e9 < t < 1e12
    assert bench.hbm_traffic("no_such_kernel") is None


def test_assembly_flop_table_covers_the_bench_workload():
    # (nd, element type, porder) of the tetrahedral p=3 workload the metric is quoted on
    assert bench.ASSEMBLY_FLOPS[(3, 0, 3)] == pytest.approx(9.636e6)
