"""bench.py's contract and measurement formulas, on the CPU (no device call):

* the per-unit figures of the measurement plan -- the reference's dense-B flop count
  per element (188,430 for the 10-node tet with 11 points, 455.4 M for the 125-node hex
  with 216) and the compulsory bytes of one K+T assembly (10.5 MB for the 3072-tet p=2
  cube, 13.90 GB for the 24^3 hex p=4 cube, with their CSR sizes) -- come out of the
  functions the bench line is computed with;
* the measured traffic the line carries is read from the committed PMC summary;
* ``--gpus N`` with N > 1 exits with status 2 and says why, before touching a device.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_dense_b_flops_per_element():
    assert bench.dense_b_flops_per_element(10, 11) == 188430
    assert bench.dense_b_flops_per_element(125, 216) == pytest.approx(455.4e6, rel=1e-3)


def test_compulsory_bytes():
    # tet p=2 cube: 3072 elements, 4913 nodes, nnz 1,253,961
    assert bench.compulsory_bytes(3072, 10, 4913, 1253961) == pytest.approx(10.5e6, rel=0.01)
    # hex p=4 cube 24^3: 13824 elements, 912,673 nodes, nnz 1,728,900,297
    assert bench.compulsory_bytes(13824, 125, 912673, 1728900297) == 13903826832


def test_measured_traffic_from_committed_profile():
    t = bench.measured_traffic(bench.PMC_SUMMARY, "stiffness")
    assert t is not None and 2e10 < t < 5e10
    assert bench.measured_traffic("no_such_summary.md", "stiffness") is None


def test_multi_gpu_request_exits():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
