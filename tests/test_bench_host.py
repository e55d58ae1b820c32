"""Host-side pieces of bench.py (no GPU): algorithmic byte counts and the
HBM traffic figures read from the committed rocprofv3 summaries."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_bytes_per_eval_matches_shell_sizes():
    # both sites' neighbour rows (int32 index + uint8 species) and the two species
    assert bench.bytes_per_eval(26) == 262  # BCC 8/6/12
    assert bench.bytes_per_eval(42) == 422  # FCC 12/6/24


def test_relax_bytes():
    fwd, bwd = bench.relax_bytes(108000, 42, 5, 4)
    assert fwd == 108000 * 42 * 4 + 108000 * 5 * 4
    assert bwd == fwd + 108000 * 5 * 4


def test_profiled_traffic_is_committed_for_measured_workloads():
    for workload, kernel, key in [("plain-bcc16", "k_local_mc", "traffic_bytes_per_launch"),
                                  ("batched-64", "k_local_mc_batch", "traffic_bytes_per_launch"),
                                  ("relax-fcc30", "k_relax", "traffic_bytes_per_step")]:
        t = bench.profiled_traffic(workload, kernel, key)
        assert isinstance(t, int) and t > 0, workload
    assert bench.profiled_traffic("plain-fcc10") is None


def test_guided_launch_bytes():
    # one round of 512 proposals with 8 candidates each and two accepts, FCC 12/6/24
    st = {"rounds": 1, "accepts": 2}
    assert bench.guided_launch_bytes(st, 512, 42, 8) == 512 * (422 + 8 * 18) + 2 * 84 * 16
