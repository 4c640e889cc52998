"""The committed counter profiles carry the figures bench.py puts into its roofline lines
(HBM traffic and VALU instructions per launch) for the kernels the bench workloads launch."""
import json
import os

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("workload", ["sand", "fluid"])
def test_profiles_hold_the_bench_kernels(workload):
    counters = bench.counters_from_profiles(workload)
    assert counters, f"profiles/counters_{workload}.json missing or empty"
    for kernel, name in bench.PROFILE_KERNELS.items():
        assert name in counters, f"{name} not in profiles/counters_{workload}.json"
        fig = counters[name]
        assert fig["traffic_bytes"] > 0 and fig["valu_insts"] > 0
        assert fig["traffic_bytes"] == pytest.approx(fig["read_bytes"] + fig["write_bytes"])


@pytest.mark.parametrize("workload", ["sand", "fluid"])
def test_profiles_name_their_commit_and_commands(workload):
    with open(os.path.join(ROOT, "profiles", f"counters_{workload}.json")) as f:
        d = json.load(f)
    assert len(d["commit"]) >= 7 and d["commands"]
    assert all(c.startswith("rocprofv3 ") for c in d["commands"])
