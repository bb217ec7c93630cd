"""The committed rocprofv3 summaries hold the fields bench.py reads."""
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = os.path.join(ROOT, "profiles")


def test_traffic_summary_fields():
    d = json.load(open(os.path.join(P, "stage2d_euler2d_1024_fp64_traffic.json")))
    assert d["bytes_per_launch"] > 0 and d["alg_bytes_per_launch"] > 0
    assert 0.9 < d["traffic_over_alg"] < 2.0
    assert all("rocprofv3 --pmc" in c for c in d["commands"])


def test_valu_summary_fields():
    d = json.load(open(os.path.join(P, "stage2d_euler2d_1024_fp64_valu.json")))
    assert "rocprofv3 --pmc" in d["command"]
    ks = d["kernels"]
    assert any(k.startswith("stage2d_kernel<4, 2, double,") for k in ks)
    for k in ks.values():
        assert k["fp64_flop_per_launch"] > 0 and k["valu_lane_instr_per_cell"] > 0
        assert k["counters"]["SQ_INSTS_VALU"] > 0


def test_kernel_stats_present():
    for f in ("stage2d_euler2d_1024_fp64_kernel_stats.txt", "batch_dambreak4096_kernel_stats.txt"):
        txt = open(os.path.join(P, f)).read()
        assert "rocprofv3 --kernel-trace --stats" in txt and "avg_us=" in txt
