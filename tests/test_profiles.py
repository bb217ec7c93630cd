"""The committed rocprofv3 summaries hold the fields bench.py reads, for the
benchmark's own command (bench.py --gpus 1 --steps 20 --warmup 5)."""
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = os.path.join(ROOT, "profiles")
CMD = "bench.py --gpus 1 --steps 20 --warmup 5"


def test_traffic_summary_fields():
    d = json.load(open(os.path.join(P, "stage2d_euler2d_1024_fp64_traffic.json")))
    w = d["windows"]["bench:replay"]
    assert w["launches"] == 60 and w["bytes_per_launch"] > 0
    # below 1 on the benchmark's state: the deep-constant tiles of a step's
    # second and third stages read one cell of q per tile
    assert 0.5 < w["traffic_over_alg"] < 2.0
    assert 0.9 < d["windows"]["bench:smooth"]["traffic_over_alg"] < 2.0
    assert d["alg_bytes_per_launch"] > 0
    assert any("--pmc FETCH_SIZE" in c and CMD in c for c in d["commands"])
    assert set(d["calibration"]) == {"FETCH_SIZE", "WRITE_SIZE"}


def test_valu_summary_fields():
    d = json.load(open(os.path.join(P, "stage2d_euler2d_1024_fp64_valu.json")))
    assert any("--pmc SQ_INSTS_VALU" in c and CMD in c for c in d["commands"])
    assert "stage2d_kernel<4, 2, double," in d["kernel"]
    r = json.load(open(os.path.join(P, "stage2d_res_euler2d_1024_fp64_counters.json")))
    assert "stage2d_res_kernel<4, 2, double," in r["kernel"] and r["windows"]["bench:timed"]["launches"] == 60
    for w in ("bench:replay", "bench:smooth"):
        e = d["windows"][w]
        assert e["fp64_flop_per_launch"] > 0 and e["valu_lane_instr_per_cell"] > 0
        assert e["counters"]["SQ_INSTS_VALU"] > 0


def test_kernel_stats_present():
    for f in ("stage2d_euler2d_1024_fp64_kernel_stats.txt", "batch_dambreak4096_kernel_stats.txt"):
        txt = open(os.path.join(P, f)).read()
        assert "rocprofv3 --kernel-trace" in txt and "avg_us=" in txt
    txt = open(os.path.join(P, "stage2d_euler2d_1024_fp64_kernel_stats.txt")).read()
    assert CMD in txt and "# range bench:timed" in txt
