"""CPU: the rocprofv3 counter summaries behind the benchmark's roofline
(pm355.profiling) on synthetic counter-collection CSVs."""
import csv
import json
import os

from pm355 import profiling

FIELDS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _write(d, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "pmc_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, FIELDS)
        w.writeheader()
        for i, (kern, name, val) in enumerate(rows):
            w.writerow(dict(Correlation_Id=i, Dispatch_Id=i // 10, Agent_Id=1, Kernel_Name=kern, Counter_Name=name,
                            Counter_Value=val))


def test_summarize_scales_and_rates(tmp_path):
    k = "(anonymous namespace)::k_knn_stream512(KnnParams)"
    # two dispatches of the kernel, one of another kernel
    _write(tmp_path / "a", [(k, "TCC_READ_SECTORS_sum", 100), (k, "TCC_HIT_sum", 30), (k, "TCC_MISS_sum", 10)]
           + [("other(int)", "TCC_READ_SECTORS_sum", 1e9)] * 7
           + [(k, "TCC_READ_SECTORS_sum", 300), (k, "TCC_HIT_sum", 90), (k, "TCC_MISS_sum", 10)])
    _write(tmp_path / "b", [(k, "FETCH_SIZE", 1000), (k, "WRITE_SIZE", 10)])
    # the copy probe: 1 GiB read shows as half a GiB of FETCH_SIZE
    _write(tmp_path / "c", [("__amd_rocclr_copyBuffer", "FETCH_SIZE", 524288),
                            ("__amd_rocclr_copyBuffer", "WRITE_SIZE", 1048576)])
    out = tmp_path / "s.json"
    s = profiling.summarize("k_knn_stream512", [str(tmp_path / "a"), str(tmp_path / "b")], [str(tmp_path / "c")],
                            {"variant": "v", "cell_over_r": 0.5}, str(out))
    assert s["l2_read_bytes_per_dispatch"] == 32 * 200  # median of 100 and 300
    assert abs(s["l2_hit_rate"] - 60 / 70) < 1e-12  # medians 60 hits, 10 misses
    assert s["hbm_read_bytes_per_dispatch"] == 1000 * 1024 * 2
    assert s["hbm_write_bytes_per_dispatch"] == 10 * 1024
    assert json.load(open(out))["source_hash"] == profiling.source_hash()
    got = profiling.load(str(out), variant="v", cell_over_r=0.5)
    assert got is not None and got["stale"] is False
    assert profiling.load(str(out), variant="w") is None
    d = json.load(open(out))
    d["source_hash"] = "0" * 16
    json.dump(d, open(out, "w"))
    assert profiling.load(str(out), variant="v")["stale"] is True
    # a summary of the gather kernel is tied to the gather's own sources only
    s2 = profiling.summarize("k_knn_stream512", [str(tmp_path / "a")], [], {"sources": "gather"},
                             str(tmp_path / "g.json"))
    assert s2["source_files"] == profiling.KERNEL_SOURCES["gather"]
    assert s2["source_hash"] == profiling.source_hash(profiling.KERNEL_SOURCES["gather"])
    assert profiling.load(str(tmp_path / "g.json"))["stale"] is False


def test_valu_summary(tmp_path):
    _write(tmp_path / "v", [("k_trace_pool<false>", "SQ_INSTS_VALU_ADD_F32", 10),
                            ("k_trace_pool<false>", "SQ_INSTS_VALU_FMA_F32", 5),
                            ("k_trace_emit", "SQ_INSTS_VALU_MUL_F32", 1),
                            ("k_trace_emit", "SQ_INSTS_VALU_TRANS_F32", 2)])
    s = profiling.summarize_valu([str(tmp_path / "v")], 2, {}, str(tmp_path / "v.json"))
    assert s["fp32_flop_per_trace"] == 64 * (10 + 1 + 2 + 2 * 5) / 2


def _bench_module():
    import importlib.util
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(root, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_bench_instruction_and_lane_figures():
    """bench.py's roofline figures from a counter summary: instructions per
    query (one wave per query) and active lanes per VALU instruction."""
    b = _bench_module()
    cs = {"per_dispatch_median": {"SQ_WAVES": 1000.0, "SQ_INSTS_VALU": 600_000.0, "SQ_INSTS_SALU": 500_000.0,
                                  "SQ_INSTS_LDS": 23_000.0, "SQ_INSTS_VMEM_RD": 32_000.0, "SQ_INSTS_SMEM": 21_000.0,
                                  "SQ_ACTIVE_INST_VALU": 625_000.0, "SQ_THREAD_CYCLES_VALU": 28_000_000.0}}
    ipq = b.insts_per_query(cs)
    assert ipq == {"valu": 600.0, "salu": 500.0, "lds": 23.0, "vmem_rd": 32.0, "smem": 21.0}
    assert abs(b.valu_lane_util(cs) - 28_000_000.0 / (64 * 625_000.0)) < 1e-12
    # missing counters give None, not a made-up figure
    assert b.insts_per_query(None) is None and b.insts_per_query({"per_dispatch_median": {}}) is None
    assert b.valu_lane_util({"per_dispatch_median": {"SQ_WAVES": 1.0}}) is None
