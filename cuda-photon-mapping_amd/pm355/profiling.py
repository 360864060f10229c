"""rocprofv3 counter summaries for the benchmark's roofline.

The benchmark times its dominant kernel live (HIP events); the bytes that
kernel moves through each level of the memory hierarchy come from hardware
counters, which need a profiler around the process, so they are collected
in separate rocprofv3 runs and summarised here into a small JSON file that
bench.py reads. Each summary records a hash of the HIP sources it was
measured on; bench.py flags the numbers as stale when the sources have
changed since.

    python -m pm355.profiling summarize --kernel k_knn_stream512 \
        --pmc DIR [DIR ...] --copy DIR [DIR ...] --meta sources=gather key=value ... --out FILE
    python -m pm355.profiling valu --pmc DIR --traces 2 --out FILE

DIRs are rocprofv3 output directories (-d) of --pmc runs; --copy runs are
of `bench.py --copy-probe` (three 1 GiB device copies), which scale
FETCH_SIZE / WRITE_SIZE to bytes.
"""
from __future__ import annotations

import argparse
import csv
import glob
import hashlib
import json
import os
import re
import statistics

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PKG, "csrc")

# aggregate L2 read bandwidth of MI355X: 8 XCDs x 16 channels x 128 B per
# clock x 2.4 GHz
L2_PEAK_GBS = 8 * 16 * 128 * 2.4
HBM_PEAK_GBS = 8000.0
COPY_BYTES = 1 << 30


# The sources each profiled kernel is compiled from: its own file and the
# shared headers (a change anywhere else cannot change the kernel).
KERNEL_SOURCES = {
    "gather": ["knn.hip", "pm_select.hpp", "pm_internal.hpp"],
    "trace": ["trace.hip", "pm_geom.hpp", "pm_internal.hpp"],
}


def source_hash(files=None) -> str:
    """sha256 over HIP sources of libpm355.so (names and contents): the
    given file names under csrc/, or every source."""
    h = hashlib.sha256()
    paths = sorted(glob.glob(os.path.join(CSRC, "*"))) if files is None else \
        [os.path.join(CSRC, f) for f in sorted(files)]
    for p in paths:
        h.update(os.path.basename(p).encode())
        with open(p, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def _rows(d):
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            yield from csv.DictReader(f)


def per_dispatch(dirs, kernel):
    """{counter: [value per dispatch of `kernel`]} over rocprofv3 output dirs."""
    # the kernel's own name, bare or after a namespace / return type, followed
    # by its argument list or template arguments
    pat = re.compile(r"(^|[\s:])" + re.escape(kernel) + r"($|[(<])")
    acc = {}
    for d in dirs:
        vals = {}
        for r in _rows(d):
            if not pat.search(r.get("Kernel_Name", "")):
                continue
            key = (r.get("Dispatch_Id"), r["Counter_Name"])
            vals[key] = vals.get(key, 0.0) + float(r["Counter_Value"])
        for (_, cname), v in vals.items():
            acc.setdefault(cname, []).append(v)
    return acc


def summarize(kernel, pmc_dirs, copy_dirs, meta, out):
    k = per_dispatch(pmc_dirs, kernel)
    c = per_dispatch(copy_dirs, "__amd_rocclr_copyBuffer")
    med = {n: statistics.median(v) for n, v in k.items()}
    files = KERNEL_SOURCES.get(meta.get("sources", ""))
    res = {"kernel": kernel, "source_hash": source_hash(files), "source_files": files,
           "dispatches": {n: len(v) for n, v in k.items()}, "per_dispatch_median": med, **meta}
    # FETCH_SIZE / WRITE_SIZE in KiB; the copy probe's 1 GiB copies give the
    # scale from counter units to bytes read / written
    if "FETCH_SIZE" in med:
        scale = COPY_BYTES / (max(c["FETCH_SIZE"]) * 1024) if c.get("FETCH_SIZE") else 1.0
        res["fetch_scale"] = scale
        res["hbm_read_bytes_per_dispatch"] = med["FETCH_SIZE"] * 1024 * scale
    if "WRITE_SIZE" in med:
        scale = COPY_BYTES / (max(c["WRITE_SIZE"]) * 1024) if c.get("WRITE_SIZE") else 1.0
        res["write_scale"] = scale
        res["hbm_write_bytes_per_dispatch"] = med["WRITE_SIZE"] * 1024 * scale
    if "TCC_READ_SECTORS_sum" in med:
        res["l2_read_bytes_per_dispatch"] = 32.0 * med["TCC_READ_SECTORS_sum"]
    if "TCC_HIT_sum" in med and "TCC_MISS_sum" in med:
        res["l2_hit_rate"] = med["TCC_HIT_sum"] / max(med["TCC_HIT_sum"] + med["TCC_MISS_sum"], 1.0)
    if "TCP_TCC_READ_REQ_sum" in med and "TCP_TOTAL_CACHE_ACCESSES_sum" in med:
        res["l1_hit_rate"] = 1.0 - med["TCP_TCC_READ_REQ_sum"] / max(med["TCP_TOTAL_CACHE_ACCESSES_sum"], 1.0)
    if "TCP_TCC_READ_REQ_LATENCY_sum" in med and "TCP_TCC_READ_REQ_sum" in med:
        res["l2_read_latency_cycles"] = med["TCP_TCC_READ_REQ_LATENCY_sum"] / max(med["TCP_TCC_READ_REQ_sum"], 1.0)
    with open(out, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    return res


def summarize_valu(pmc_dirs, traces, meta, out):
    """FP32 VALU work of a run that traced `traces` photon maps: the
    SQ_INSTS_VALU_{ADD,MUL,TRANS,FMA}_F32 wave instructions of every
    dispatch, x 64 lanes (FMA x 2 flops), per trace."""
    tot = {}
    for d in pmc_dirs:
        for r in _rows(d):
            tot[r["Counter_Name"]] = tot.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    flop = 64.0 * (tot.get("SQ_INSTS_VALU_ADD_F32", 0) + tot.get("SQ_INSTS_VALU_MUL_F32", 0)
                   + tot.get("SQ_INSTS_VALU_TRANS_F32", 0) + 2 * tot.get("SQ_INSTS_VALU_FMA_F32", 0))
    files = KERNEL_SOURCES["trace"]
    res = {"fp32_flop_per_trace": flop / traces, "counter_totals": tot, "traces": traces,
           "source_hash": source_hash(files), "source_files": files, **meta}
    with open(out, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    return res


def load(path, **match):
    """A summary written by summarize(), with "stale" set when the HIP
    sources changed since it was measured; None if the file is missing or
    any `match` field differs."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        s = json.load(f)
    for key, want in match.items():
        got = s.get(key)
        if isinstance(want, float) and isinstance(got, (int, float)):
            if abs(got - want) > 1e-6:
                return None
        elif got != want:
            return None
    s["stale"] = s.get("source_hash") != source_hash(s.get("source_files"))
    return s


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("summarize")
    s.add_argument("--kernel", required=True)
    s.add_argument("--pmc", nargs="+", required=True)
    s.add_argument("--copy", nargs="*", default=[])
    s.add_argument("--meta", nargs="*", default=[], help="key=value (numbers parsed)")
    s.add_argument("--out", required=True)
    v = sub.add_parser("valu")
    v.add_argument("--pmc", nargs="+", required=True)
    v.add_argument("--traces", type=int, required=True)
    v.add_argument("--meta", nargs="*", default=[])
    v.add_argument("--out", required=True)
    hs = sub.add_parser("hash")
    hs.add_argument("--sources", default="", help="kernel source set (gather, trace); empty = every source")
    a = ap.parse_args(argv)
    if a.cmd == "hash":
        print(source_hash(KERNEL_SOURCES.get(a.sources)))
        return
    meta = {}
    for kv in a.meta:
        key, val = kv.split("=", 1)
        try:
            meta[key] = float(val)
        except ValueError:
            meta[key] = val
    if a.cmd == "valu":
        print(json.dumps(summarize_valu(a.pmc, a.traces, meta, a.out), sort_keys=True))
    else:
        print(json.dumps(summarize(a.kernel, a.pmc, a.copy, meta, a.out), sort_keys=True))


if __name__ == "__main__":
    main()
