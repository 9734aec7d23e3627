"""bench.py's host-side contract, without a GPU: a run on more than one GPU is refused with one JSON
line, and the HBM traffic it reports comes from the committed rocprofv3 summary of the frame
kernel, whose figures agree with each other."""
import json
import os
import re
import subprocess
import sys
import types

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_profiled_traffic_matches_the_summary():
    import bench
    args = types.SimpleNamespace(res=4096, width=1920, height=1080)
    t = bench.profiled_traffic(args)
    text = open(os.path.join(ROOT, bench.TRAFFIC_PROFILE)).read()
    # the summary's total is its read + write lines (MB, one decimal) and its steady-state average
    # is the mean of the last two kernel durations
    m = re.search(r"read ([0-9.]+) MB \+ write ([0-9.]+) MB = \*\*([0-9.]+) MB\*\*", text)
    assert m and t is not None
    assert abs(float(m.group(1)) + float(m.group(2)) - float(m.group(3))) < 0.15
    assert abs(t / 1e6 - float(m.group(3))) < 0.06
    durs = [float(x) for x in re.search(r"\| k_trace<false> \| 3 \| ([0-9., ]+) \|", text).group(1).split(",")]
    avg = float(re.search(r"trace_avg_us: ([0-9.]+)", text).group(1))
    assert abs((durs[1] + durs[2]) / 2 - avg) < 0.1
    # another workload has no profiled traffic
    assert bench.profiled_traffic(types.SimpleNamespace(res=1024, width=1920, height=1080)) is None


def test_profiled_valu_issue_matches_the_counters():
    import bench
    args = types.SimpleNamespace(res=4096, width=1920, height=1080)
    f = bench.profiled_valu_issue(args)
    text = open(os.path.join(ROOT, bench.VALU_PROFILE)).read()
    trace = open(os.path.join(ROOT, bench.TRAFFIC_PROFILE)).read()
    # recompute from the table: mean SQ_INSTS_VALU of frames 2-3 over the issue capacity of the
    # kernel-trace average (1024 SIMDs, 4 cycles per wave64 instruction, 2.4 GHz)
    row = re.search(r"\| SQ_INSTS_VALU \| ([^\n]+) \|", text).group(1).split("|")
    valu = (float(row[1]) + float(row[2])) / 2
    avg_us = float(re.search(r"trace_avg_us: ([0-9.]+)", trace).group(1))
    cap = 256 * 4 * avg_us * 1e-6 * 2.4e9 / 4
    assert f is not None and 0.0 < f <= 1.0
    assert abs(valu / cap - f) < 0.01
    assert bench.profiled_valu_issue(types.SimpleNamespace(res=1024, width=1920, height=1080)) is None
