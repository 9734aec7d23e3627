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


def test_edit_roofline_from_the_algorithmic_bytes_table():
    # the edit path's roofline numerator comes from the committed oracle count of the benchmark's
    # steps (oracle/edit_bytes.py); the table's byte column equals 4 x the sum of its word columns
    import bench
    eb = bench._profile_values(bench.EDIT_BYTES_PROFILE)
    steps = eb["plain_step_bytes"]
    assert len(steps) >= 25 and all(b > 0 for b in steps)
    text = open(os.path.join(ROOT, bench.EDIT_BYTES_PROFILE)).read()
    rows = re.findall(r"^\| (\d+) \| (\d+) \| (\d+) \| (\d+) \| (\d+) \| (\d+) \| (\d+) \| (\d+) \| (\d+) \|$", text, re.M)
    assert len(rows) == len(steps)
    for r in rows:
        v = [int(x) for x in r]
        assert v[8] == 4 * (v[2] + v[3] + v[4] + v[5] + v[6]) == steps[v[0]]
        assert v[1] == eb["plain_step_voxels"][v[0]]
    args = types.SimpleNamespace(res=4096, edits=256, mode="plain", warmup=1, steps=3)
    r = bench.edit_roofline(args, [40.0, 40.0, 40.0])
    assert r["bound"] == "hbm" and r["unit"] == "GB/s"
    assert abs(r["achieved"] - sum(steps[1:4]) / 3 / 0.040 / 1e9) < 1e-6
    assert abs(r["frac"] - r["achieved"] / 8000.0) < 1e-12
    assert bench.edit_roofline(types.SimpleNamespace(res=1024, edits=256, mode="plain", warmup=1, steps=3),
                               [1.0]) is None
