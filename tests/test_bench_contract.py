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
