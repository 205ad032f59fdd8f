"""The parts of bench.py's output contract that run without a GPU: the exit
for more than one GPU, the traffic figures it reads from the committed
rocprof summaries, and the CPU baseline object on a small graph."""
import importlib.util
import os
import subprocess
import sys

import trust_amd as T
from conftest import ROOT


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one MI355X" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line


def test_committed_traffic_matches_the_profiles():
    bench = load_bench()
    prof = os.path.join(ROOT, "profiles")
    found = 0
    for name in sorted(os.listdir(prof)):
        if not (name.startswith("tc_count_") and name.endswith("_traffic.txt")):
            continue
        found += 1
        workload = name[len("tc_count_"):-len("_traffic.txt")]
        fields = {}
        for line in open(os.path.join(prof, name)):
            if not line.startswith("#") and ":" in line:
                k, v = line.split(":", 1)
                fields[k.strip()] = float(v)
        t = bench.committed_traffic(workload)
        assert t == fields["traffic_bytes_per_launch"]
        # bytes = (2 x FETCH_SIZE + 1 x WRITE_SIZE) KiB: the copy-kernel calibration
        want = (2.0 * fields["fetch_kib_per_launch"] + fields["write_kib_per_launch"]) * 1024
        assert abs(t - want) <= 1e-4 * want
        # the summary the figure comes from is committed beside it
        assert os.path.exists(os.path.join(prof, f"tc_count_{workload}.md"))
    assert found >= 2  # the headline and the batch
    assert bench.committed_traffic("no_such_workload") is None


def test_cpu_baseline_object_on_a_small_graph():
    bench = load_bench()
    beg, adj = T.rmat_csr(12, 8, seed=1)
    line = bench.cpu_baseline(beg, adj, len(adj), 0, 0.05)
    assert set(line) == {"value", "unit", "cores", "kind", "sample"}
    assert line["kind"] == "port" and line["unit"] == "GTEPS"
    assert line["value"] > 0 and line["cores"] >= 1
    assert "oracle" in line["sample"]


def test_profile_commands_run_committed_files_only():
    # every summary under profiles/ names the command that made it; each runs
    # bench.py (under rocprofv3 or alone) with options bench.py defines
    import re

    src = open(os.path.join(ROOT, "bench.py")).read()
    known = set(re.findall(r'add_argument\("(--[a-z0-9-]+)"', src))
    prof = os.path.join(ROOT, "profiles")
    for name in sorted(os.listdir(prof)):
        cmds = []
        for line in open(os.path.join(prof, name)):
            line = line.strip().lstrip("#").strip()
            if line.startswith(("python3 ", "rocprofv3 ")):
                cmds.append(line)
        assert cmds, f"{name} names no command"
        for cmd in cmds:
            assert "_work" not in cmd and "/root/" not in cmd, cmd
            tail = cmd.split(" -- ", 1)[1] if cmd.startswith("rocprofv3") else cmd
            words = tail.split()
            assert words[0] == "python3" and "bench.py" in words, cmd
            for w in words:
                if w.startswith("--"):
                    assert w in known, (name, w)
