"""AddressSanitizer / UndefinedBehaviorSanitizer builds of the host code
(CPU only): the oracle with a driver that runs every oracle function on
graphs with known answers (oracle/san_check.c), and the input tools
(trust_tool: undirect, preprocess in three row orders, txt2bin,
preprocess-bin, partition, the generators) on the golden fixtures and on
inputs they must reject.  Any sanitizer report makes the process exit
non-zero (-fno-sanitize-recover=all)."""
import filecmp
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT, golden_cases

SAN_TOOL = os.path.join(ROOT, "trust_amd", "bin", "trust_tool_san")
SAN_ORACLE = os.path.join(ROOT, "oracle", "_san", "san_check")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")


@pytest.fixture(scope="module")
def san_builds():
    for d in ("oracle", os.path.join("trust_amd", "csrc")):
        subprocess.run(["make", "-C", os.path.join(ROOT, d), "san"], check=True,
                       capture_output=True, timeout=600)
    return SAN_TOOL


def run(args, cwd, ok=True):
    r = subprocess.run(args, cwd=cwd, env=ENV, capture_output=True, text=True, timeout=300)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
    if ok:
        assert r.returncode == 0, r.stderr[-2000:]
    return r


def test_oracle_under_sanitizers(san_builds):
    r = run([SAN_ORACLE], ROOT)
    assert "ok" in r.stdout


@pytest.mark.parametrize("name", golden_cases())
def test_input_tools_under_sanitizers(san_builds, name, tmp_path):
    edges = os.path.join(GOLDEN, name, "edges.txt")
    run([SAN_TOOL, "undirect", edges, "1.mmio"], tmp_path)
    run([SAN_TOOL, "preprocess", "1.mmio", "."], tmp_path)
    for f in ("begin.bin", "adjacent.bin"):
        assert filecmp.cmp(tmp_path / f, os.path.join(GOLDEN, name, f), shallow=False)
    for order in ("plain", "indegree"):
        d = tmp_path / order
        d.mkdir()
        run([SAN_TOOL, "preprocess", str(tmp_path / "1.mmio"), str(d), "--order", order], tmp_path)
    b = tmp_path / "bin"
    b.mkdir()
    run([SAN_TOOL, "txt2bin", edges], b)
    run([SAN_TOOL, "preprocess-bin", str(b)], tmp_path)
    run([SAN_TOOL, "partition", "3", "."], tmp_path)


def test_generators_and_rejections_under_sanitizers(san_builds, tmp_path):
    run([SAN_TOOL, "gen", "rmat", "10", "8", "1", str(tmp_path / "r")], tmp_path)
    run([SAN_TOOL, "gen", "chunglu", "3000", "9000", "2.5", "80", "1", str(tmp_path / "c")], tmp_path)
    run([SAN_TOOL, "gen", "er", "2000", "8000", "1", str(tmp_path / "e")], tmp_path)
    run([SAN_TOOL, "gen", "ba", "2000", "4", "1", str(tmp_path / "b")], tmp_path)
    run([SAN_TOOL, "partition", "4", str(tmp_path / "r")], tmp_path)
    bad = {"neg.txt": "3 -5\n", "huge.txt": "1 99999999999\n", "odd.txt": "1 2 3\n"}
    for f, text in bad.items():
        (tmp_path / f).write_text(text)
    assert run([SAN_TOOL, "undirect", "neg.txt", "x.mmio"], tmp_path, ok=False).returncode != 0
    assert run([SAN_TOOL, "txt2bin", "neg.txt"], tmp_path, ok=False).returncode != 0
    assert run([SAN_TOOL, "txt2bin", "huge.txt"], tmp_path, ok=False).returncode != 0
    assert run([SAN_TOOL, "txt2bin", "odd.txt"], tmp_path, ok=False).returncode != 0
    (tmp_path / "m.mmio").write_text("3 3 1\n0 2 1\n")
    assert run([SAN_TOOL, "preprocess", "m.mmio", "."], tmp_path, ok=False).returncode != 0
    (tmp_path / "empty.txt").write_text("")
    run([SAN_TOOL, "undirect", "empty.txt", "e.mmio"], tmp_path)
    run([SAN_TOOL, "preprocess", "e.mmio", str(tmp_path)], tmp_path)
