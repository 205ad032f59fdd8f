"""Out-of-core preprocessing (trust_rmat_csr_ooc, trust_tool gen-ooc): the
R-MAT stream preprocessed through bucket files must give exactly the files
the in-memory preprocessing gives (which are byte-identical to the
reference's tools, tests/test_preprocess*.py), for bucket sizes that force
many buckets, and its tiles must equal the in-memory partition's; its peak
host memory must stay far below the in-memory path's."""
import os
import subprocess
import sys

import numpy as np
import pytest

import trust_amd as T

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TOOL = os.path.join(ROOT, "trust_amd", "bin", "trust_tool")


def read_csr(folder):
    return (np.fromfile(os.path.join(folder, "begin.bin"), np.int64),
            np.fromfile(os.path.join(folder, "adjacent.bin"), np.int32))


@pytest.mark.parametrize("scale,ef,seed,bucket", [
    (8, 4, 3, 1024), (10, 16, 1, 1024), (12, 16, 2, 2048), (14, 16, 1, 4096),
    (16, 16, 1, 1 << 20), (15, 8, 7, 1 << 12),
])
def test_ooc_equals_in_memory(tmp_path, scale, ef, seed, bucket):
    out, work = str(tmp_path / "g"), str(tmp_path / "w")
    os.makedirs(work)
    T.rmat_csr_ooc(scale, ef, seed, out, work, bucket_pairs=bucket)
    beg, adj = T.preprocess_edges(*T.gen_rmat(scale, ef, seed=seed))
    b2, a2 = read_csr(out)
    assert np.array_equal(beg, b2) and np.array_equal(adj, a2)
    assert os.listdir(work) == []  # every bucket file removed


def test_ooc_rejects_bad_arguments(tmp_path):
    with pytest.raises(T.TrustError) as e:
        T.rmat_csr_ooc(10, 16, 1, str(tmp_path / "g"), str(tmp_path), bucket_pairs=10)
    assert e.value.code == T.ERR_ARG
    with pytest.raises(T.TrustError) as e:
        T.rmat_csr_ooc(10, 16, 1, str(tmp_path / "g"), str(tmp_path / "missing"), bucket_pairs=4096)
    assert e.value.code == T.ERR_IO


def test_cli_ooc_then_partition_equals_in_memory_tiles(tmp_path):
    out = str(tmp_path / "g")
    r = subprocess.run([TOOL, "gen-ooc", "rmat", "13", "16", "5", out, str(tmp_path), "2048"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([TOOL, "partition", "3", out], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    beg, adj = T.preprocess_edges(*T.gen_rmat(13, 16, seed=5))
    ref = str(tmp_path / "ref")
    T.partition_write(ref, beg, adj, 3)
    for i in range(3):
        for j in range(3):
            for f in ("begin.bin", "adjacent.bin"):
                a = open(os.path.join(out, f"partition{i}_{j}", f), "rb").read()
                b = open(os.path.join(ref, f"partition{i}_{j}", f), "rb").read()
                assert a == b, (i, j, f)


PEAK = r"""
import resource, sys
sys.path.insert(0, {root!r})
import trust_amd as T
T.lib()
base = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss
if sys.argv[1] == "ooc":
    T.rmat_csr_ooc(21, 16, 1, sys.argv[2], sys.argv[3], bucket_pairs=1 << 21)
else:
    beg, adj = T.preprocess_edges(*T.gen_rmat(21, 16, seed=1))
print(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss - base)
"""


def test_ooc_peak_memory_is_bounded(tmp_path):
    # R-MAT 21 (33.5M pairs): the in-memory path holds the pairs, their keys
    # and a sort buffer; the out-of-core path holds vertex arrays and buckets
    # of 2^21 keys
    code = PEAK.format(root=ROOT)
    os.makedirs(tmp_path / "w")
    peak = {}
    for mode in ("ooc", "mem"):
        r = subprocess.run([sys.executable, "-c", code, mode, str(tmp_path / "g"),
                            str(tmp_path / "w")], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        peak[mode] = int(r.stdout.split()[-1]) * 1024  # ru_maxrss is in KiB on Linux
    print(f"peak host memory above the loaded library: out of core {peak['ooc'] >> 20} MiB, "
          f"in memory {peak['mem'] >> 20} MiB")
    assert peak["ooc"] * 4 < peak["mem"]
