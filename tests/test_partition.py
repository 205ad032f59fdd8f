"""2D hash-partition tiles (Preprocess/partition.cpp layout), CPU side."""
import filecmp
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
import trust_amd as T
from conftest import GOLDEN, golden_cases, load_golden


@pytest.mark.parametrize("name", golden_cases())
@pytest.mark.parametrize("n", [2, 3])
def test_tiles_byte_identical_to_reference_golden(name, n, tmp_path):
    beg, adj, _ = load_golden(name)
    T.partition_write(str(tmp_path), beg, adj, n)
    ref = os.path.join(GOLDEN, name, f"part{n}")
    for i in range(n):
        for j in range(n):
            for fn in ("begin.bin", "adjacent.bin", "map", "division"):
                a = os.path.join(ref, f"partition{i}_{j}", fn)
                b = os.path.join(str(tmp_path), f"partition{i}_{j}", fn)
                assert filecmp.cmp(a, b, shallow=False), (i, j, fn)


@pytest.mark.skipif(oracle.ref_tool("partition") is None, reason="reference tools not built")
@pytest.mark.parametrize("n", [2, 4])
def test_tiles_byte_identical_to_reference_tool(n, tmp_path):
    beg, adj = T.rmat_csr(12, 8, seed=11)
    ref_dir = tmp_path / "ref"
    ours = tmp_path / "ours"
    T.save_folder(str(ref_dir), beg, adj)
    for i in range(n):
        for j in range(n):
            os.makedirs(ref_dir / f"partition{i}_{j}")
    subprocess.run([oracle.ref_tool("partition"), str(n)], cwd=ref_dir, check=True,
                   capture_output=True)
    T.partition_write(str(ours), beg, adj, n)
    for i in range(n):
        for j in range(n):
            for fn in ("begin.bin", "adjacent.bin", "map", "division"):
                assert filecmp.cmp(ref_dir / f"partition{i}_{j}" / fn,
                                   ours / f"partition{i}_{j}" / fn, shallow=False)


@pytest.mark.parametrize("name", ["rmat10", "ba2k", "tiny"])
@pytest.mark.parametrize("n", [2, 3])
def test_folder_writer_matches_golden(name, n, tmp_path):
    # the mapped, row-at-a-time writer (trust_tool partition) gives the same
    # bytes as the reference tool
    beg, adj, _ = load_golden(name)
    T.save_folder(str(tmp_path), beg, adj)
    tool = os.path.join(T.BIN_DIR, "trust_tool")
    subprocess.run([tool, "partition", str(n), str(tmp_path)], check=True, capture_output=True)
    ref = os.path.join(GOLDEN, name, f"part{n}")
    for i in range(n):
        for j in range(n):
            for fn in ("begin.bin", "adjacent.bin", "map", "division"):
                a = os.path.join(ref, f"partition{i}_{j}", fn)
                b = os.path.join(str(tmp_path), f"partition{i}_{j}", fn)
                assert filecmp.cmp(a, b, shallow=False), (i, j, fn)


@pytest.mark.parametrize("n", [11, 16])
def test_folder_writer_beyond_ten_parts(n, tmp_path):
    # past the reference's n <= 10: every tile matches the in-memory tiles
    beg, adj = T.rmat_csr(11, 8, seed=4)
    T.save_folder(str(tmp_path), beg, adj)
    T.partition_folder(str(tmp_path), n)
    tiles = T.partition_csr(beg, adj, n)
    for i in range(n):
        for j in range(n):
            tb, ta = T.load_folder(str(tmp_path / f"partition{i}_{j}"))
            assert np.array_equal(tb, tiles[i * n + j][0])
            assert np.array_equal(ta, tiles[i * n + j][1])
    assert sum(len(t[1]) for t in tiles) == len(adj)


def test_folder_writer_rejects_bad_input(tmp_path):
    with pytest.raises(T.TrustError) as e:
        T.partition_folder(str(tmp_path / "missing"), 2)
    assert e.value.code == T.ERR_IO
    np.array([0, 2, 3], np.int64).tofile(tmp_path / "begin.bin")
    np.array([1, 5, 0], np.int32).tofile(tmp_path / "adjacent.bin")  # id 5 >= V = 2
    with pytest.raises(T.TrustError) as e:
        T.partition_folder(str(tmp_path), 2)
    assert e.value.code == T.ERR_FORMAT
    T.save_folder(str(tmp_path), np.array([0, 1, 1]), np.array([1]))
    with pytest.raises(T.TrustError) as e:
        T.partition_folder(str(tmp_path), 1025)
    assert e.value.code == T.ERR_ARG


def test_tile_layout_definition():
    beg, adj = T.rmat_csr(9, 8, seed=2)
    n = 3
    tiles = T.partition_csr(beg, adj, n)
    V = len(beg) - 1
    for i in range(n):
        for j in range(n):
            tb, ta = tiles[i * n + j]
            rows = len(range(i, V, n))
            assert len(tb) - 1 == rows
            for r in range(rows):
                u = r * n + i
                want = [v // n for v in adj[beg[u]:beg[u + 1]] if v % n == j]
                assert ta[tb[r]:tb[r + 1]].tolist() == want


def test_partition_rejects_bad_n():
    beg, adj = T.rmat_csr(8, 4, seed=1)
    with pytest.raises(T.TrustError) as e:
        T.partition_csr(beg, adj, 0)
    assert e.value.code == T.ERR_ARG
