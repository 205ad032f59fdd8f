"""GPU parity of the 2D partition path: tile-streamed sums equal the
unpartitioned count (With-graph-partition/TC.cu semantics, one GPU)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import trust_amd as T
from conftest import GOLDEN, golden_cases, load_golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat14():
    beg, adj = T.rmat_csr(14, 16, seed=3)
    return beg, adj, oracle.count_hash(beg, adj)


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_stream_all_jobs(gpu, rmat14, n):
    beg, adj, want = rmat14
    tiles = T.partition_csr(beg, adj, n)
    print(f"[gpu] tiles n={n}", flush=True)
    r = T.count_tiles(tiles, n, device=gpu)
    assert r.triangles == want


@pytest.mark.parametrize("kw", [dict(reserved=[3]), dict(reserved=[5], threads=128),
                                dict(flags=T.FLAG_FORCE_SPILL), dict(hub_threshold=0),
                                dict(flags=T.FLAG_OFFSETS64)])
def test_stream_small_cache_and_params(gpu, rmat14, kw):
    beg, adj, want = rmat14
    tiles = T.partition_csr(beg, adj, 3)
    assert T.count_tiles(tiles, 3, device=gpu, **kw).triangles == want


def test_job_counts_match_oracle(gpu, rmat14):
    beg, adj, _ = rmat14
    n = 2
    tiles = T.partition_csr(beg, adj, n)
    folder = "/tmp/trust_part_jobs"
    T.partition_write(folder, beg, adj, n)
    for rank in range(n ** 3):
        i, j, k = rank % n, rank // n % n, rank // n // n % n
        want = oracle.count_tile_job(tiles[i * n + j], tiles[i * n + k], tiles[k * n + j])
        a = T.triangle_count_partition(rank, folder, n ** 3, 1024, 1024, 1, n)
        assert a.count == want, (i, j, k)


@pytest.mark.parametrize("name", golden_cases())
def test_reference_tile_folders(gpu, name):
    # tiles written by the reference's own partition tool
    _, _, meta = load_golden(name)
    for n in (2, 3):
        folder = os.path.join(GOLDEN, name, f"part{n}")
        total = sum(T.triangle_count_partition(r, folder, n ** 3, 256, 64, 1, n).count
                    for r in range(n ** 3))
        assert total == meta["triangles_networkx"], n


def test_cli_partitioned(gpu, rmat14, tmp_path):
    beg, adj, want = rmat14
    n = 2
    T.partition_write(str(tmp_path), beg, adj, n)
    exe = os.path.join(T.BIN_DIR, "trianglecounting")
    out = subprocess.run([exe, str(tmp_path), str(2 * n ** 3), "1024", "1024", "1", str(n)],
                         capture_output=True, text=True, check=True, timeout=120).stdout
    line = out.strip().splitlines()[-1]
    assert line.split(",")[1] == f"triangle:{want}"
