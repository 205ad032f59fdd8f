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


def test_journal_resume(gpu, rmat14, tmp_path):
    # an interrupted stream resumes from its journal: stop after a few jobs
    # (forced from the host), resume twice, and every job's count, the sum
    # and the journal's lines must match the uninterrupted stream
    beg, adj, want = rmat14
    n = 3
    tiles = T.partition_csr(beg, adj, n)
    full = T.count_tiles(tiles, n)
    assert full.triangles == want
    path = str(tmp_path / "jobs.journal")
    left, r1, _ = T.count_tiles_journal(tiles, n, path, max_new_jobs=5)
    assert left == n ** 3 - 5 and r1.triangles <= want
    lines = open(path).read().splitlines()
    assert lines[0].startswith("trust-tiles-journal 1 n=3 ") and len(lines) == 6
    left, r2, _ = T.count_tiles_journal(tiles, n, path, max_new_jobs=10)
    assert left == n ** 3 - 15 and r1.triangles <= r2.triangles <= want
    # a write cut short: the last line loses its newline and is counted again
    text = open(path).read()
    open(path, "w").write(text[:-3])
    left, r3, jobs = T.count_tiles_journal(tiles, n, path)
    assert left == 0 and r3.triangles == want and r3.probes == full.probes
    lines = open(path).read().splitlines()
    assert len(lines) == 1 + n ** 3
    ids = sorted(int(x.split()[1]) for x in lines[1:])
    assert ids == list(range(n ** 3))
    for x in lines[1:]:
        _, job, i, j, k, c, _w = x.split()
        job, i, j, k, c = int(job), int(i), int(j), int(k), int(c)
        assert (i, j, k) == (job // (n * n), job // n % n, job % n)
        assert c == jobs[job]
    assert int(jobs.sum()) == want
    # a finished journal counts nothing more and gives the same result
    left, r4, _ = T.count_tiles_journal(tiles, n, path)
    assert left == 0 and r4.triangles == want and r4.count_seconds == 0
    # a journal of another partition is refused
    with pytest.raises(T.TrustError) as e:
        T.count_tiles_journal(T.partition_csr(beg, adj, 2), 2, path)
    assert e.value.code == T.ERR_FORMAT
