"""GPU parity at the full workload sizes (the five benchmark configurations).

Each graph is generated on the host exactly as bench.py builds it, checked
to be the graph the committed oracle counts were made from (V, E and xxh64
fingerprints in tests/golden/large_counts.json, written by
tests/golden/make_large_counts.py with the CPU oracle), then counted on the
GPU through the C ABI and compared with the oracle's count:

  R-MAT 16                 resident count (every walk), and the oracle live
  Chung-Lu 3.8M            resident count, fresh and with the index kept,
                           at the default and the reference's launch
  R-MAT 24                 resident count; 4 x 4 tiles streamed through the
                           GPU sum to the same count (the partition config's
                           parity scale)
  R-MAT 26                 resident count (the headline graph)
  10,000-graph ER/BA batch every per-graph count, eager and hipGraph replays
"""
import json
import os

import numpy as np
import pytest
import xxhash

import oracle
import trust_amd as T
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

FIXTURES = json.load(open(os.path.join(GOLDEN, "large_counts.json")))


def fingerprint_ok(name, beg, adj):
    fx = FIXTURES[name]
    assert (len(beg) - 1, len(adj)) == (fx["V"], fx["E"]), name
    assert xxhash.xxh64(memoryview(np.ascontiguousarray(adj))).hexdigest() == fx["adj_xxh64"]
    assert xxhash.xxh64(memoryview(np.ascontiguousarray(beg))).hexdigest() == fx["beg_xxh64"]
    return fx


def count_all_ways(ctx, fx, label):
    want = fx["triangles"]
    runs = {
        "default, fresh": ctx.count(flags=T.FLAG_FRESH_INDEX),
        "default, index kept": ctx.count(),
        "reference walk": ctx.count(flags=T.FLAG_REFERENCE_WALK),
        "short-side walk": ctx.count(flags=T.FLAG_SHORT_WALK),
        "reference launch": ctx.count(threads=1024, blocks=1024, chunk=1),
    }
    for k, r in runs.items():
        print(f"[gpu] {label} {k}: {r.triangles} ({r.seconds * 1e3:.2f} ms, index "
              f"{r.index_seconds * 1e3:.2f} ms, walk {r.walk_used}, core {r.core_vertices} "
              f"vertices, {r.dense_edges} dense edges)", flush=True)
        assert r.triangles == want, (label, k)
    assert runs["reference walk"].probes == fx["probes_reference_walk"]
    # the dense core takes some edges' probes off the short-side walk
    short = runs["short-side walk"]
    assert short.probes + short.dense_probes == fx["probes_short_walk"]
    return runs


def test_rmat16_full_config(gpu):
    print("[gpu] config 1: R-MAT 16", flush=True)
    beg, adj = T.rmat_csr(16, 16, seed=1)
    fx = fingerprint_ok("rmat16", beg, adj)
    assert oracle.count_hash(beg, adj) == fx["triangles"]
    with T.Context(gpu) as c:
        c.load(beg, adj)
        count_all_ways(c, fx, "rmat16")


def test_chung_lu_full_config(gpu):
    print("[gpu] config 2: Chung-Lu 3.8M", flush=True)
    beg, adj = T.preprocess_edges(*T.gen_chung_lu(3_800_000, 16_500_000, 2.5, 800.0, seed=1))
    fx = fingerprint_ok("chunglu3.8M", beg, adj)
    assert oracle.count_hash(beg, adj) == fx["triangles"]  # 0.2 s on the host
    with T.Context(gpu) as c:
        c.load(beg, adj)
        runs = count_all_ways(c, fx, "chung-lu")
        assert runs["default, fresh"].index_seconds > 0
        assert runs["default, index kept"].index_seconds == 0


@pytest.fixture(scope="module")
def rmat24():
    beg, adj = T.rmat_csr(24, 16, seed=1)
    return beg, adj, fingerprint_ok("rmat24", beg, adj)


@pytest.mark.timeout(600)
def test_rmat24_resident(gpu, rmat24):
    print("[gpu] config 4 parity scale: R-MAT 24 resident", flush=True)
    beg, adj, fx = rmat24
    with T.Context(gpu) as c:
        c.load(beg, adj)
        count_all_ways(c, fx, "rmat24")


@pytest.mark.timeout(600)
def test_rmat24_tiles_4x4(gpu, rmat24):
    print("[gpu] config 4 parity scale: R-MAT 24, 4 x 4 tiles streamed", flush=True)
    beg, adj, fx = rmat24
    tiles = T.partition_csr(beg, adj, 4)
    r = T.count_tiles(tiles, 4)
    print(f"[gpu] rmat24 tiles: {r.triangles} in {r.seconds * 1e3:.1f} ms", flush=True)
    assert r.triangles == fx["triangles"]


@pytest.mark.timeout(900)
def test_rmat26_resident(gpu):
    print("[gpu] config 3: R-MAT 26 resident", flush=True)
    beg, adj = T.rmat_csr(26, 16, seed=1)
    fx = fingerprint_ok("rmat26", beg, adj)
    with T.Context(gpu) as c:
        c.load(beg, adj)
        del beg, adj
        for flags in (T.FLAG_FRESH_INDEX, 0, T.FLAG_REFERENCE_WALK):
            r = c.count(flags=flags)
            print(f"[gpu] rmat26 flags {flags}: {r.triangles} in {r.seconds * 1e3:.1f} ms "
                  f"(index {r.index_seconds * 1e3:.1f} ms, walk {r.walk_used}, core "
                  f"{r.core_vertices} vertices, {r.dense_edges} dense edges)", flush=True)
            assert r.triangles == fx["triangles"]
            assert r.probes + r.dense_probes == (fx["probes_short_walk"] if r.walk_used
                                                 else fx["probes_reference_walk"])
            if r.walk_used:
                assert r.core_vertices > 0 and r.dense_edges > 0  # the automatic core


@pytest.mark.timeout(900)
def test_batch10000_full_config(gpu):
    print("[gpu] config 5: 10,000-graph ER/BA batch", flush=True)
    fx = FIXTURES["batch10000"]
    want = np.load(os.path.join(GOLDEN, fx["per_graph_counts"]))
    beg, adj, vb = T.batch_generate(10000, seed=1, nmin=1000, nmax=200000, dmin=8.0, dmax=64.0)
    fingerprint_ok("batch10000", beg, adj)
    assert xxhash.xxh64(memoryview(vb)).hexdigest() == fx["vbase_xxh64"]
    with T.Context(gpu) as c:
        c.load_batch(beg, adj, vb)
        del beg, adj
        for flags in (0, T.FLAG_HIP_GRAPH, T.FLAG_HIP_GRAPH):
            counts, done, r = c.count_batch(flags=flags)
            print(f"[gpu] batch10000 flags {flags}: {r.triangles} in {r.seconds * 1e3:.1f} ms",
                  flush=True)
            bad = np.nonzero(counts != want)[0]
            assert len(bad) == 0, f"graphs {bad[:10].tolist()} differ from the oracle"
            assert r.triangles == fx["triangles"]
