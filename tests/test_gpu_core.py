"""GPU parity of the dense core (core.hip) against the CPU oracle.

The core counts the edges from dense sources to dense targets with bit rows
and the walk leaves them out, so every count here must equal the oracle's,
and the walk's probes plus the probes the core took over must equal the
oracle's short-side walk probes.  Core sizes and degree bounds are forced
small so that graphs of a few thousand vertices have a core; the automatic
core (single graphs of at least 2^23 edges) is exercised by the full-size
configurations (test_gpu_configs.py) and bench.py.
"""
import numpy as np
import pytest

import oracle
import trust_amd as T
from conftest import CLOSED_FORMS, closed_form_csr, golden_cases, load_golden

pytestmark = pytest.mark.gpu

SHAPES = [(256, 2), (256, 16), (1024, 8), (4096, 64), (16384, 2), (16384, 64)]


@pytest.fixture(scope="module")
def rmat16():
    beg, adj = T.rmat_csr(16, 16, seed=1)
    return beg, adj, oracle.count_hash(beg, adj), oracle.probes_short(beg, adj)


@pytest.fixture(scope="module")
def ctx16(gpu, rmat16):
    beg, adj, _, _ = rmat16
    c = T.Context(gpu)
    c.load(beg, adj)
    yield c
    c.close()


@pytest.mark.parametrize("K,thr", SHAPES)
def test_core_counts_exactly_rmat16(ctx16, rmat16, K, thr):
    _, _, want, short = rmat16
    kw = dict(flags=T.FLAG_SHORT_WALK, core_vertices=K, core_min_degree=thr)
    print(f"[gpu] core rmat16 K={K} thr={thr}", flush=True)
    r = ctx16.count(**kw)
    print(f"[gpu]   core {r.core_vertices} vertices, {r.dense_edges} dense edges, "
          f"{r.dense_probes} of {short} probes", flush=True)
    assert r.triangles == want
    assert 0 < r.core_vertices <= 32768 and r.dense_edges > 0  # the seed, closed
    assert r.probes + r.dense_probes == short
    # kept index, then the gather-only pass over the same walk rows
    r2 = ctx16.count(**kw)
    assert r2.triangles == want and r2.index_seconds == 0 and r2.probes == r.probes
    assert ctx16.gather(**kw).probes == r.probes


@pytest.mark.parametrize("kw", [dict(flags=T.FLAG_FORCE_SPILL), dict(flags=T.FLAG_WAVE_ONLY),
                                dict(flags=T.FLAG_REBUILD_LINEAR),
                                dict(flags=T.FLAG_LINEAR_PROBING),
                                dict(threads=1024, blocks=1024, chunk=1),
                                dict(threads=64, blocks=7), dict(hub_threshold=0)])
def test_core_with_every_tier(ctx16, rmat16, kw):
    _, _, want, short = rmat16
    kw = dict(kw)
    kw["flags"] = kw.get("flags", 0) | T.FLAG_SHORT_WALK
    print(f"[gpu] core rmat16 {kw}", flush=True)
    r = ctx16.count(core_vertices=1024, core_min_degree=8, **kw)
    assert r.triangles == want and r.dense_edges > 0
    assert r.probes + r.dense_probes == short


def test_core_int64_offsets(gpu, rmat16):
    beg, adj, want, short = rmat16
    with T.Context(gpu) as c:
        c.load(beg, adj, flags=T.FLAG_OFFSETS64)
        print("[gpu] core rmat16 int64 offsets", flush=True)
        r = c.count(flags=T.FLAG_SHORT_WALK, core_vertices=4096, core_min_degree=16)
        assert r.triangles == want and r.dense_edges > 0
        assert r.probes + r.dense_probes == short


def test_core_switching_rebuilds_the_walk(ctx16, rmat16):
    # a count that asks for another core (or none) makes the walk rows again
    _, _, want, short = rmat16
    seq = [(2048, 16), (-1, 0), (256, 4), (2048, 16)]
    for K, thr in seq:
        print(f"[gpu] core switch K={K} thr={thr}", flush=True)
        r = ctx16.count(flags=T.FLAG_SHORT_WALK, core_vertices=K, core_min_degree=thr)
        assert r.triangles == want
        if K < 0:
            assert r.core_vertices == 0 and r.dense_edges == 0 and r.probes == short
        else:
            assert r.dense_edges > 0 and r.probes + r.dense_probes == short
            assert r.index_seconds > 0


def test_core_reference_walk_and_learned(ctx16, rmat16):
    _, _, want, short = rmat16
    kw = dict(core_vertices=2048, core_min_degree=8)
    print("[gpu] core with the reference walk and the learned order", flush=True)
    r = ctx16.count(flags=T.FLAG_REFERENCE_WALK, **kw)
    assert r.triangles == want and r.walk_used == 0 and r.dense_edges == 0
    ctx16.train_cost_model(epochs=3, max_samples=20000, seed=1, flags=T.FLAG_SHORT_WALK, **kw)
    r = ctx16.count(schedule=T.SCHED_LEARNED, flags=T.FLAG_SHORT_WALK, **kw)
    assert r.triangles == want and r.dense_edges > 0
    assert r.probes + r.dense_probes == short


@pytest.mark.parametrize("name", golden_cases())
def test_core_golden(gpu, name):
    beg, adj, meta = load_golden(name)
    print(f"[gpu] core golden {name}", flush=True)
    for K, thr in ((256, 2), (1024, 4)):
        r = T.count_csr(beg, adj, device=gpu, flags=T.FLAG_SHORT_WALK, core_vertices=K,
                        core_min_degree=thr)
        assert r.triangles == meta["triangles_networkx"]
        assert r.probes + r.dense_probes == oracle.probes_short(beg, adj)


@pytest.mark.parametrize("name", sorted(CLOSED_FORMS))
def test_core_closed_forms(gpu, name):
    beg, adj, want = closed_form_csr(name)
    print(f"[gpu] core closed form {name}", flush=True)
    r = T.count_csr(beg, adj, device=gpu, flags=T.FLAG_SHORT_WALK, core_vertices=256,
                    core_min_degree=2)
    assert r.triangles == want


def test_core_other_orderings(gpu):
    # orientations not by degree: a core vertex's list may leave the core,
    # which only makes fewer dense targets
    src, dst = T.gen_rmat(14, 16, seed=4)
    for order in ("plain", "indegree"):
        beg, adj = T.preprocess_edges(src, dst, order=order)
        want = oracle.count_hash(beg, adj)
        print(f"[gpu] core order {order}", flush=True)
        r = T.count_csr(beg, adj, device=gpu, flags=T.FLAG_SHORT_WALK, core_vertices=1024,
                        core_min_degree=4)
        assert r.triangles == want
        assert r.probes + r.dense_probes == oracle.probes_short(beg, adj)


def test_core_repeated_neighbours(gpu):
    # rows with a repeated id are never dense (sources or targets): the count
    # equals the count without a core, whatever repeats mean to the walk
    beg, adj = T.rmat_csr(12, 16, seed=2)
    beg = np.asarray(beg, np.int64)
    adj = np.asarray(adj, np.int32)
    deg = np.diff(beg)
    rows = np.argsort(-deg)[:40]  # repeat the first id of the longest rows
    nb, na = [0], []
    for u in range(len(beg) - 1):
        lst = list(adj[beg[u]:beg[u + 1]])
        if u in set(rows.tolist()) and lst:
            lst.append(lst[0])
        na += lst
        nb.append(len(na))
    nb, na = np.array(nb, np.int64), np.array(na, np.int32)
    print("[gpu] core with repeated neighbours", flush=True)
    off = T.count_csr(nb, na, device=gpu, flags=T.FLAG_SHORT_WALK, core_vertices=-1)
    on = T.count_csr(nb, na, device=gpu, flags=T.FLAG_SHORT_WALK, core_vertices=512,
                     core_min_degree=4)
    assert on.dense_edges > 0 and on.triangles == off.triangles
    assert on.probes + on.dense_probes == off.probes


def test_core_not_for_batches_or_ranks(gpu, rmat16):
    beg, adj, want, _ = rmat16
    with T.Context(gpu) as c:
        c.load(beg, adj)
        tot = 0
        for rank in range(3):
            r = c.count(rank=rank, total_process=3, flags=T.FLAG_SHORT_WALK, core_vertices=1024)
            assert r.core_vertices == 0
            tot += r.triangles
        assert tot == want


def test_core_params_rejected(ctx16):
    for kw in (dict(core_vertices=300), dict(core_vertices=128), dict(core_vertices=32768),
               dict(core_vertices=-2), dict(core_min_degree=1), dict(core_min_degree=-1)):
        with pytest.raises(T.TrustError):
            ctx16.count(**kw)


def test_core_closure_past_its_limit_keeps_the_seed(gpu):
    # vertex 0 points at 40000 vertices, each at the next: the closure of a
    # seed holding 0 passes 32768 vertices, so the seed alone is the core
    # (whose lists leave it: no dense targets), and the count stays exact:
    # one triangle (0, i, i + 1) per i
    n = 40000
    beg = np.zeros(n + 2, np.int64)
    beg[1] = n
    beg[2:] = n + np.minimum(np.arange(1, n + 1), n - 1)
    adj = np.concatenate([np.arange(1, n + 1), np.arange(2, n + 1)]).astype(np.int32)
    assert len(adj) == beg[-1]
    want = oracle.count_hash(beg, adj)
    assert want == n - 1
    print("[gpu] core closure past 32768 vertices", flush=True)
    r = T.count_csr(beg, adj, device=gpu, flags=T.FLAG_SHORT_WALK, core_vertices=256,
                    core_min_degree=2)
    assert r.triangles == want
    assert r.core_vertices <= 256 and r.dense_edges == 0
    assert r.probes + r.dense_probes == oracle.probes_short(beg, adj)


def test_core_closure_retried_from_a_halved_seed(gpu):
    # 200 vertices (d+ 1001, no in-edges) point at 1001 targets with lists of
    # two inside them; 100 vertices (d+ 600) point at 60000 others.  A seed of
    # 512 holds both groups and closes past 32768 vertices; halved to 256 it
    # holds the first group only, which closes to 1201 vertices (the seed
    # alone would be the core without the retry).  Dense
    # targets: the 999 targets with two entries; dense edges: 200 x 999 from
    # the first group and the 998 + 997 between targets
    K, T_, B, U = 200, 1001, 100, 60000
    t0, b0, u0 = K, K + T_, K + T_ + B
    rows = [list(range(t0, t0 + T_)) for _ in range(K)]
    rows += [[x for x in (t + 1, t + 2) if x < t0 + T_] for t in range(t0, t0 + T_)]
    rows += [list(range(u0 + 600 * i, u0 + 600 * (i + 1))) for i in range(B)]
    rows += [[] for _ in range(U)]
    beg = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    adj = np.array([x for r in rows for x in r], np.int32)
    want = oracle.count_hash(beg, adj)
    assert want == K * (2 * (T_ - 2) + 1) + (T_ - 2)
    print("[gpu] core closure retried from a halved seed", flush=True)
    r = T.count_csr(beg, adj, device=gpu, flags=T.FLAG_SHORT_WALK, core_vertices=512,
                    core_min_degree=2)
    assert r.triangles == want
    assert r.core_vertices == K + T_
    assert r.dense_edges == K * (T_ - 2) + (T_ - 3) + (T_ - 4)
    assert r.probes + r.dense_probes == oracle.probes_short(beg, adj)
