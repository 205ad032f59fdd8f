"""GPU parity of batched counting: one persistent count over a batch of
graphs must give every graph exactly its own oracle count (uint64), under
every schedule and launch shape, eagerly and as replays of a captured
hipGraph."""
import numpy as np
import pytest

import oracle
import trust_amd as T
from batch_util import concat_csr
from conftest import closed_form_csr

pytestmark = pytest.mark.gpu


def per_graph_oracle(beg, adj, vb):
    return np.array([oracle.count_hash(*T.batch_graph(beg, adj, vb, g))
                     for g in range(len(vb) - 1)], np.uint64)


@pytest.fixture(scope="module")
def batch():
    # 300 ER/BA graphs, n log-uniform in [50, 3000], mean degree in [4, 48]
    beg, adj, vb = T.batch_generate(300, seed=11, nmin=50, nmax=3000, dmin=4, dmax=48)
    return beg, adj, vb, per_graph_oracle(beg, adj, vb)


@pytest.fixture(scope="module")
def bctx(gpu, batch):
    beg, adj, vb, _ = batch
    c = T.Context(gpu)
    c.load_batch(beg, adj, vb)
    yield c
    c.close()


@pytest.mark.parametrize("kw", [
    {}, dict(threads=64), dict(threads=1024), dict(chunk=1), dict(chunk=7),
    dict(flags=T.FLAG_FORCE_SPILL), dict(flags=T.FLAG_WAVE_ONLY), dict(hub_threshold=0),
    dict(flags=T.FLAG_LINEAR_PROBING), dict(flags=T.FLAG_OFFSETS64),
    dict(blocks=7), dict(flags=T.FLAG_SHORT_WALK), dict(flags=T.FLAG_SHORT_WALK, threads=64),
    dict(flags=T.FLAG_REFERENCE_WALK),
], ids=lambda kw: ",".join(f"{k}={v}" for k, v in kw.items()) or "default")
def test_batch_counts_match_oracle(gpu, bctx, batch, kw):
    beg, adj, vb, want = batch
    print(f"[gpu] batch 300 graphs {kw}", flush=True)
    if kw.get("flags") == T.FLAG_OFFSETS64:  # needs a reload with int64 offsets
        with T.Context(gpu) as c:
            c.load_batch(beg, adj, vb, flags=T.FLAG_OFFSETS64)
            counts, done, r = c.count_batch()
    else:
        counts, done, r = bctx.count_batch(**kw)
    assert np.array_equal(counts, want)
    assert r.triangles == int(want.sum())
    assert np.all(done >= 0) and np.all(done[want > 0] > 0)
    # every graph completes inside the counting launches (clock slack 1 ms)
    assert done.max() <= r.seconds + 1e-3


def test_batch_rank_split(bctx, batch):
    _, _, _, want = batch
    parts = [bctx.count_batch(rank=r, total_process=3)[0] for r in range(3)]
    assert np.array_equal(parts[0] + parts[1] + parts[2], want)


def test_batch_learned_schedule(bctx, batch):
    _, _, _, want = batch
    bctx.train_cost_model(epochs=10, max_samples=50000, seed=3)
    counts, _, r = bctx.count_batch(schedule=T.SCHED_LEARNED)
    assert r.schedule_used == T.SCHED_LEARNED
    assert np.array_equal(counts, want)


def test_hip_graph_replays_match_eager(bctx, batch):
    _, _, _, want = batch
    eager, _, _ = bctx.count_batch()
    assert np.array_equal(eager, want)
    for i in range(6):
        counts, done, r = bctx.count_batch(flags=T.FLAG_HIP_GRAPH)
        print(f"[gpu] hipGraph replay {i}: {r.seconds * 1e3:.3f} ms", flush=True)
        assert np.array_equal(counts, want), i
        assert r.triangles == int(want.sum())
        # eager work and a different count between replays
        if i % 2 == 0:
            assert np.array_equal(bctx.count_batch(threads=512)[0], want)
            assert bctx.count().triangles == int(want.sum())
    # a parameter change re-captures; the learned schedule and the short-side
    # walk (its walk rows built before capture) replay too
    for kw in (dict(threads=1024), dict(schedule=T.SCHED_LEARNED),
               dict(flags=T.FLAG_SHORT_WALK), {}):
        rest = {k: v for k, v in kw.items() if k != "flags"}
        for _ in range(3):
            counts, _, _ = bctx.count_batch(flags=T.FLAG_HIP_GRAPH | kw.get("flags", 0), **rest)
            assert np.array_equal(counts, want), kw


def test_closed_form_batch_with_empty_graphs(gpu):
    names = ["K10", "W30", "K150", "C9", "F25", "K5_6_7", "K3", "K40_60", "K50"]
    parts, want = [], []
    for n in names:
        b, a, w = closed_form_csr(n)
        parts.append((b, a))
        want.append(w)
        parts.append((np.zeros(1, np.int64), np.zeros(0, np.int32)))  # empty graph
        want.append(0)
    parts.append((np.zeros(6, np.int64), np.zeros(0, np.int32)))  # 5 isolated vertices
    want.append(0)
    beg, adj, vb = concat_csr(parts)
    with T.Context(gpu) as c:
        c.load_batch(beg, adj, vb)
        for kw in ({}, dict(flags=T.FLAG_FORCE_SPILL), dict(threads=1024),
                   dict(flags=T.FLAG_HIP_GRAPH)):
            counts, done, _ = c.count_batch(**kw)
            assert counts.tolist() == want, kw
            assert np.all(done[np.array(want) == 0] >= 0)


def test_many_tiny_graphs_page_table(gpu):
    # 5000 graphs of 3..40 vertices: many graphs per 4096-vertex page
    rng = np.random.default_rng(8)
    parts = []
    for _ in range(5000):
        n = int(rng.integers(3, 40))
        m = int(rng.integers(1, 4 * n))
        parts.append(T.preprocess_edges(rng.integers(0, n, m).astype(np.int32),
                                        rng.integers(0, n, m).astype(np.int32)))
    beg, adj, vb = concat_csr(parts)
    want = per_graph_oracle(beg, adj, vb)
    with T.Context(gpu) as c:
        c.load_batch(beg, adj, vb)
        for kw in ({}, dict(chunk=64), dict(flags=T.FLAG_HIP_GRAPH)):
            assert np.array_equal(c.count_batch(**kw)[0], want), kw


def test_single_graph_batch_equals_count(gpu):
    beg, adj = T.rmat_csr(14, 16, seed=2)
    with T.Context(gpu) as c:
        c.load_batch(beg, adj, np.array([0, len(beg) - 1], np.int64))
        counts, done, r = c.count_batch()
        assert counts[0] == r.triangles == c.count().triangles == oracle.count_hash(beg, adj)


def test_load_batch_rejects_bad_layout(gpu):
    beg, adj, vb = T.batch_generate(5, seed=1, nmin=30, nmax=60, dmin=4, dmax=8)
    with T.Context(gpu) as c:
        bad = vb.copy()
        bad[[1, 2]] = bad[[2, 1]]
        with pytest.raises(T.TrustError):
            c.load_batch(beg, adj, bad)  # decreasing
        with pytest.raises(T.TrustError):
            c.load_batch(beg, adj, vb[:-1].copy())  # does not end at V
        cross = adj.copy()
        cross[0] = int(vb[3])  # an edge of graph 0 into graph 3
        with pytest.raises(T.TrustError):
            c.load_batch(beg, cross, vb)
        with pytest.raises(T.TrustError):
            c.count_batch()  # nothing loaded


@pytest.mark.parametrize("schedule", [T.SCHED_DEGREE, T.SCHED_LEARNED])
def test_graphs_complete_progressively(gpu, schedule):
    # graphs are drained in vertex-range order (the learned queue by
    # superblocks, longest-first inside each), so half of the graphs finish
    # well before the last one
    beg, adj, vb = T.batch_generate(2000, seed=5, nmin=300, nmax=3000, dmin=8, dmax=32)
    with T.Context(gpu) as c:
        c.load_batch(beg, adj, vb)
        if schedule == T.SCHED_LEARNED:
            c.train_cost_model(epochs=5, max_samples=50000, seed=1)
        for _ in range(2):
            counts, done, r = c.count_batch(schedule=schedule, flags=T.FLAG_HIP_GRAPH)
        assert r.schedule_used == schedule
        p50, p100 = np.percentile(done, 50), done.max()
        print(f"[gpu] schedule {schedule}: p50 {p50 * 1e3:.3f} ms, max {p100 * 1e3:.3f} ms", flush=True)
        assert p50 < 0.8 * p100


def greedy_segments(beg, vb, max_edges):
    """Segments of whole graphs with at most max_edges edges each, cut
    greedily in graph order; 0 when some graph alone exceeds the limit."""
    eb = np.asarray(beg)[np.asarray(vb)]
    ng, n, g = len(vb) - 1, 0, 0
    max_edges = min(max_edges, (1 << 30) - 1)
    if max_edges <= 0:
        return 0
    while g < ng:
        if eb[g + 1] - eb[g] > max_edges:
            return 0
        g0 = g
        while g < ng and eb[g + 1] - eb[g0] <= max_edges:
            g += 1
        n += vb[g] > vb[g0]
    return n


@pytest.fixture(scope="module")
def sctx(gpu, batch):
    beg, adj, vb, _ = batch
    c = T.Context(gpu)
    c.load_batch(beg, adj, vb, flags=T.FLAG_OFFSETS64)
    yield c
    c.close()


@pytest.mark.parametrize("limit", ["below_largest", "largest", "3x_largest", "all", "cap", "off"])
def test_int32_segments_match_oracle(sctx, batch, limit):
    # an int64 batch counted as int32 segments of whole graphs must give every
    # graph its own count, whatever the cut
    beg, adj, vb, want = batch
    e = np.diff(beg[vb])
    max_edges = {"below_largest": int(e.max()) - 1, "largest": int(e.max()),
                 "3x_largest": 3 * int(e.max()), "all": len(adj), "cap": 1 << 40, "off": 0}[limit]
    nseg = sctx.batch_segments(max_edges)
    print(f"[gpu] int64 batch, segments of <= {max_edges} edges: {nseg}", flush=True)
    assert nseg == greedy_segments(beg, vb, max_edges)
    if limit in ("all", "cap"):
        assert nseg == 1
    if limit == "largest":
        assert nseg > 10
    for kw in ({}, dict(threads=1024), dict(flags=T.FLAG_FORCE_SPILL), dict(chunk=1),
               dict(flags=T.FLAG_WAVE_ONLY), dict(flags=T.FLAG_SHORT_WALK),
               dict(flags=T.FLAG_HIP_GRAPH)):
        counts, done, r = sctx.count_batch(**kw)
        assert np.array_equal(counts, want), kw
        assert r.triangles == int(want.sum()), kw
        assert np.all(done[want > 0] > 0) and done.max() <= r.seconds + 1e-3, kw
    for i in range(3):  # replays with eager work between them
        assert np.array_equal(sctx.count_batch(flags=T.FLAG_HIP_GRAPH)[0], want), i
        r = sctx.count()
        assert r.triangles == int(want.sum())
    # tier sizes add up over segments: every vertex with d+ >= 2 once
    assert r.n_spill + r.n_hub + r.n_wave == int(np.sum(np.diff(beg) >= 2))
    assert sctx.gather().probes == r.probes  # the gather-only pass walks the same lists
    assert sctx.count(rank=1, total_process=2).triangles + \
        sctx.count(rank=0, total_process=2).triangles == int(want.sum())


def test_segments_with_empty_graphs(gpu):
    names = ["K10", "W30", "K150", "C9", "F25", "K50"]
    parts, want = [], []
    for n in names:
        b, a, w = closed_form_csr(n)
        parts += [(b, a), (np.zeros(1, np.int64), np.zeros(0, np.int32)),
                  (np.zeros(4, np.int64), np.zeros(0, np.int32))]
        want += [w, 0, 0]
    beg, adj, vb = concat_csr(parts)
    with T.Context(gpu) as c:
        c.load_batch(beg, adj, vb, flags=T.FLAG_OFFSETS64)
        for max_edges in (int(np.diff(beg[vb]).max()), 2 * int(np.diff(beg[vb]).max()), 1 << 30):
            nseg = c.batch_segments(max_edges)
            assert nseg == greedy_segments(beg, vb, max_edges) >= 1
            for kw in ({}, dict(flags=T.FLAG_HIP_GRAPH)):
                assert c.count_batch(**kw)[0].tolist() == want, (max_edges, kw)


def test_batch_segments_rules(gpu):
    beg, adj, vb = T.batch_generate(20, seed=2, nmin=50, nmax=200, dmin=4, dmax=8)
    with T.Context(gpu) as c:
        with pytest.raises(T.TrustError):
            c.batch_segments(1000)  # no batch loaded
        c.load_batch(beg, adj, vb)  # int32 offsets: nothing to cut
        assert c.batch_segments(1000) == 0
        c.load_batch(beg, adj, vb, flags=T.FLAG_OFFSETS64)  # the flag alone does not cut
        assert c.batch_segments(len(adj)) == 1
        with pytest.raises(T.TrustError):
            c.batch_segments(-1)
        c.load(beg, adj)  # a plain load ends the batch
        with pytest.raises(T.TrustError):
            c.batch_segments(1000)


def test_segments_offset_lookup_flag(sctx, batch):
    # batches always look the offsets up; the flag is accepted and changes nothing
    beg, adj, vb, want = batch
    sctx.batch_segments(int(np.diff(beg[vb]).max()))
    for kw in (dict(flags=T.FLAG_OFFSET_LOOKUP), dict(flags=T.FLAG_HIP_GRAPH | T.FLAG_OFFSET_LOOKUP)):
        counts, _, r = sctx.count_batch(**kw)
        assert np.array_equal(counts, want), kw


def test_segments_beside_learned_schedule(sctx, batch):
    # each int32 segment gets its own plan, learned or degree split, and the
    # two schedules alternate, eagerly and captured
    beg, adj, vb, want = batch
    assert sctx.batch_segments(int(np.diff(beg[vb]).max())) > 1
    sctx.train_cost_model(epochs=3, max_samples=20000, seed=2)
    for kw in (dict(schedule=T.SCHED_LEARNED), dict(schedule=T.SCHED_DEGREE),
               dict(schedule=T.SCHED_LEARNED, flags=T.FLAG_HIP_GRAPH),
               dict(schedule=T.SCHED_DEGREE, flags=T.FLAG_HIP_GRAPH),
               dict(schedule=T.SCHED_LEARNED, flags=T.FLAG_HIP_GRAPH),
               dict(schedule=T.SCHED_LEARNED, threads=1024),
               dict(schedule=T.SCHED_LEARNED, flags=T.FLAG_FORCE_SPILL)):
        counts, _, r = sctx.count_batch(**kw)
        assert np.array_equal(counts, want), kw
        assert r.schedule_used == kw["schedule"]
        # every segment's plan lists each of its vertices with d+ >= 2 once
        assert r.n_spill + r.n_hub + r.n_wave == int(np.sum(np.diff(beg) >= 2)), kw
    r = sctx.count(schedule=T.SCHED_LEARNED)
    assert r.triangles == int(want.sum()) and r.schedule_used == T.SCHED_LEARNED
