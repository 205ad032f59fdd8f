"""CPU tests of batches of graphs: the seeded ER/BA batch generator, the
batch layout (global ids, vertex ranges), and per-graph oracle counts."""
import numpy as np
import pytest

import oracle
import trust_amd as T
from batch_util import concat_csr
from conftest import closed_form_csr


def test_shape_ranges_and_mix():
    n, d, k = T.batch_shape(4000, seed=5, nmin=1000, nmax=200000, dmin=8, dmax=64)
    assert n.min() >= 1000 and n.max() <= 200000
    assert d.min() >= 8 and d.max() <= 64
    assert np.array_equal(k, np.arange(4000) % 2)
    # log-uniform n: log n is uniform, so its mean sits mid-way
    assert abs(np.log(n).mean() - (np.log(1000) + np.log(200000)) / 2) < 0.1
    assert abs(d.mean() - 36) < 1.0
    n2, d2, _ = T.batch_shape(4000, seed=5, nmin=1000, nmax=200000, dmin=8, dmax=64)
    assert np.array_equal(n, n2) and np.array_equal(d, d2)


def test_shape_rejects_bad_ranges():
    with pytest.raises(T.TrustError):
        T.batch_shape(10, nmin=100, nmax=50)
    with pytest.raises(T.TrustError):
        T.batch_shape(0)


def test_generate_layout_and_determinism():
    beg, adj, vb = T.batch_generate(60, seed=2, nmin=40, nmax=2000, dmin=4, dmax=30)
    T.validate(beg, adj)
    assert vb[0] == 0 and vb[-1] == len(beg) - 1 and np.all(np.diff(vb) > 0)
    # every edge stays inside its graph
    g_of = np.searchsorted(vb, np.arange(len(beg) - 1), side="right") - 1
    src = np.repeat(np.arange(len(beg) - 1), np.diff(beg))
    assert np.array_equal(g_of[src], g_of[adj])
    b2, a2, v2 = T.batch_generate(60, seed=2, nmin=40, nmax=2000, dmin=4, dmax=30)
    assert np.array_equal(beg, b2) and np.array_equal(adj, a2) and np.array_equal(vb, v2)


def test_generated_graphs_follow_the_input_contract():
    # each graph is preprocess_edges of its generator's pairs (ER: m = n d / 2
    # pairs; BA: n k pairs with k = round(d / 2)); deduplication only removes
    beg, adj, vb = T.batch_generate(6, seed=9, nmin=50, nmax=500, dmin=6, dmax=20)
    n, d, kind = T.batch_shape(6, seed=9, nmin=50, nmax=500, dmin=6, dmax=20)
    for g in range(6):
        b, a = T.batch_graph(beg, adj, vb, g)
        T.validate(b, a)
        assert len(b) - 1 <= n[g]
        if kind[g] == 0:
            assert len(a) <= round(n[g] * d[g] / 2)
        else:
            k = max(1, round(d[g] / 2))
            assert 0.5 * k * n[g] <= len(a) <= k * n[g]


def test_per_graph_counts_sum_to_whole():
    beg, adj, vb = T.batch_generate(40, seed=4, nmin=30, nmax=1500, dmin=4, dmax=40)
    per = [oracle.count_hash(*T.batch_graph(beg, adj, vb, g)) for g in range(40)]
    assert sum(per) == oracle.count_hash(beg, adj) == oracle.count_merge(beg, adj)
    # the batch counter of the oracle: graph by graph, each graph on its own
    assert oracle.count_hash_batch(beg, adj, vb).tolist() == per
    assert [oracle.count_merge(*T.batch_graph(beg, adj, vb, g)) for g in range(40)] == per


def test_oracle_batch_counter_closed_forms():
    names = ["K10", "W30", "K150", "C9", "F25", "K5_6_7", "K40_60"]
    graphs = [closed_form_csr(n) for n in names]
    parts = []
    for b, a, _ in graphs:
        parts += [(b, a), (np.zeros(1, np.int64), np.zeros(0, np.int32))]
    beg, adj, vb = concat_csr(parts)
    want = []
    for _, _, w in graphs:
        want += [w, 0]
    assert oracle.count_hash_batch(beg, adj, vb).tolist() == want


def test_concat_closed_forms():
    graphs = [closed_form_csr(n) for n in ("K10", "W30", "K150", "C9", "F25")]
    beg, adj, vb = concat_csr([(b, a) for b, a, _ in graphs])
    T.validate(beg, adj)
    for g, (_, _, want) in enumerate(graphs):
        assert oracle.count_hash(*T.batch_graph(beg, adj, vb, g)) == want
