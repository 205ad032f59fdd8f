"""Pins the CPU oracle before anything is checked against it.

Each oracle function is held to values it cannot derive from its own code:
closed-form triangle counts, the O(n^3) brute force on the raw edge list,
networkx counts committed with the golden fixtures, and agreement between
the hashed transcription and the independent sort-merge counter.
"""
import numpy as np
import pytest

import oracle
import trust_amd as T
from conftest import CLOSED_FORMS, closed_form_csr, golden_cases, load_golden


@pytest.mark.parametrize("name", sorted(CLOSED_FORMS))
def test_closed_forms(name):
    beg, adj, want = closed_form_csr(name)
    assert oracle.count_hash(beg, adj) == want
    assert oracle.count_merge(beg, adj) == want


@pytest.mark.parametrize("name", sorted(CLOSED_FORMS))
def test_brute_force_closed_forms(name):
    e, want = CLOSED_FORMS[name]
    e = np.array(e, np.int32)
    n = int(e.max()) + 1
    assert oracle.count_brute(n, e[:, 0], e[:, 1]) == want


@pytest.mark.parametrize("name", golden_cases())
def test_golden_networkx(name):
    beg, adj, meta = load_golden(name)
    want = meta["triangles_networkx"]
    assert meta["triangles_brute"] == want
    assert oracle.count_hash(beg, adj) == want
    assert oracle.count_merge(beg, adj) == want


@pytest.mark.parametrize("seed", range(6))
def test_random_graphs_three_methods(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 400))
    m = int(rng.integers(n, n * 12))
    src = rng.integers(0, n, m).astype(np.int32)
    dst = rng.integers(0, n, m).astype(np.int32)
    brute = oracle.count_brute(n, src, dst)
    beg, adj = T.preprocess_edges(src, dst)
    assert oracle.count_hash(beg, adj) == brute
    assert oracle.count_merge(beg, adj) == brute


def test_hash_bucket_paths():
    # K150 has d+ up to 149 > 100: the 1024-bucket (CTA) branch; K50 only the
    # 32-bucket (warp) branch.  Both must agree with the closed form.
    beg, adj, want = closed_form_csr("K150")
    c, maxb = oracle.count_hash(beg, adj, return_max_bucket=True)
    assert c == want and maxb >= 1
    d = np.diff(beg)
    assert d.max() > 100


def test_thread_count_invariance():
    beg, adj = T.rmat_csr(12, 8, seed=3)
    counts = {oracle.count_hash(beg, adj, threads=t) for t in (1, 2, 5)}
    assert len(counts) == 1


def test_probes_closed_form_complete_graph():
    # Degree orientation of K_n is a transitive tournament (out-degrees
    # n-1..0); every probe hits, so W = C(n,3) exactly.
    for name, n in (("K10", 10), ("K50", 50), ("K150", 150)):
        beg, adj, want = closed_form_csr(name)
        assert oracle.probes(beg, adj) == n * (n - 1) * (n - 2) // 6 == want


def test_probes_per_vertex_sum_and_skip():
    beg, adj = T.rmat_csr(10, 8, seed=2)
    w, per = oracle.probes(beg, adj, per_vertex=True)
    assert per.sum() == w
    d = np.diff(beg)
    assert np.all(per[d < 2] == 0)  # vertices with d+ < 2 are skipped


def test_probes_short_closed_form_complete_graph():
    # in K_n's transitive tournament every edge u -> v has d+(v) < d+(u), so
    # the short side is always the reference's side: W is C(n,3) again
    for name, n in (("K10", 10), ("K50", 50)):
        beg, adj, want = closed_form_csr(name)
        assert oracle.probes_short(beg, adj) == want


def _explicit_short_walk(beg, adj, in_ties=False):
    """Brute force: build each vertex's two-part walk explicitly (out-part
    with d+(v) <= d+(u), in-part = lower-d+ in-neighbours with d+ >= 2), probe
    every listed N+(x) against a set, return (triangles, probes per vertex).
    in_ties=True plants a mistake: the in-part also takes equal-d+
    in-neighbours, so an edge between equal out-degrees is walked twice."""
    d = np.diff(beg)
    V = len(d)
    ins = [[] for _ in range(V)]
    for u in range(V):
        for v in adj[beg[u]:beg[u + 1]]:
            ins[v].append(u)
    tri, per = 0, np.zeros(V, np.int64)
    for o in range(V):
        if d[o] < 2:
            continue
        tab = set(adj[beg[o]:beg[o + 1]].tolist())
        walk = [v for v in adj[beg[o]:beg[o + 1]] if d[v] <= d[o]]
        walk += [w for w in ins[o] if 2 <= d[w] and (d[w] < d[o] or (in_ties and d[w] == d[o]))]
        for x in walk:
            lst = adj[beg[x]:beg[x + 1]]
            per[o] += len(lst)
            tri += sum(1 for y in lst.tolist() if y in tab)
    return tri, per


@pytest.mark.parametrize("seed", range(3))
def test_probes_short_against_explicit_walk(seed):
    beg, adj = T.rmat_csr(9, 8, seed=seed + 20)
    tri, per = _explicit_short_walk(beg, adj)
    assert tri == oracle.count_hash(beg, adj)  # the short-side walk finds every triangle once
    w, per_o = oracle.probes_short(beg, adj, per_vertex=True)
    assert np.array_equal(per_o, per) and w == per.sum()
    assert w < oracle.probes(beg, adj)  # R-MAT: strictly fewer probes than the reference walk


@pytest.mark.parametrize("seed", range(2))
def test_features_short_against_explicit_walk(seed):
    # sum = the vertex's short-side probes; max from the explicit walk rows
    beg, adj = T.rmat_csr(9, 8, seed=seed + 30)
    f = oracle.features_short(beg, adj)
    _, per = oracle.probes_short(beg, adj, per_vertex=True)
    assert np.array_equal(f[:, 1], per.astype(np.float64))
    d = np.diff(beg)
    assert np.array_equal(f[:, 0], d.astype(np.float64))
    ins = [[] for _ in range(len(d))]
    for u in range(len(d)):
        for v in adj[beg[u]:beg[u + 1]]:
            ins[v].append(u)
    for o in range(len(d)):
        walk = ([d[v] for v in adj[beg[o]:beg[o + 1]] if d[v] <= d[o]] if d[o] >= 2 else [])
        walk += [d[w] for w in ins[o] if 2 <= d[w] < d[o]]
        assert f[o, 2] == (max(walk) if walk else 0), o


def test_short_walk_tie_rule_mistake_is_caught():
    beg, adj = T.rmat_csr(9, 8, seed=21)
    bad, _ = _explicit_short_walk(beg, adj, in_ties=True)
    assert bad > oracle.count_hash(beg, adj)


def test_features_closed_form():
    # transitive tournament: the vertex with d+ = k points at vertices with
    # d+ = k-1..0, so sum = k(k-1)/2 and max = k-1.
    beg, adj, _ = closed_form_csr("K50")
    f = oracle.features(beg, adj)
    k = f[:, 0]
    assert sorted(k.astype(int).tolist()) == list(range(50))
    assert np.array_equal(f[:, 1], k * (k - 1) / 2)
    assert np.array_equal(f[:, 2], np.maximum(k - 1, 0))


def test_tile_jobs_sum_to_total():
    beg, adj = T.rmat_csr(11, 8, seed=4)
    total = oracle.count_hash(beg, adj)
    for n in (1, 2, 3, 4):
        tiles = T.partition_csr(beg, adj, n)
        assert oracle.count_partitioned(tiles, n) == total


def test_tile_job_planted_mistake_is_caught():
    # Swapping the walk and probe tiles (a plausible indexing slip) must
    # change the total, so the additivity test above can fail.
    beg, adj = T.rmat_csr(11, 8, seed=4)
    n = 3
    tiles = T.partition_csr(beg, adj, n)
    total = oracle.count_hash(beg, adj)
    wrong = 0
    for i in range(n):
        for j in range(n):
            for k in range(n):
                wrong += oracle.count_tile_job(tiles[i * n + j], tiles[k * n + j],
                                               tiles[i * n + k])
    assert wrong != total


def test_sample_counter_full_stride_is_total():
    beg, adj = T.rmat_csr(11, 8, seed=6)
    c, p = oracle.count_hash_sample(beg, adj, 0, 1)
    assert c == oracle.count_hash(beg, adj)
    assert p == oracle.probes(beg, adj)
    parts = [oracle.count_hash_sample(beg, adj, f, 4)[0] for f in range(4)]
    assert sum(parts) == c


def test_mlp_oracle_gradient_finite_differences():
    rng = np.random.default_rng(0)
    P = oracle.MLP_H ** 2 + 6 * oracle.MLP_H + 1
    p = rng.normal(0, 0.3, P)
    x = rng.uniform(0, 1, (40, 3))
    t = rng.uniform(0, 1, 40)
    g, loss = oracle.mlp_loss_grad(p, x, t)
    assert np.isclose(loss, ((oracle.mlp_forward(p, x) - t) ** 2).mean())
    eps = 1e-6
    for i in rng.choice(P, 60, replace=False):
        e = np.zeros(P)
        e[i] = eps
        num = (oracle.mlp_loss_grad(p + e, x, t)[1] - oracle.mlp_loss_grad(p - e, x, t)[1]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), i


def test_mlp_oracle_linear_regime_closed_form():
    # all weights zero: y = b3 for every x, so d loss / d b3 = 2 mean(b3 - t)
    # and every other gradient is zero (ReLU inputs are 0, not > 0)
    P = oracle.MLP_H ** 2 + 6 * oracle.MLP_H + 1
    p = np.zeros(P)
    p[-1] = 0.7
    x = np.random.default_rng(1).uniform(0, 1, (10, 3))
    t = np.linspace(0, 1, 10)
    g, loss = oracle.mlp_loss_grad(p, x, t)
    assert np.isclose(g[-1], 2 * np.mean(0.7 - t))
    assert np.all(g[:-1] == 0)
    assert np.isclose(loss, np.mean((0.7 - t) ** 2))
