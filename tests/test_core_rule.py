"""The dense core's exactness rule, on the CPU (trust_amd/csrc/core.hip).

The GPU counts an edge u -> v from a dense source to a dense target as
popcount(B(u) & B(v)) over bit rows of a core C and hashes every other edge.
That is exact for ANY core C as long as a dense target's whole out-list lies
in C (and rows have no repeats): every common out-neighbour of u and v is
then in C.  These tests split the oracle's hashed count by that rule for
arbitrary cores (random, the top degrees, a closure) on small graphs and
check the parts sum to the oracle's count; a planted mistake -- targets
chosen without the out-list condition -- must lose triangles.
"""
import numpy as np
import pytest

import oracle
import trust_amd as T


def rows_of(beg, adj):
    return [set(adj[beg[u]:beg[u + 1]].tolist()) for u in range(len(beg) - 1)]


def split_count(beg, adj, core, t, require_closed=True):
    """(dense part, hashed part) of the count under the core rule."""
    rows = rows_of(beg, adj)
    d = np.diff(beg)
    in_core = np.zeros(len(beg) - 1, bool)
    in_core[list(core)] = True
    target = [d[v] >= t and in_core[v] and (not require_closed or all(in_core[w] for w in rows[v]))
              for v in range(len(rows))]
    dense = hashed = 0
    for u, ru in enumerate(rows):
        bu = {w for w in ru if in_core[w]}
        for v in ru:
            if d[u] >= t and target[v]:
                bv = {w for w in rows[v] if in_core[w]}
                dense += len(bu & bv)  # popcount(B(u) & B(v))
            else:
                hashed += len(ru & rows[v])
    return dense, hashed


@pytest.fixture(scope="module")
def graphs():
    out = []
    for s, seed in ((9, 1), (10, 3), (11, 5)):
        beg, adj = T.rmat_csr(s, 16, seed=seed)
        out.append((np.asarray(beg, np.int64), np.asarray(adj, np.int64),
                    oracle.count_hash(beg, adj)))
    return out


def closure(beg, adj, seed):
    core = set(seed)
    todo = list(seed)
    while todo:
        v = todo.pop()
        for w in adj[beg[v]:beg[v + 1]].tolist():
            if w not in core:
                core.add(w)
                todo.append(w)
    return core


@pytest.mark.parametrize("kind", ["random", "top", "closed"])
@pytest.mark.parametrize("t", [2, 8])
def test_parts_sum_to_the_oracle(graphs, kind, t):
    rng = np.random.default_rng(7)
    for beg, adj, want in graphs:
        V = len(beg) - 1
        deg = np.diff(beg) + np.bincount(adj, minlength=V)
        if kind == "random":
            core = set(rng.choice(V, size=V // 8, replace=False).tolist())
        elif kind == "top":
            core = set(np.argsort(-deg, kind="stable")[:V // 16].tolist())
        else:
            core = closure(beg, adj, np.argsort(-deg, kind="stable")[:V // 64].tolist())
        dense, hashed = split_count(beg, adj, core, t)
        assert dense + hashed == want, (kind, t)
        if kind != "random" and t == 2:
            assert dense > 0  # the rule takes edges on these graphs


def test_planted_mistake_targets_without_the_out_list_condition(graphs):
    # a "target" whose out-list leaves the core loses the triangles closed
    # outside it: the parts no longer sum to the count
    lost = False
    rng = np.random.default_rng(11)
    for beg, adj, want in graphs:
        V = len(beg) - 1
        core = set(rng.choice(V, size=V // 4, replace=False).tolist())
        dense, hashed = split_count(beg, adj, core, 2, require_closed=False)
        lost |= dense + hashed < want
    assert lost
