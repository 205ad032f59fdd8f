"""GPU edge cases of the hash tables, each compared with the oracle.

* Natural cuckoo failure.  A vertex whose out-neighbours all fall into the
  same two buckets of its table (same bucket1 = id & mask, same bucket2 =
  the multiplicative hash of the id's low 24 bits; ids found by search) has
  more keys than the four slots of that bucket pair, so its cuckoo insertion
  fails and the table is cleared and rebuilt with linear probing.  Checked
  in the wave tier, the workgroup tier and the HBM-table tier, through
  trust_result.linear_rebuilds.
* Forced rebuild (TRUST_FLAG_REBUILD_LINEAR): every table takes that path.
* The reference's drop case: more than 100 out-neighbours with the same id
  mod 1024, i.e. in one bucket of the reference's CTA table, which keeps 100
  entries per bucket and drops the rest
  (Without-graph-partition/TC.cu:284-301).  The oracle reports the bucket's
  occupancy (> 100) and counts exactly, as must the GPU.
"""
import numpy as np
import pytest

import oracle
import trust_amd as T
from conftest import GOLDEN, golden_cases, load_golden

pytestmark = pytest.mark.gpu


def bucket1(key, lgb):
    return key & ((1 << lgb) - 1)


def bucket2(key, lgb):
    # the kernel's second bucket: low 32 bits of (key mod 2^24) * 0x9E3779, top lgb bits
    return (((key & 0xFFFFFF) * 0x9E3779) & 0xFFFFFFFF) >> (32 - lgb)


def colliding_keys(n, lgb, lo, hi, skip=()):
    """n ids in [lo, hi) sharing bucket1 and bucket2 (the two distinct)."""
    groups = {}
    for k in range(lo, hi):
        if k in skip:
            continue
        b = (bucket1(k, lgb), bucket2(k, lgb))
        if b[0] == b[1]:
            continue
        groups.setdefault(b, []).append(k)
        if len(groups[b]) == n:
            return groups[b]
    raise AssertionError("no colliding ids in range")


def csr_from_oriented(V, edges):
    e = np.array(sorted(set(edges)), np.int64).reshape(-1, 2)
    assert np.all(e[:, 0] != e[:, 1])
    beg = np.zeros(V + 1, np.int64)
    np.add.at(beg, e[:, 0] + 1, 1)
    beg = np.cumsum(beg)
    adj = e[:, 1].astype(np.int32)  # rows sorted by source, ids ascending within a row
    T.validate(beg, adj)
    return beg, adj


def cuckoo_failure_graph(hubs=3, keys=5):
    """`hubs` disjoint K_{keys+1}: hub h -> its keys, key_i -> key_j (i < j).
    Each hub's keys collide in both buckets of its 64-slot table."""
    edges, used = [], set(range(hubs))
    for h in range(hubs):
        ks = colliding_keys(keys, 5, 1000 + 50000 * h, 1000 + 50000 * (h + 1), used)
        used |= set(ks)
        edges += [(h, k) for k in ks]
        edges += [(ks[i], ks[j]) for i in range(keys) for j in range(i + 1, keys)]
    V = max(max(e) for e in edges) + 1
    return csr_from_oriented(V, edges), hubs * (keys + 1) * keys * (keys - 1) // 6


TIERS = {
    "wave": dict(),
    "workgroup": dict(hub_threshold=0),
    "hbm": dict(flags=T.FLAG_FORCE_SPILL),
}


@pytest.mark.parametrize("tier", sorted(TIERS))
@pytest.mark.parametrize("walk", [T.FLAG_REFERENCE_WALK, T.FLAG_SHORT_WALK])
def test_natural_cuckoo_failure_rebuilds_linear(gpu, tier, walk):
    (beg, adj), want = cuckoo_failure_graph()
    assert oracle.count_merge(beg, adj) == oracle.count_hash(beg, adj) == want
    kw = dict(TIERS[tier])
    kw["flags"] = kw.get("flags", 0) | walk
    print(f"[gpu] natural cuckoo failure: tier {tier}, walk flag {walk}", flush=True)
    r = T.count_csr(beg, adj, **kw)
    assert r.triangles == want
    assert r.linear_rebuilds == 3  # one per hub, no other table fails
    assert r.probes == (oracle.probes_short(beg, adj) if walk == T.FLAG_SHORT_WALK
                        else oracle.probes(beg, adj))


@pytest.mark.parametrize("tier", sorted(TIERS))
def test_forced_rebuild_rmat(gpu, tier):
    beg, adj = T.rmat_csr(14, 16, seed=3)
    want = oracle.count_hash(beg, adj)
    kw = dict(TIERS[tier])
    kw["flags"] = kw.get("flags", 0) | T.FLAG_REBUILD_LINEAR
    print(f"[gpu] forced linear rebuild, rmat14, tier {tier}", flush=True)
    with T.Context(gpu) as c:
        c.load(beg, adj)
        plain = c.count(**TIERS[tier])
        r = c.count(**kw)
        again = c.count(**kw)
    # (a few tables fail their cuckoo insertion on their own -- small tables
    # at load 1/2, concurrent inserts -- and take the same rebuild)
    print(f"[gpu] natural rebuilds {plain.linear_rebuilds} of {plain.n_spill + plain.n_hub + plain.n_wave}"
          f" tables", flush=True)
    assert plain.linear_rebuilds < (plain.n_spill + plain.n_hub + plain.n_wave) // 100
    assert r.triangles == again.triangles == plain.triangles == want
    # every hashed vertex took the rebuild path
    assert r.linear_rebuilds == again.linear_rebuilds == r.n_spill + r.n_hub + r.n_wave > 0


@pytest.mark.parametrize("name", golden_cases())
def test_forced_rebuild_golden(gpu, name):
    beg, adj, meta = load_golden(name)
    for kw in TIERS.values():
        flags = kw.get("flags", 0) | T.FLAG_REBUILD_LINEAR
        r = T.count_csr(beg, adj, **dict(kw, flags=flags))
        assert r.triangles == meta["triangles_networkx"], (name, kw)


def reference_drop_graph(d=150, c=5):
    """Hub 0 with d out-neighbours n_j = c + 1024 (j + 1): all in bucket c of
    the reference's 1024-bucket CTA table.  n_j -> n_{j+1} and n_j -> n_{j+2}."""
    n = [c + 1024 * (j + 1) for j in range(d)]
    edges = [(0, x) for x in n]
    edges += [(n[j], n[j + 1]) for j in range(d - 1)]
    edges += [(n[j], n[j + 2]) for j in range(d - 2)]
    return csr_from_oriented(n[-1] + 1, edges)


@pytest.mark.parametrize("kw", [dict(), dict(flags=T.FLAG_WAVE_ONLY), dict(hub_threshold=0),
                                dict(flags=T.FLAG_FORCE_SPILL), dict(flags=T.FLAG_REFERENCE_WALK),
                                dict(flags=T.FLAG_LINEAR_PROBING),
                                dict(threads=1024, blocks=1024, chunk=1)],
                         ids=lambda kw: ",".join(f"{k}={v}" for k, v in kw.items()) or "default")
def test_reference_bucket_overflow_case(gpu, kw):
    beg, adj = reference_drop_graph()
    want, max_bucket = oracle.count_hash(beg, adj, return_max_bucket=True)
    assert max_bucket == 150 > 100  # the reference would drop 50 entries of this bucket
    assert want == oracle.count_merge(beg, adj) == 149 + 148 + 148
    print(f"[gpu] reference drop case {kw}", flush=True)
    r = T.count_csr(beg, adj, **kw)
    assert r.triangles == want
