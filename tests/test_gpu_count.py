"""GPU parity: the HIP counting path against the CPU oracle (bit-exact uint64).

Every count here goes through the C ABI into the gfx950 kernels; the oracle
is the checker only.  The count must not depend on any launch or schedule
parameter, so the sweeps below double as invariance tests of the hash-table
sizing, the tier split, the queues and the logical-rank split.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import trust_amd as T
from conftest import (CLOSED_FORMS, GOLDEN, closed_form_csr, expected_probes, golden_cases,
                      load_golden)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat16():
    # R-MAT scale 16, edge factor 16, (a,b,c,d) = (.57,.19,.19,.05), seed 1
    beg, adj = T.rmat_csr(16, 16, seed=1)
    want = oracle.count_hash(beg, adj)
    assert want == oracle.count_merge(beg, adj)
    # probes of the walk a default count takes (the rule in expected_probes)
    return beg, adj, want, expected_probes(beg, adj)


@pytest.fixture(scope="module")
def ctx16(gpu, rmat16):
    beg, adj, want, w = rmat16
    c = T.Context(gpu)
    c.load(beg, adj)
    yield c
    c.close()


@pytest.mark.parametrize("name", sorted(CLOSED_FORMS))
def test_closed_forms(gpu, name):
    beg, adj, want = closed_form_csr(name)
    print(f"[gpu] closed form {name}", flush=True)
    assert T.count_csr(beg, adj, device=gpu).triangles == want
    assert T.count_csr(beg, adj, device=gpu, flags=T.FLAG_FORCE_SPILL).triangles == want


@pytest.mark.parametrize("name", golden_cases())
def test_golden(gpu, name):
    beg, adj, meta = load_golden(name)
    print(f"[gpu] golden {name}", flush=True)
    for fl in (0, T.FLAG_REFERENCE_WALK, T.FLAG_SHORT_WALK):
        r = T.count_csr(beg, adj, device=gpu, flags=fl)
        assert r.triangles == meta["triangles_networkx"]
        assert r.probes == expected_probes(beg, adj, fl)
        short = oracle.probes_short(beg, adj) <= 0.45 * oracle.probes(beg, adj)
        want_walk = 1 if fl == T.FLAG_SHORT_WALK else 0 if fl == T.FLAG_REFERENCE_WALK else int(short)
        assert r.walk_used == want_walk


def test_rmat16_default(ctx16, rmat16):
    beg, adj, want, w = rmat16
    r = ctx16.count()
    print(f"[gpu] rmat16 tiers {r.n_spill}/{r.n_hub}/{r.n_wave} t={r.seconds:.6f}s", flush=True)
    assert r.triangles == want
    assert r.probes == w
    assert r.n_hub > 0 and r.n_wave > 0
    assert r.n_hub + r.n_wave + r.n_spill == int((np.diff(beg) >= 2).sum())


@pytest.mark.parametrize("threads", [64, 128, 256, 512, 1024])
def test_rmat16_threads(ctx16, rmat16, threads):
    assert ctx16.count(threads=threads).triangles == rmat16[2]


@pytest.mark.parametrize("kw", [
    dict(threads=1024, blocks=1024, chunk=1),       # the reference README launch
    dict(blocks=1), dict(blocks=7, chunk=3), dict(chunk=64),
    dict(hub_threshold=0), dict(hub_threshold=32), dict(hub_threshold=1 << 30),
    dict(wave_slots=64, threads=256), dict(wave_slots=2048, threads=256),
    dict(lds_budget=4096, threads=64, wave_slots=256),
    dict(flags=T.FLAG_FORCE_SPILL), dict(flags=T.FLAG_WAVE_ONLY),
    dict(flags=T.FLAG_FORCE_SPILL, threads=64, blocks=3),
    dict(flags=T.FLAG_LINEAR_PROBING), dict(flags=T.FLAG_LINEAR_PROBING | T.FLAG_FORCE_SPILL),
    dict(flags=T.FLAG_LINEAR_PROBING, hub_threshold=0),
    dict(flags=T.FLAG_REFERENCE_WALK), dict(flags=T.FLAG_REFERENCE_WALK | T.FLAG_FORCE_SPILL),
    dict(flags=T.FLAG_SHORT_WALK), dict(flags=T.FLAG_SHORT_WALK | T.FLAG_FORCE_SPILL),
    dict(flags=T.FLAG_SHORT_WALK, threads=1024, blocks=1024, chunk=1),
    dict(flags=T.FLAG_SHORT_WALK, threads=64, blocks=7, chunk=3),
    dict(flags=T.FLAG_SHORT_WALK, hub_threshold=0), dict(flags=T.FLAG_SHORT_WALK, hub_threshold=1 << 30),
    dict(flags=T.FLAG_SHORT_WALK | T.FLAG_WAVE_ONLY, wave_slots=2048, threads=256),
    dict(flags=T.FLAG_SHORT_WALK | T.FLAG_LINEAR_PROBING),
    dict(flags=T.FLAG_SHORT_WALK, lds_budget=4096, threads=64, wave_slots=256),
])
def test_rmat16_invariance(ctx16, rmat16, kw):
    print(f"[gpu] rmat16 {kw}", flush=True)
    r = ctx16.count(**kw)
    assert r.triangles == rmat16[2]
    assert r.probes == expected_probes(rmat16[0], rmat16[1], kw.get("flags", 0))


def test_short_side_walk_saves_probes(ctx16, rmat16):
    beg, adj, want, _ = rmat16
    ref = ctx16.count(flags=T.FLAG_REFERENCE_WALK)
    short = ctx16.count(flags=T.FLAG_SHORT_WALK)
    again = ctx16.count(flags=T.FLAG_SHORT_WALK)
    print(f"[gpu] rmat16 probes reference {ref.probes} short-side {short.probes}; "
          f"count {ref.count_seconds * 1e3:.3f} ms -> {short.count_seconds * 1e3:.3f} ms, "
          f"index {short.index_seconds * 1e3:.3f} ms", flush=True)
    assert ref.triangles == short.triangles == again.triangles == want
    assert short.probes == oracle.probes_short(beg, adj) < ref.probes == oracle.probes(beg, adj)
    assert short.walk_used == 1 and ref.walk_used == 0
    assert again.index_seconds == 0.0  # the in-walk rows are built once per graph


@pytest.mark.parametrize("walk", [T.FLAG_SHORT_WALK, T.FLAG_REFERENCE_WALK, 0])
def test_fresh_index_is_timed(ctx16, rmat16, walk):
    # FLAG_FRESH_INDEX rebuilds the walk's per-graph index inside the count's
    # seconds (index + plan + count); without it a context reuses the index
    beg, adj, want, _ = rmat16
    kept = ctx16.count(flags=walk)
    f1 = ctx16.count(flags=walk | T.FLAG_FRESH_INDEX)
    f2 = ctx16.count(flags=walk | T.FLAG_FRESH_INDEX)
    again = ctx16.count(flags=walk)
    print(f"[gpu] rmat16 walk flag {walk}: fresh {f1.seconds * 1e3:.3f} ms (index "
          f"{f1.index_seconds * 1e3:.3f} ms), kept {again.seconds * 1e3:.3f} ms", flush=True)
    for r in (kept, f1, f2, again):
        assert r.triangles == want
        assert r.seconds == pytest.approx(r.index_seconds + r.plan_seconds + r.count_seconds,
                                          rel=1e-9, abs=1e-12)
    assert f1.index_seconds > 0 and f2.index_seconds > 0
    assert again.index_seconds == 0
    assert f1.walk_used == again.walk_used


def test_walk_flags_together_rejected(ctx16):
    with pytest.raises(T.TrustError):
        ctx16.count(flags=T.FLAG_REFERENCE_WALK | T.FLAG_SHORT_WALK)
    with pytest.raises(T.TrustError):
        T.count_tiles(T.partition_csr(*T.rmat_csr(8, 8, seed=1), 2), 2,
                      flags=T.FLAG_REFERENCE_WALK | T.FLAG_SHORT_WALK)


def test_default_walk_choice_rmat21(gpu):
    # R-MAT 21: the short side makes ~42% of the reference walk's probes, so a
    # default count takes it; the reference walk must find the same triangles
    beg, adj = T.rmat_csr(21, 16, seed=1)
    ref_w, short_w = oracle.probes(beg, adj), oracle.probes_short(beg, adj)
    assert short_w <= 0.45 * ref_w
    with T.Context(gpu) as c:
        c.load(beg, adj)
        auto = c.count()
        ref = c.count(flags=T.FLAG_REFERENCE_WALK)
        print(f"[gpu] rmat21 auto walk {auto.walk_used}: {auto.count_seconds * 1e3:.2f} ms vs "
              f"reference {ref.count_seconds * 1e3:.2f} ms, index {auto.index_seconds * 1e3:.1f} ms",
              flush=True)
        # (the automatic dense core takes some of the short walk's probes)
        assert auto.walk_used == 1 and auto.probes + auto.dense_probes == short_w
        assert ref.walk_used == 0 and ref.probes == ref_w
        assert auto.triangles == ref.triangles
        gauto, gref = c.gather(), c.gather(flags=T.FLAG_REFERENCE_WALK)
        assert gauto.probes + gauto.dense_probes == short_w and gref.probes == ref_w


def test_unknown_flags_rejected(ctx16):
    for bad in (32, 1 << 12, 1 << 31):
        with pytest.raises(T.TrustError):
            ctx16.count(flags=bad)


def test_rmat16_offsets64(gpu, rmat16):
    beg, adj, want, w = rmat16
    with T.Context(gpu) as c:
        c.load(beg, adj, flags=T.FLAG_OFFSETS64)
        assert c.count().triangles == want
        assert c.count(flags=T.FLAG_FORCE_SPILL).triangles == want
        # int64 walk rows: build, count, gather and walk features
        r = c.count(flags=T.FLAG_SHORT_WALK)
        assert r.triangles == want and r.probes == oracle.probes_short(beg, adj)
        assert c.gather(flags=T.FLAG_SHORT_WALK).probes == r.probes
        assert np.allclose(c.walk_features(), oracle.mlp_scaled_features_short(beg, adj),
                           rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("P", [2, 3, 8])
def test_logical_ranks_sum(ctx16, rmat16, P):
    parts = [ctx16.count(rank=r, total_process=P, chunk=2).triangles for r in range(P)]
    assert sum(parts) == rmat16[2]
    assert min(parts) > 0


def test_repeat_is_deterministic(ctx16, rmat16):
    counts = {ctx16.count().triangles for _ in range(5)}
    assert counts == {rmat16[2]}


def test_instrumented_probes_match_oracle(ctx16, rmat16):
    beg, adj, want, w = rmat16
    for fl in (0, T.FLAG_SHORT_WALK, T.FLAG_REFERENCE_WALK):
        cyc, prb, r = ctx16.profile_vertices(flags=fl)
        assert r.triangles == want
        _, per = expected_probes(beg, adj, fl, per_vertex=True)
        assert np.array_equal(prb.astype(np.int64), per)
    d = np.diff(beg)
    assert np.all(cyc[d >= 2] > 0) and np.all(cyc[d < 2] == 0)


@pytest.mark.parametrize("seed", range(3))
def test_random_dense_graphs(gpu, seed):
    # dense random graphs: many hub-tier vertices and long probe lists
    rng = np.random.default_rng(seed)
    n = int(rng.integers(300, 900))
    m = int(rng.integers(20 * n, 60 * n))
    src = rng.integers(0, n, m).astype(np.int32)
    dst = rng.integers(0, n, m).astype(np.int32)
    beg, adj = T.preprocess_edges(src, dst)
    want = oracle.count_hash(beg, adj)
    for kw in ({}, dict(flags=T.FLAG_FORCE_SPILL), dict(threads=64), dict(hub_threshold=8),
               dict(flags=T.FLAG_LINEAR_PROBING)):
        assert T.count_csr(beg, adj, device=gpu, **kw).triangles == want


def test_big_hub_spills_to_hbm(gpu):
    # K_400 has d+ up to 399: with a 4 KB LDS budget the biggest tables do
    # not fit and take the HBM tier without being asked to.
    beg, adj, _ = closed_form_csr("K150")
    e = np.array([(i, j) for i in range(400) for j in range(i + 1, 400)], np.int32)
    beg, adj = T.preprocess_edges(e[:, 0], e[:, 1])
    r = T.count_csr(beg, adj, device=gpu, lds_budget=4096, threads=64, wave_slots=64)
    assert r.triangles == 400 * 399 * 398 // 6
    assert r.n_spill > 0


def test_unsorted_general_csr(gpu):
    # any valid oriented CSR, not only the reference's class order: reverse
    # the vertex order and shuffle every list
    beg, adj = T.rmat_csr(12, 16, seed=5)
    want = oracle.count_hash(beg, adj)
    V = len(beg) - 1
    rng = np.random.default_rng(0)
    perm = rng.permutation(V)            # new id of old vertex
    inv = np.argsort(perm)
    d = np.diff(beg)
    nb = np.zeros(V + 1, np.int64)
    nb[1:] = np.cumsum(d[inv])
    na = np.empty_like(adj)
    for new in range(V):
        old = inv[new]
        lst = perm[adj[beg[old]:beg[old + 1]]]
        rng.shuffle(lst)
        na[nb[new]:nb[new + 1]] = lst
    assert T.count_csr(nb, na, device=gpu).triangles == want


@pytest.mark.parametrize("name", golden_cases())
def test_golden_other_input_tools(gpu, name):
    # the CSRs written by the reference's otherPreprocess and speedupIO tools
    # (other row orders, self-loops weighed once): same triangles
    _, _, meta = load_golden(name)
    for sub in ("plain", "indegree", "speedup"):
        d = os.path.join(GOLDEN, name, sub)
        beg = np.fromfile(os.path.join(d, "begin.bin"), np.int64)
        adj = np.fromfile(os.path.join(d, "adjacent.bin"), np.int32)
        print(f"[gpu] golden {name}/{sub}", flush=True)
        for fl in (0, T.FLAG_SHORT_WALK):
            r = T.count_csr(beg, adj, device=gpu, flags=fl)
            assert r.triangles == meta["triangles_networkx"]
            assert r.probes == expected_probes(beg, adj, fl)


@pytest.mark.parametrize("kw", [dict(order="plain"), dict(order="indegree"),
                                dict(self_loop_once=True)])
def test_reference_row_orders_rmat14(gpu, kw):
    src, dst = T.gen_rmat(14, 16, seed=3)
    beg, adj = T.preprocess_edges(src, dst, **kw)
    want = oracle.count_hash(beg, adj)
    assert want == oracle.count_merge(beg, adj)
    print(f"[gpu] rmat14 {kw}", flush=True)
    for ckw in ({}, dict(threads=1024, blocks=1024, chunk=1), dict(flags=T.FLAG_FORCE_SPILL),
                dict(flags=T.FLAG_SHORT_WALK)):
        r = T.count_csr(beg, adj, device=gpu, **ckw)
        assert r.triangles == want and r.probes == expected_probes(beg, adj, ckw.get("flags", 0))


def test_degenerate_graphs(gpu):
    for beg, adj in (([0], []), ([0, 0], []), ([0, 1, 1], [1]), ([0, 2, 3, 3], [1, 2, 2])):
        r = T.count_csr(np.array(beg), np.array(adj), device=gpu)
        want = 1 if len(adj) == 3 else 0
        assert r.triangles == want


def test_drop_in_entry_point_and_cli(gpu, tmp_path, rmat16):
    beg, adj, want, _ = rmat16
    folder = str(tmp_path)
    T.save_folder(folder, beg, adj)
    a = T.triangle_count(0, folder, 1, 1024, 1024, 1)
    assert a.count == want and a.edge_count == len(adj) and a.vertices == len(beg) - 1
    # a call on a fresh context builds its walk's index and times it with the count
    r = T.count_csr(beg, adj, threads=1024, blocks=1024, chunk=1)
    assert r.triangles == want and r.index_seconds > 0
    assert r.seconds == pytest.approx(r.index_seconds + r.plan_seconds + r.count_seconds)
    assert a.time > 0
    exe = os.path.join(T.BIN_DIR, "trianglecounting")
    out = subprocess.run([exe, folder, "3", "1024", "1024", "1"], capture_output=True, text=True,
                         check=True, timeout=120).stdout
    fields = out.strip().split(",")
    assert fields[0] == folder
    assert int(fields[1]) == len(beg) - 1 and int(fields[2]) == len(adj)
    assert int(fields[3]) == want


@pytest.mark.parametrize("kw", [{}, dict(threads=1024), dict(flags=T.FLAG_FORCE_SPILL),
                                dict(flags=T.FLAG_WAVE_ONLY), dict(chunk=1),
                                dict(flags=T.FLAG_LINEAR_PROBING)],
                         ids=lambda kw: ",".join(f"{k}={v}" for k, v in kw.items()) or "default")
def test_start_index_matches_offset_lookup(gpu, kw):
    # the reference walk reading list starts from the start index must count
    # and probe exactly what the walk that looks up the offsets does
    beg, adj = T.rmat_csr(14, 16, seed=9)
    want = oracle.count_hash(beg, adj)
    w = oracle.probes(beg, adj)
    flags = kw.get("flags", 0) | T.FLAG_REFERENCE_WALK
    rest = {k: v for k, v in kw.items() if k != "flags"}
    with T.Context(gpu) as c:
        c.load(beg, adj)
        first = c.count(flags=flags, **rest)
        assert first.index_seconds > 0  # built here, once
        again = c.count(flags=flags, **rest)
        assert again.index_seconds == 0
        look = c.count(flags=flags | T.FLAG_OFFSET_LOOKUP, **rest)
        for r in (first, again, look):
            assert r.triangles == want and r.probes == w and r.walk_used == 0
        g1 = c.gather(flags=flags, **rest)
        g2 = c.gather(flags=flags | T.FLAG_OFFSET_LOOKUP, **rest)
        assert g1.probes == g2.probes == w
        assert g1.triangles == g2.triangles  # same lists read: same checksum


def test_start_index_default_walk_graphs(gpu):
    # graphs whose automatic walk is the reference one (balanced degrees)
    # take the start index by default
    for n, m, seed in ((3000, 40000, 1), (20000, 160000, 2)):
        rng = np.random.default_rng(seed)
        beg, adj = T.preprocess_edges(rng.integers(0, n, m).astype(np.int32),
                                      rng.integers(0, n, m).astype(np.int32))
        with T.Context(gpu) as c:
            c.load(beg, adj)
            r = c.count()
            assert r.walk_used == 0
            assert r.triangles == oracle.count_hash(beg, adj)
            assert r.triangles == c.count(flags=T.FLAG_OFFSET_LOOKUP).triangles


def test_membench_read_rate(gpu):
    # the L2 read reference: a 2 MiB buffer every workgroup re-reads stays in
    # the L2s, whose read rate is well above HBM's 8 TB/s
    l2 = T.membench_read(2 << 20)
    hbm = T.membench_read(1 << 30, reps=1)
    print(f"[gpu] read rate: 2 MiB {l2 / 1e12:.1f} TB/s, 1 GiB {hbm / 1e12:.1f} TB/s", flush=True)
    assert l2 > 10e12 and l2 > hbm
