"""Out-of-memory paths on the GPU, forced from the host with the library's
device-memory cap (trust_set_device_memory_cap): a failed allocation must
come back as TRUST_ERR_OOM with the context still usable, and the walk
index must fall back to its low-scratch build and still count exactly.
The reference exits on any allocation failure
(Without-graph-partition/herror.h:1-8)."""
import pytest

import oracle
import trust_amd as T

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat18():
    beg, adj = T.rmat_csr(18, 16, seed=1)
    return beg, adj, oracle.count_hash(beg, adj)


def test_oom_error_then_fallback_then_recovery(gpu, rmat18):
    beg, adj, want = rmat18
    V, E = len(beg) - 1, len(adj)
    flags = T.FLAG_SHORT_WALK | T.FLAG_FRESH_INDEX
    try:
        with T.Context(gpu) as c:
            c.load(beg, adj)
            held = T.device_memory_held()
            assert held >= 4 * E
            # no room for anything the count allocates: a clean ERR_OOM
            print(f"[gpu] cap at the loaded graph's {held} bytes", flush=True)
            T.set_device_memory_cap(held)
            with pytest.raises(T.TrustError) as e:
                c.count(flags=flags)
            assert e.value.code == T.ERR_OOM and "cap" in str(e.value)
            # room for the atomic build of the walk rows (8 bytes per entry,
            # per-row counters) and the count's own buffers, not for the
            # binned build's scratch (per-edge d+, 12-byte records): the
            # count falls back and stays exact
            cap = held + 8 * E + 24 * V + (8 << 20)
            print(f"[gpu] cap {cap} bytes: fallback build", flush=True)
            T.set_device_memory_cap(cap)
            r = c.count(flags=flags)
            assert r.triangles == want and r.walk_used == 1
            assert r.probes == oracle.probes_short(beg, adj)
            assert T.device_memory_held() <= cap
            # the cap lifted: the binned build again, same count
            T.set_device_memory_cap(0)
            r2 = c.count(flags=flags)
            assert r2.triangles == want and r2.probes == r.probes
            assert T.device_memory_held() > cap - (8 << 20) - 24 * V
    finally:
        T.set_device_memory_cap(0)


def test_dense_core_without_room_is_left_out(gpu, rmat18):
    # a core asked for on a graph whose core buffers (ids, target lists: 2
    # bytes per entry) do not fit: the count goes on without it, exactly
    beg, adj, want = rmat18
    V, E = len(beg) - 1, len(adj)
    kw = dict(flags=T.FLAG_SHORT_WALK | T.FLAG_FRESH_INDEX, core_vertices=1024,
              core_min_degree=16)
    try:
        with T.Context(gpu) as c:
            c.load(beg, adj)
            full = c.count(**kw)
            assert full.triangles == want and full.core_vertices > 0
        with T.Context(gpu) as c:
            c.load(beg, adj)
            held = T.device_memory_held()
            cap = held + 8 * E + 24 * V + (8 << 20)
            print(f"[gpu] cap {cap} bytes: no room for the core", flush=True)
            T.set_device_memory_cap(cap)
            r = c.count(**kw)
            assert r.triangles == want and r.walk_used == 1
            assert r.core_vertices == 0 and r.dense_edges == 0
            assert r.probes == oracle.probes_short(beg, adj)
            assert T.device_memory_held() <= cap
            # with room again, the next fresh count has its core
            T.set_device_memory_cap(0)
            r2 = c.count(**kw)
            assert r2.triangles == want and r2.core_vertices == full.core_vertices
            assert r2.probes + r2.dense_probes == oracle.probes_short(beg, adj)
    finally:
        T.set_device_memory_cap(0)


def test_oom_in_batch_and_tiles_is_clean(gpu, rmat18):
    beg, adj, want = rmat18
    tiles = T.partition_csr(beg, adj, 2)
    try:
        with T.Context(gpu) as c:
            # no room beyond what the process holds now: loading a graph and
            # streaming tiles fail cleanly
            T.set_device_memory_cap(T.device_memory_held() + 1)
            with pytest.raises(T.TrustError) as e:
                c.load(beg, adj)
            assert e.value.code == T.ERR_OOM
            with pytest.raises(T.TrustError) as e:
                T.count_tiles(tiles, 2)
            assert e.value.code == T.ERR_OOM
            T.set_device_memory_cap(0)
            c.load(beg, adj)
            assert c.count().triangles == want
            assert T.count_tiles(tiles, 2).triangles == want
    finally:
        T.set_device_memory_cap(0)
