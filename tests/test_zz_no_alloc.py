"""The allocation-free entry point (trust_count_csr_noalloc) against
trust_count_csr and the oracle, through the C ABI.

The CSR is placed in device memory by the test itself with hipMalloc from
the HIP runtime the library links (the caller's allocation, as a caller that
cannot allocate would hold it already); the library call allocates nothing.  Two calls in a row on one device share the library's
scratch array, the second on a smaller graph, so a stale value left by the
first call (queue cursors, tier bounds, counts) would show.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import trust_amd as T

pytestmark = pytest.mark.gpu


_hip = None


def hip():
    """The HIP runtime libtrust_amd.so is linked against (same instance)."""
    global _hip
    if _hip is None:
        T.lib()
        _hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipFree.argtypes = [C.c_void_p]
    return _hip


class DeviceArray:
    """A caller-owned device copy of a host array (freed with the object)."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), max(host.nbytes, 16)) == 0
        if host.nbytes:
            assert hip().hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, 1) == 0  # H2D

    def __del__(self):
        if self.ptr:
            hip().hipFree(self.ptr)


def on_device(beg, adj):
    return DeviceArray(np.asarray(beg, np.int64)), DeviceArray(np.asarray(adj, np.int32))


def noalloc(b, a, V, E, **params):
    p = T.make_params(**params)
    r = T.Result()
    rc = T.lib().trust_count_csr_noalloc(b.ptr, a.ptr, C.c_int64(V), C.c_int64(E), C.byref(p),
                                         C.byref(r), C.c_void_p(0))
    return rc, r


@pytest.mark.parametrize("kw", [dict(), dict(threads=1024), dict(flags=T.FLAG_WAVE_ONLY),
                                dict(hub_threshold=0), dict(threads=1024, blocks=1024, chunk=1)],
                         ids=lambda kw: ",".join(f"{k}={v}" for k, v in kw.items()) or "default")
def test_noalloc_matches_count_csr_twice(gpu, kw):
    big = T.rmat_csr(16, 16, seed=1)
    small = T.rmat_csr(12, 8, seed=5)
    for beg, adj in (big, small):
        print(f"[gpu] noalloc V={len(beg) - 1} E={len(adj)} {kw}", flush=True)
        want = oracle.count_hash(beg, adj)
        rkw = dict(kw, flags=kw.get("flags", 0) | T.FLAG_REFERENCE_WALK | T.FLAG_OFFSET_LOOKUP)
        ref = T.count_csr(beg, adj, **rkw)
        b, a = on_device(beg, adj)
        rc, r = noalloc(b, a, len(beg) - 1, len(adj), **kw)
        assert rc == 0, T.lib().trust_last_error().decode()
        assert r.triangles == ref.triangles == want
        assert r.probes == ref.probes == oracle.probes(beg, adj)
        assert (r.n_hub, r.n_wave) == (ref.n_hub, ref.n_wave)
        assert r.seconds > 0


def test_noalloc_closed_forms_and_empty(gpu):
    from conftest import CLOSED_FORMS, closed_form_csr

    for name in sorted(CLOSED_FORMS):
        beg, adj, want = closed_form_csr(name)
        b, a = on_device(beg, adj)
        rc, r = noalloc(b, a, len(beg) - 1, len(adj))
        assert rc == 0 and r.triangles == want, name
    b, a = on_device(np.zeros(4, np.int64), np.zeros(0, np.int32))
    rc, r = noalloc(b, a, 3, 0)
    assert rc == 0 and r.triangles == 0


def test_noalloc_rejects_what_it_cannot_hold(gpu):
    beg, adj = T.rmat_csr(10, 8, seed=1)
    b, a = on_device(beg, adj)
    V, E = len(beg) - 1, len(adj)
    assert noalloc(b, a, (1 << 22) + 1, E)[0] == T.ERR_OVERFLOW  # checked before any device work
    assert noalloc(b, a, V, 1 << 30)[0] == T.ERR_OVERFLOW
    for f in (T.FLAG_FORCE_SPILL, T.FLAG_SHORT_WALK, T.FLAG_HIP_GRAPH, T.FLAG_OFFSETS64):
        assert noalloc(b, a, V, E, flags=f)[0] == T.ERR_ARG
    assert noalloc(b, a, V, E, schedule=T.SCHED_LEARNED)[0] == T.ERR_ARG
    # a CSR that fails the device-side checks is refused, not counted
    bad = adj.copy()
    bad[len(bad) // 2] = V + 7  # an id past V: read by the check kernel only
    b2, a2 = on_device(beg, bad)
    assert noalloc(b2, a2, V, E)[0] == T.ERR_FORMAT
    # and the scratch still serves a good call afterwards
    rc, r = noalloc(b, a, V, E)
    assert rc == 0 and r.triangles == oracle.count_hash(beg, adj)
