"""CPU oracle (test infrastructure only).

ctypes binding of oracle/liboracle.so (tc_oracle.c) plus pure-Python
restatements used as checkers.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this package; the product library
(trust_amd/) never does.

Functions and the reference code they restate (paths relative to the
wzbxpy/TRUST root):
  count_hash        Without-graph-partition/TC.cu:262-305,330-359,435-545,70-110
  count_merge       Without-graph-partition/graph.hpp:121-149 (independent method)
  count_brute       no reference counterpart (O(n^3/64) bitset)
  count_tile_job    With-graph-partition/TC.cu:630-973, 985-994
  probes            number of linear_search calls (TC.cu:354,534)
  probes_short      probes of the product's short-side walk (no reference
                    counterpart; min(d+(u), d+(v)) per oriented edge)
  features          per-vertex d+(u), sum d+(v), max d+(v)
  features_short    the same over the short-side walk's rows
  mlp_*             numpy float64 forward/backward of the cost model
                    (no reference counterpart)
Pinning: tests/test_oracle.py checks every function against closed forms,
brute force and the networkx triangle counts committed under tests/golden/.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_DIR = os.path.join(_HERE, "_ref")

_P64, _P32 = C.POINTER(C.c_int64), C.POINTER(C.c_int32)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        h = C.CDLL(LIB_PATH)
        h.oracle_count_hash.restype = C.c_uint64
        h.oracle_count_hash.argtypes = [_P64, _P32, C.c_int64, C.c_int, _P64]
        h.oracle_count_merge.restype = C.c_uint64
        h.oracle_count_merge.argtypes = [_P64, _P32, C.c_int64, C.c_int]
        h.oracle_count_brute.restype = C.c_uint64
        h.oracle_count_brute.argtypes = [C.c_int64, _P32, _P32, C.c_int64]
        h.oracle_count_tile_job.restype = C.c_uint64
        h.oracle_count_tile_job.argtypes = [_P64, _P32, _P64, _P32, _P64, _P32, C.c_int64,
                                            C.c_int64, C.c_int]
        h.oracle_probes.restype = C.c_uint64
        h.oracle_probes.argtypes = [_P64, _P32, C.c_int64, _P64]
        h.oracle_probes_short.restype = C.c_uint64
        h.oracle_probes_short.argtypes = [_P64, _P32, C.c_int64, _P64]
        h.oracle_features_short.restype = None
        h.oracle_features_short.argtypes = [_P64, _P32, C.c_int64, C.POINTER(C.c_double)]
        h.oracle_features.restype = None
        h.oracle_features.argtypes = [_P64, _P32, C.c_int64, C.POINTER(C.c_double)]
        h.oracle_count_hash_sample.restype = C.c_uint64
        h.oracle_count_hash_sample.argtypes = [_P64, _P32, C.c_int64, C.c_int64, C.c_int64,
                                               C.c_int, C.POINTER(C.c_uint64)]
        h.oracle_count_hash_batch.restype = None
        h.oracle_count_hash_batch.argtypes = [_P64, _P32, C.c_int64, _P64, C.c_int64, C.c_int,
                                              C.POINTER(C.c_uint64)]
        h.oracle_max_threads.restype = C.c_int
        _lib = h
    return _lib


def _i64(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(_P64)


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_P32)


def count_hash(beg, adj, threads: int = 0, return_max_bucket: bool = False):
    beg, pb = _i64(beg)
    adj, pa = _i32(adj)
    mb = C.c_int64(0)
    c = lib().oracle_count_hash(pb, pa, len(beg) - 1, threads, C.byref(mb))
    return (int(c), mb.value) if return_max_bucket else int(c)


def count_hash_batch(beg, adj, vbase, threads: int = 0):
    """Per-graph hashed counts of a batch (uint64 array, one per graph)."""
    beg, pb = _i64(beg)
    adj, pa = _i32(adj)
    vbase, pv = _i64(vbase)
    out = np.zeros(len(vbase) - 1, np.uint64)
    lib().oracle_count_hash_batch(pb, pa, len(beg) - 1, pv, len(vbase) - 1, threads,
                                  out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out


def count_merge(beg, adj, threads: int = 0) -> int:
    beg, pb = _i64(beg)
    adj, pa = _i32(adj)
    return int(lib().oracle_count_merge(pb, pa, len(beg) - 1, threads))


def count_brute(n: int, src, dst) -> int:
    src, ps = _i32(src)
    dst, pd = _i32(dst)
    return int(lib().oracle_count_brute(n, ps, pd, len(src)))


def count_tile_job(ht, walk, probe, threads: int = 0) -> int:
    (hb, ph), (ha, pha) = _i64(ht[0]), _i32(ht[1])
    (wb, pw), (wa, pwa) = _i64(walk[0]), _i32(walk[1])
    (qb, pq), (qa, pqa) = _i64(probe[0]), _i32(probe[1])
    return int(lib().oracle_count_tile_job(ph, pha, pw, pwa, pq, pqa, len(hb) - 1, len(qb) - 1,
                                           threads))


def count_partitioned(tiles, n: int, threads: int = 0) -> int:
    """Sum of all n^3 (i,j,k) jobs; tiles[i*n+j] = (beg, adj)."""
    total = 0
    for i in range(n):
        for j in range(n):
            for k in range(n):
                total += count_tile_job(tiles[i * n + j], tiles[i * n + k], tiles[k * n + j],
                                        threads)
    return total


def probes(beg, adj, per_vertex: bool = False):
    beg, pb = _i64(beg)
    adj, pa = _i32(adj)
    V = len(beg) - 1
    out = np.zeros(V, np.int64) if per_vertex else None
    w = lib().oracle_probes(pb, pa, V, out.ctypes.data_as(_P64) if per_vertex else None)
    return (int(w), out) if per_vertex else int(w)


def probes_short(beg, adj, per_vertex: bool = False):
    beg, pb = _i64(beg)
    adj, pa = _i32(adj)
    V = len(beg) - 1
    out = np.zeros(V, np.int64) if per_vertex else None
    w = lib().oracle_probes_short(pb, pa, V, out.ctypes.data_as(_P64) if per_vertex else None)
    return (int(w), out) if per_vertex else int(w)


def features(beg, adj) -> np.ndarray:
    beg, pb = _i64(beg)
    adj, pa = _i32(adj)
    V = len(beg) - 1
    out = np.zeros((V, 3), np.float64)
    lib().oracle_features(pb, pa, V, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def count_hash_sample(beg, adj, first: int, stride: int, threads: int = 0):
    """(partial count, probes) over u = first, first+stride, ..."""
    beg, pb = _i64(beg)
    adj, pa = _i32(adj)
    pr = C.c_uint64(0)
    c = lib().oracle_count_hash_sample(pb, pa, len(beg) - 1, first, stride, threads, C.byref(pr))
    return int(c), int(pr.value)


def max_threads() -> int:
    return int(lib().oracle_max_threads())


def ref_tool(name: str) -> str | None:
    """Path of a reference tool built into oracle/_ref (None when absent)."""
    p = os.path.join(REF_DIR, name)
    return p if os.path.exists(p) else None


# ---------------------------------------------------------------- cost-model MLP
# float64 restatement of the network of trust_amd/csrc/mlp.hip (no reference
# counterpart): x (3) -> relu(W1 x + b1) (H) -> relu(W2 h1 + b2) (H) -> W3 h2 + b3,
# parameters laid out W1[H][3], b1[H], W2[H][H], b2[H], W3[H], b3.  Pinned
# by central finite differences in tests/test_oracle.py.
MLP_H = 32


def mlp_unpack(p, H=MLP_H):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + 3 * H].reshape(H, 3); o += 3 * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def mlp_forward(p, x, H=MLP_H):
    W1, b1, W2, b2, W3, b3 = mlp_unpack(p, H)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3 + b3


def mlp_loss_grad(p, x, t, H=MLP_H):
    """(gradient of mean((y-t)^2) w.r.t. the flat parameters, loss)."""
    W1, b1, W2, b2, W3, b3 = mlp_unpack(p, H)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    n = len(x)
    a1 = x @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    y = h2 @ W3 + b3
    err = y - t
    dy = 2.0 * err / n
    gW3 = h2.T @ dy
    gb3 = dy.sum()
    dh2 = np.outer(dy, W3) * (a2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (a1 > 0)
    gW1 = dh1.T @ x
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return g, float((err ** 2).mean())


def features_short(beg, adj) -> np.ndarray:
    beg, pb = _i64(beg)
    adj, pa = _i32(adj)
    V = len(beg) - 1
    out = np.zeros((V, 3), np.float64)
    lib().oracle_features_short(pb, pa, V, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def mlp_scaled_features_short(beg, adj):
    """log2(1 + (d+, sum, max)) / 32 over the short-side walk's rows."""
    return (np.log2(1.0 + features_short(beg, adj)) / 32.0).astype(np.float64)


def mlp_scaled_features(beg, adj):
    """log2(1 + (d+, sum d+(v), max d+(v))) / 32, as the GPU feature kernel."""
    return (np.log2(1.0 + features(beg, adj)) / 32.0).astype(np.float64)
