"""Python binding of libtrust_amd.so (the MI355X TRUST triangle-counting library).

Thin ctypes layer over the C ABI in include/trust.h, used by the tests and
bench.py.  The library is loaded from this package directory; when it is
missing every call raises -- there is no Python or CPU fallback for any
counting path.

Reference interfaces mirrored here (paths relative to the wzbxpy/TRUST root):
  triangle_count            Triangle_count, Without-graph-partition/TC.cuh:10
  triangle_count_partition  Triangle_count, With-graph-partition/TC.cuh:10
  preprocess_edges          Preprocess/fromDirectToUndirect.cpp + preprocess_outdegree.cpp;
                            order= selects otherPreprocess/preprocess{,_indegree}.cpp,
                            self_loop_once= the speedupIO/preprocess_out.cpp degree rule
  partition_csr / write     Preprocess/partition.cpp
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtrust_amd.so")
BIN_DIR = os.path.join(_HERE, "bin")

TRUST_OK = 0
ERR_ARG, ERR_IO, ERR_FORMAT, ERR_OOM, ERR_HIP, ERR_OVERFLOW, ERR_NODEV = (-1, -2, -3, -4, -5, -6, -7)
SCHED_DEGREE, SCHED_LEARNED, SCHED_AUTO = 0, 1, 2
FLAG_FORCE_SPILL, FLAG_OFFSETS64, FLAG_NO_VALIDATE, FLAG_WAVE_ONLY, FLAG_LINEAR_PROBING = 1, 2, 4, 8, 16
FLAG_HIP_GRAPH = 64
FLAG_REFERENCE_WALK = 128
FLAG_SHORT_WALK = 256
FLAG_OFFSET_LOOKUP = 512
FLAG_FRESH_INDEX = 1024
FLAG_REBUILD_LINEAR = 2048
ORDER_OUTDEGREE, ORDER_PLAIN, ORDER_INDEGREE = 0, 1, 2
ORDERS = {"outdegree": ORDER_OUTDEGREE, "plain": ORDER_PLAIN, "indegree": ORDER_INDEGREE}
PRE_SELF_LOOP_ONCE = 1
# launch parameters the autotuner searches
TUNED_FIELDS = ("threads", "wave_slots", "hub_threshold", "chunk")


class TrustError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"trust error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("threads", C.c_int32), ("blocks", C.c_int32), ("chunk", C.c_int32),
        ("rank", C.c_int32), ("total_process", C.c_int32), ("schedule", C.c_int32),
        ("hub_threshold", C.c_int32), ("wave_slots", C.c_int32), ("lds_budget", C.c_int32),
        ("flags", C.c_uint32), ("device", C.c_int32), ("reserved", C.c_int32 * 3),
        ("core_vertices", C.c_int32), ("core_min_degree", C.c_int32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("vertices", C.c_int64), ("edges", C.c_int64), ("triangles", C.c_uint64),
        ("probes", C.c_uint64), ("seconds", C.c_double), ("count_seconds", C.c_double),
        ("plan_seconds", C.c_double), ("h2d_seconds", C.c_double), ("gteps", C.c_double),
        ("alg_bytes", C.c_double), ("n_spill", C.c_int64), ("n_hub", C.c_int64),
        ("n_wave", C.c_int64), ("threads", C.c_int32), ("blocks", C.c_int32),
        ("hub_slots", C.c_int32), ("wave_slots", C.c_int32),
        ("schedule_used", C.c_int32), ("walk_used", C.c_int32),
        ("index_seconds", C.c_double),
        ("linear_rebuilds", C.c_uint64),
        ("core_vertices", C.c_int64), ("dense_edges", C.c_uint64), ("dense_probes", C.c_uint64),
        ("dense_seconds", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Csr(C.Structure):
    _fields_ = [("vertices", C.c_int64), ("edges", C.c_int64),
                ("beg", C.POINTER(C.c_int64)), ("adj", C.POINTER(C.c_int32))]


class Arguments(C.Structure):
    _fields_ = [("edge_count", C.c_int64), ("count", C.c_uint64), ("time", C.c_double),
                ("degree", C.c_int64), ("vertices", C.c_int64)]


_P64, _P32 = C.POINTER(C.c_int64), C.POINTER(C.c_int32)
_PU64 = C.POINTER(C.c_uint64)
_PF = C.POINTER(C.c_float)

# name -> (restype, argtypes); every symbol of include/trust.h
SIGNATURES = {
    "trust_last_error": (C.c_char_p, []),
    "trust_set_device_memory_cap": (C.c_int, [C.c_uint64]),
    "trust_device_memory_held": (C.c_uint64, []),
    "trust_version": (C.c_char_p, []),
    "trust_params_init": (None, [C.POINTER(Params)]),
    "trust_device_count": (C.c_int, []),
    "trust_triangle_count": (C.c_int, [C.c_int, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(Arguments)]),
    "trust_triangle_count_partition": (C.c_int, [C.c_int, C.c_char_p, C.c_int, C.c_int, C.c_int,
                                                 C.c_int, C.c_int, C.POINTER(Arguments)]),
    "trust_count_csr": (C.c_int, [_P64, _P32, C.c_int64, C.c_int64, C.POINTER(Params),
                                  C.POINTER(Result)]),
    "trust_count_tiles": (C.c_int, [C.POINTER(Csr), C.c_int, C.POINTER(Params),
                                    C.POINTER(Result)]),
    "trust_count_tiles_journal": (C.c_int, [C.POINTER(Csr), C.c_int, C.POINTER(Params),
                                            C.c_char_p, C.c_int64, _PU64, C.POINTER(Result)]),
    # device pointers (beg int64, adj int32), sizes, params, result, HIP stream
    "trust_count_csr_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64,
                                          C.POINTER(Params), C.POINTER(Result), C.c_void_p]),
    "trust_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "trust_ctx_destroy": (None, [C.c_void_p]),
    "trust_ctx_load_csr": (C.c_int, [C.c_void_p, _P64, _P32, C.c_int64, C.c_int64, C.c_uint32,
                                     C.POINTER(C.c_double)]),
    "trust_ctx_count": (C.c_int, [C.c_void_p, C.POINTER(Params), C.POINTER(Result)]),
    "trust_ctx_profile_vertices": (C.c_int, [C.c_void_p, C.POINTER(Params), _PU64, _PU64,
                                             C.POINTER(Result)]),
    "trust_ctx_gather": (C.c_int, [C.c_void_p, C.POINTER(Params), C.POINTER(Result)]),
    "trust_mlp_param_count": (C.c_int64, [C.c_int]),
    "trust_mlp_init": (C.c_int, [C.c_int, C.c_uint64, _PF]),
    "trust_mlp_forward": (C.c_int, [C.c_int, C.c_int, _PF, _PF, C.c_int64, _PF]),
    "trust_mlp_loss_grad": (C.c_int, [C.c_int, C.c_int, _PF, _PF, _PF, C.c_int64, _PF,
                                      C.POINTER(C.c_double)]),
    "trust_mlp_train": (C.c_int, [C.c_int, C.c_int, _PF, _PF, _PF, C.c_int64, C.c_int, C.c_int,
                                  C.c_float, C.c_uint64, C.POINTER(C.c_double)]),
    "trust_ctx_features": (C.c_int, [C.c_void_p, _PF]),
    "trust_ctx_walk_features": (C.c_int, [C.c_void_p, _PF]),
    "trust_ctx_train_cost_model": (C.c_int, [C.c_void_p, C.POINTER(Params), C.c_int, C.c_int64,
                                             C.c_uint64, C.POINTER(C.c_double),
                                             C.POINTER(C.c_double)]),
    "trust_ctx_set_cost_model": (C.c_int, [C.c_void_p, C.c_int, _PF]),
    "trust_ctx_schedule_judgement": (C.c_int, [C.c_void_p, C.POINTER(C.c_int),
                                               C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "trust_ctx_get_cost_model": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), _PF]),
    "trust_ctx_predict_costs": (C.c_int, [C.c_void_p, _PF]),
    "trust_ctx_autotune": (C.c_int, [C.c_void_p, C.POINTER(Params), C.c_int, C.c_double,
                                     C.POINTER(Params), C.POINTER(C.c_double),
                                     C.POINTER(C.c_int)]),
    "trust_batch_shape": (C.c_int, [C.c_int64, C.c_uint64, C.c_int64, C.c_int64, C.c_double,
                                    C.c_double, _P64, C.POINTER(C.c_double), _P32]),
    "trust_batch_generate": (C.c_int, [C.c_int64, C.c_uint64, C.c_int64, C.c_int64, C.c_double,
                                       C.c_double, C.POINTER(Csr), _P64]),
    "trust_ctx_load_batch": (C.c_int, [C.c_void_p, _P64, _P32, C.c_int64, C.c_int64, _P64,
                                       C.c_int64, C.c_uint32]),
    "trust_ctx_count_batch": (C.c_int, [C.c_void_p, C.POINTER(Params), _PU64,
                                        C.POINTER(C.c_double), C.POINTER(Result)]),
    "trust_ctx_batch_segments": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "trust_membench_copy": (C.c_int, [C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_double)]),
    "trust_membench_read": (C.c_int, [C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_double)]),
    "trust_csr_load": (C.c_int, [C.c_char_p, C.POINTER(Csr)]),
    "trust_csr_save": (C.c_int, [C.c_char_p, C.POINTER(Csr)]),
    "trust_csr_validate": (C.c_int, [_P64, _P32, C.c_int64, C.c_int64]),
    "trust_csr_free": (None, [C.POINTER(Csr)]),
    "trust_preprocess_edges": (C.c_int, [_P32, _P32, C.c_int64, C.POINTER(Csr)]),
    "trust_preprocess_edges_ex": (C.c_int, [_P32, _P32, C.c_int64, C.c_int, C.c_uint32,
                                            C.POINTER(Csr)]),
    "trust_partition_csr": (C.c_int, [C.POINTER(Csr), C.c_int, C.POINTER(Csr)]),
    "trust_partition_write": (C.c_int, [C.c_char_p, C.POINTER(Csr), C.c_int]),
    "trust_partition_folder": (C.c_int, [C.c_char_p, C.c_int]),
    "trust_rmat_csr_ooc": (C.c_int, [C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double,
                                     C.c_uint64, C.c_char_p, C.c_char_p, C.c_int64]),
    "trust_gen_rmat": (C.c_int, [C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double,
                                 C.c_uint64, _P32, _P32]),
    "trust_gen_chung_lu": (C.c_int, [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_uint64,
                                     _P32, _P32]),
    "trust_gen_erdos_renyi": (C.c_int, [C.c_int64, C.c_int64, C.c_uint64, _P32, _P32]),
    "trust_gen_barabasi_albert": (C.c_int, [C.c_int64, C.c_int, C.c_uint64, _P32, _P32,
                                            C.POINTER(C.c_int64)]),
}

_lib = None


def lib():
    """Load libtrust_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -c 'import __graft_entry__ as g; g.build()'`")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def _check(rc: int):
    if rc != TRUST_OK:
        raise TrustError(rc, lib().trust_last_error().decode())


def _i64(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(_P64)


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_P32)


def membench_copy(nbytes: int, reps: int = 5, device: int = 0) -> float:
    """Seconds per device-to-device streaming copy of nbytes."""
    t = C.c_double(0)
    _check(lib().trust_membench_copy(device, nbytes, reps, C.byref(t)))
    return t.value


def membench_read(nbytes: int, reps: int = 8, device: int = 0) -> float:
    """Bytes per second of every workgroup (8 per CU) reading an nbytes buffer."""
    t = C.c_double(0)
    _check(lib().trust_membench_read(device, nbytes, reps, C.byref(t)))
    return t.value


def device_count() -> int:
    n = lib().trust_device_count()
    if n < 0:
        _check(n)
    return n


def set_device_memory_cap(nbytes: int) -> None:
    """Cap on the device bytes the library's buffers may hold together (0: none);
    allocations beyond it fail with ERR_OOM."""
    _check(lib().trust_set_device_memory_cap(nbytes))


def device_memory_held() -> int:
    """Device bytes the library's buffers hold now."""
    return int(lib().trust_device_memory_held())


def make_params(**kw) -> Params:
    p = Params()
    lib().trust_params_init(C.byref(p))
    for k, v in kw.items():
        if k == "reserved":
            for i, x in enumerate(v):
                p.reserved[i] = x
        else:
            setattr(p, k, v)
    return p


# ---------------------------------------------------------------- CSR helpers
def _take_csr(c: Csr):
    try:
        V, E = c.vertices, c.edges
        beg = np.ctypeslib.as_array(c.beg, shape=(V + 1,)).copy()
        adj = (np.ctypeslib.as_array(c.adj, shape=(E,)).copy() if E > 0
               else np.zeros(0, np.int32))
    finally:
        lib().trust_csr_free(C.byref(c))
    return beg, adj


def _as_csr(beg, adj):
    beg, pb = _i64(beg)
    adj, pa = _i32(adj)
    return Csr(len(beg) - 1, len(adj), pb, pa), (beg, adj)


def load_folder(folder: str):
    c = Csr()
    _check(lib().trust_csr_load(folder.encode(), C.byref(c)))
    return _take_csr(c)


def save_folder(folder: str, beg, adj):
    c, keep = _as_csr(beg, adj)
    _check(lib().trust_csr_save(folder.encode(), C.byref(c)))


def validate(beg, adj):
    beg, pb = _i64(beg)
    adj, pa = _i32(adj)
    _check(lib().trust_csr_validate(pb, pa, len(beg) - 1, len(adj)))


def preprocess_edges(src, dst, order="outdegree", self_loop_once=False):
    """Raw edge pairs -> (beg int64, adj int32) in the reference's oriented layout.

    order: "outdegree" (preprocess_outdegree.cpp, the default), "plain"
    (otherPreprocess/preprocess.cpp) or "indegree"
    (otherPreprocess/preprocess_indegree.cpp).  self_loop_once: a self-loop
    adds 1 to the degree, as after speedupIO/preprocess_out.cpp's list dedup.
    """
    src, ps = _i32(src)
    dst, pd = _i32(dst)
    c = Csr()
    if order == "outdegree" and not self_loop_once:
        _check(lib().trust_preprocess_edges(ps, pd, len(src), C.byref(c)))
    else:
        if order not in ORDERS:
            raise ValueError(f"unknown order {order!r}")
        flags = PRE_SELF_LOOP_ONCE if self_loop_once else 0
        _check(lib().trust_preprocess_edges_ex(ps, pd, len(src), ORDERS[order], flags,
                                               C.byref(c)))
    return _take_csr(c)


def partition_csr(beg, adj, n: int):
    """n x n hash-partition tiles; returns list of (beg, adj), index i*n+j."""
    c, keep = _as_csr(beg, adj)
    tiles = (Csr * (n * n))()
    _check(lib().trust_partition_csr(C.byref(c), n, tiles))
    return [_take_csr(tiles[k]) for k in range(n * n)]


def partition_folder(folder: str, n: int):
    """Tile folders written from folder/begin.bin + adjacent.bin, one tile row in memory."""
    _check(lib().trust_partition_folder(folder.encode(), n))


def partition_write(folder: str, beg, adj, n: int):
    c, keep = _as_csr(beg, adj)
    _check(lib().trust_partition_write(folder.encode(), C.byref(c), n))


# ---------------------------------------------------------------- generators
RMAT_ABC = (0.57, 0.19, 0.19)


def rmat_csr_ooc(scale: int, edge_factor: int, seed: int, out_folder: str, work_dir: str,
                 bucket_pairs: int = 1 << 27, abc=RMAT_ABC) -> None:
    """gen_rmat + preprocess_edges out of core: writes out_folder/begin.bin and
    adjacent.bin through bucket files of at most bucket_pairs keys in work_dir."""
    m = edge_factor << scale
    _check(lib().trust_rmat_csr_ooc(scale, m, abc[0], abc[1], abc[2], seed,
                                    os.fsencode(out_folder), os.fsencode(work_dir), bucket_pairs))


def gen_rmat(scale: int, edge_factor: int = 16, seed: int = 1, abc=RMAT_ABC):
    m = edge_factor << scale
    src = np.empty(m, np.int32)
    dst = np.empty(m, np.int32)
    _check(lib().trust_gen_rmat(scale, m, abc[0], abc[1], abc[2], seed,
                                src.ctypes.data_as(_P32), dst.ctypes.data_as(_P32)))
    return src, dst


def gen_chung_lu(n: int, m: int, gamma: float = 2.5, max_degree: float = 800.0, seed: int = 1):
    src = np.empty(m, np.int32)
    dst = np.empty(m, np.int32)
    _check(lib().trust_gen_chung_lu(n, m, gamma, max_degree, seed,
                                    src.ctypes.data_as(_P32), dst.ctypes.data_as(_P32)))
    return src, dst


def gen_erdos_renyi(n: int, m: int, seed: int = 1):
    src = np.empty(m, np.int32)
    dst = np.empty(m, np.int32)
    _check(lib().trust_gen_erdos_renyi(n, m, seed, src.ctypes.data_as(_P32),
                                       dst.ctypes.data_as(_P32)))
    return src, dst


def gen_barabasi_albert(n: int, k: int, seed: int = 1):
    src = np.empty(n * k, np.int32)
    dst = np.empty(n * k, np.int32)
    m = C.c_int64(0)
    _check(lib().trust_gen_barabasi_albert(n, k, seed, src.ctypes.data_as(_P32),
                                           dst.ctypes.data_as(_P32), C.byref(m)))
    return src[: m.value], dst[: m.value]


def rmat_csr(scale: int, edge_factor: int = 16, seed: int = 1):
    return preprocess_edges(*gen_rmat(scale, edge_factor, seed))


# ---------------------------------------------------------------- batches
def batch_shape(ngraphs: int, seed: int = 1, nmin: int = 1000, nmax: int = 200000,
                dmin: float = 8.0, dmax: float = 64.0):
    """Per-graph (n, mean degree, kind) of a seeded batch; kind 0 ER, 1 BA."""
    n = np.zeros(ngraphs, np.int64)
    d = np.zeros(ngraphs, np.float64)
    k = np.zeros(ngraphs, np.int32)
    _check(lib().trust_batch_shape(ngraphs, seed, nmin, nmax, dmin, dmax, n.ctypes.data_as(_P64),
                                   d.ctypes.data_as(C.POINTER(C.c_double)),
                                   k.ctypes.data_as(_P32)))
    return n, d, k


def batch_generate(ngraphs: int, seed: int = 1, nmin: int = 1000, nmax: int = 200000,
                   dmin: float = 8.0, dmax: float = 64.0):
    """A seeded ER/BA batch as one CSR with global ids: (beg, adj, vbase)."""
    c = Csr()
    vbase = np.zeros(ngraphs + 1, np.int64)
    _check(lib().trust_batch_generate(ngraphs, seed, nmin, nmax, dmin, dmax, C.byref(c),
                                      vbase.ctypes.data_as(_P64)))
    beg, adj = _take_csr(c)
    return beg, adj, vbase


def batch_graph(beg, adj, vbase, g: int):
    """Graph g of a batch as its own CSR with local ids."""
    lo, hi = int(vbase[g]), int(vbase[g + 1])
    b = np.asarray(beg[lo: hi + 1], np.int64)
    return b - b[0], np.asarray(adj[b[0]: b[-1]], np.int32) - np.int32(lo)


# ---------------------------------------------------------------- counting
def count_csr(beg, adj, **params) -> Result:
    beg, pb = _i64(beg)
    adj, pa = _i32(adj)
    p = make_params(**params)
    r = Result()
    _check(lib().trust_count_csr(pb, pa, len(beg) - 1, len(adj), C.byref(p), C.byref(r)))
    return r


def count_tiles(tiles, n: int, **params) -> Result:
    arr = (Csr * (n * n))()
    keep = []
    for k, (b, a) in enumerate(tiles):
        c, kk = _as_csr(b, a)
        arr[k] = c
        keep.append(kk)
    p = make_params(**params)
    r = Result()
    _check(lib().trust_count_tiles(arr, n, C.byref(p), C.byref(r)))
    return r


def count_tiles_journal(tiles, n: int, journal: str, max_new_jobs: int = 0, **params):
    """count_tiles that records each finished job in a journal file and skips
    the jobs already there: (jobs still missing, Result, per-job counts)."""
    arr = (Csr * (n * n))()
    keep = []
    for k, (b, a) in enumerate(tiles):
        c, kk = _as_csr(b, a)
        arr[k] = c
        keep.append(kk)
    p = make_params(**params)
    r = Result()
    jobs = np.zeros(n ** 3, np.uint64)
    rc = lib().trust_count_tiles_journal(arr, n, C.byref(p), os.fsencode(journal), max_new_jobs,
                                         jobs.ctypes.data_as(_PU64), C.byref(r))
    if rc < 0:
        _check(rc)
    return rc, r, jobs


def triangle_count(rank: int, folder: str, total_process: int = 1, threads: int = 0,
                   blocks: int = 0, chunk_size: int = 0) -> Arguments:
    a = Arguments()
    _check(lib().trust_triangle_count(rank, folder.encode(), total_process, threads, blocks,
                                      chunk_size, C.byref(a)))
    return a


def triangle_count_partition(rank: int, folder: str, total_process: int, threads: int,
                             blocks: int, chunk_size: int, partition_num: int) -> Arguments:
    a = Arguments()
    _check(lib().trust_triangle_count_partition(rank, folder.encode(), total_process, threads,
                                                blocks, chunk_size, partition_num, C.byref(a)))
    return a


class Context:
    """A device context holding one resident graph (trust_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().trust_ctx_create(device, C.byref(h)))
        self._h = h
        self.V = self.E = self.ngraphs = 0

    def close(self):
        if self._h:
            lib().trust_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, beg, adj, flags: int = 0) -> float:
        beg, pb = _i64(beg)
        adj, pa = _i32(adj)
        t = C.c_double(0)
        _check(lib().trust_ctx_load_csr(self._h, pb, pa, len(beg) - 1, len(adj), flags,
                                        C.byref(t)))
        self.V, self.E = len(beg) - 1, len(adj)
        return t.value

    def count(self, **params) -> Result:
        p = make_params(**params)
        r = Result()
        _check(lib().trust_ctx_count(self._h, C.byref(p), C.byref(r)))
        return r

    def autotune(self, reps: int = 2, max_seconds: float = 0.0, **base):
        """Fastest launch parameters for the resident graph: (params dict,
        seconds per count, candidates timed)."""
        p = make_params(**base)
        best = Params()
        sec, n = C.c_double(0), C.c_int(0)
        _check(lib().trust_ctx_autotune(self._h, C.byref(p), reps, max_seconds, C.byref(best),
                                        C.byref(sec), C.byref(n)))
        tuned = {k: getattr(best, k) for k in TUNED_FIELDS}
        return tuned, sec.value, n.value

    def load_batch(self, beg, adj, vbase, flags: int = 0):
        beg, pb = _i64(beg)
        adj, pa = _i32(adj)
        vbase, pv = _i64(vbase)
        _check(lib().trust_ctx_load_batch(self._h, pb, pa, len(beg) - 1, len(adj), pv,
                                          len(vbase) - 1, flags))
        self.V, self.E, self.ngraphs = len(beg) - 1, len(adj), len(vbase) - 1

    def count_batch(self, **params):
        """One count of the resident batch: (counts[ng] uint64, done_seconds[ng], Result)."""
        p = make_params(**params)
        r = Result()
        counts = np.zeros(self.ngraphs, np.uint64)
        done = np.zeros(self.ngraphs, np.float64)
        _check(lib().trust_ctx_count_batch(self._h, C.byref(p), counts.ctypes.data_as(_PU64),
                                           done.ctypes.data_as(C.POINTER(C.c_double)),
                                           C.byref(r)))
        return counts, done, r

    def batch_segments(self, max_edges: int) -> int:
        """Re-cut the resident int64 batch into int32 segments of at most
        max_edges edges (0: count it whole); returns the segment count."""
        n = C.c_int64(0)
        _check(lib().trust_ctx_batch_segments(self._h, max_edges, C.byref(n)))
        return n.value

    def gather(self, **params) -> Result:
        """Gather-only pass (same walk, no hashing): the load-path time."""
        p = make_params(**params)
        r = Result()
        _check(lib().trust_ctx_gather(self._h, C.byref(p), C.byref(r)))
        return r

    def features(self) -> np.ndarray:
        """GPU features, shape (V, 3): log2(1 + (d+, sum d+(v), max d+(v))) / 32."""
        out = np.zeros((self.V, 3), np.float32)
        _check(lib().trust_ctx_features(self._h, out.ctypes.data_as(_PF)))
        return out

    def walk_features(self) -> np.ndarray:
        """GPU features over the short-side walk's rows, shape (V, 3)."""
        out = np.zeros((self.V, 3), np.float32)
        _check(lib().trust_ctx_walk_features(self._h, out.ctypes.data_as(_PF)))
        return out

    def train_cost_model(self, epochs: int = 30, max_samples: int = 200000, seed: int = 1,
                         **params):
        """Instrumented wave-only count + GPU training; returns (mse, seconds)."""
        p = make_params(**params)
        loss, sec = C.c_double(0), C.c_double(0)
        _check(lib().trust_ctx_train_cost_model(self._h, C.byref(p), epochs, max_samples, seed,
                                                C.byref(loss), C.byref(sec)))
        return loss.value, sec.value

    def set_cost_model(self, params):
        params = np.ascontiguousarray(params, np.float32)
        _check(lib().trust_ctx_set_cost_model(self._h, MLP_HIDDEN, params.ctypes.data_as(_PF)))

    def schedule_judgement(self):
        """(decision, learned seconds, degree seconds) of the last training on this
        graph; decision is SCHED_LEARNED, SCHED_DEGREE or -1 (not judged)."""
        d, ls, ds = C.c_int(0), C.c_double(0), C.c_double(0)
        _check(lib().trust_ctx_schedule_judgement(self._h, C.byref(d), C.byref(ls), C.byref(ds)))
        return d.value, ls.value, ds.value

    def get_cost_model(self) -> np.ndarray:
        h = C.c_int(0)
        out = np.zeros(mlp_param_count(), np.float32)
        _check(lib().trust_ctx_get_cost_model(self._h, C.byref(h), out.ctypes.data_as(_PF)))
        return out

    def predict_costs(self) -> np.ndarray:
        out = np.zeros(self.V, np.float32)
        _check(lib().trust_ctx_predict_costs(self._h, out.ctypes.data_as(_PF)))
        return out

    def profile_vertices(self, **params):
        """Instrumented count: (cycles[V], probes[V], Result)."""
        p = make_params(**params)
        r = Result()
        cyc = np.zeros(self.V, np.uint64)
        prb = np.zeros(self.V, np.uint64)
        _check(lib().trust_ctx_profile_vertices(self._h, C.byref(p), cyc.ctypes.data_as(_PU64),
                                                prb.ctypes.data_as(_PU64), C.byref(r)))
        return cyc, prb, r


# ---------------------------------------------------------------- cost model
MLP_HIDDEN = 32


def mlp_param_count(hidden: int = MLP_HIDDEN) -> int:
    return int(lib().trust_mlp_param_count(hidden))


def mlp_init(seed: int = 1, hidden: int = MLP_HIDDEN) -> np.ndarray:
    p = np.zeros(mlp_param_count(hidden), np.float32)
    _check(lib().trust_mlp_init(hidden, seed, p.ctypes.data_as(_PF)))
    return p


def _f32(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(_PF)


def mlp_forward(params, x, device: int = 0) -> np.ndarray:
    params, pp = _f32(params)
    x, px = _f32(x)
    y = np.zeros(len(x), np.float32)
    _check(lib().trust_mlp_forward(device, MLP_HIDDEN, pp, px, len(x), y.ctypes.data_as(_PF)))
    return y


def mlp_loss_grad(params, x, target, device: int = 0):
    params, pp = _f32(params)
    x, px = _f32(x)
    target, pt = _f32(target)
    g = np.zeros(len(params), np.float32)
    loss = C.c_double(0)
    _check(lib().trust_mlp_loss_grad(device, MLP_HIDDEN, pp, px, pt, len(x),
                                     g.ctypes.data_as(_PF), C.byref(loss)))
    return g, loss.value


def mlp_train(params, x, target, epochs=20, batch=1024, lr=3e-3, seed=1, device: int = 0):
    params = np.array(params, np.float32, copy=True)
    x, px = _f32(x)
    target, pt = _f32(target)
    loss = C.c_double(0)
    _check(lib().trust_mlp_train(device, MLP_HIDDEN, params.ctypes.data_as(_PF), px, pt, len(x),
                                 epochs, batch, lr, seed, C.byref(loss)))
    return params, loss.value
