"""C-ABI boundary on the CPU: the library loads, exports every function that
include/trust.h declares, and every failure path returns an error code
(never an exit) -- the reference exits on any CUDA error (herror.h:1-8) and
validates nothing it reads (graph.hpp:38-54)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import trust_amd as T
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "trust.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(trust_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "trust_triangle_count" in names and "trust_count_csr" in names
    assert len(names) >= 20


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports_symbol(name):
    lib = ctypes.CDLL(T.LIB_PATH)
    assert hasattr(lib, name)
    assert name in T.SIGNATURES  # the Python binding types every entry point


def header_struct(name):
    """(field, C type, array length) of `typedef struct name {...}` in trust.h."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct " + name + r" \{(.*?)\} " + name + ";", text, re.S).group(1)
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"(.+?)\s*\b(\w+)(\[(\d+)\])?$", decl)
        ctype, field, n = m.group(1).replace("*", "").strip(), m.group(2), m.group(4)
        out.append((field, ctype, int(n) if n else 1))
    return out


@pytest.mark.parametrize("name,py", [("trust_params", T.Params), ("trust_result", T.Result),
                                     ("trust_arguments", T.Arguments)])
def test_struct_layout_matches_header(name, py):
    # the ctypes mirrors bind the same fields, in order, with the same sizes
    sizes = {"int32_t": 4, "uint32_t": 4, "int64_t": 8, "uint64_t": 8, "double": 8}
    fields = header_struct(name)
    assert [f for f, _, _ in fields] == [f for f, _ in py._fields_]
    for (field, ctype, n), (_, pytype) in zip(fields, py._fields_):
        assert ctypes.sizeof(pytype) == sizes[ctype] * n, field


def test_version_and_defaults():
    assert b"gfx950" in T.lib().trust_version()
    p = T.make_params()
    assert p.total_process == 1 and p.hub_threshold == 100 and p.schedule == T.SCHED_DEGREE


def test_device_memory_cap_setting():
    # no device needed: the cap and the held-bytes counter are host state
    assert T.device_memory_held() >= 0
    try:
        T.set_device_memory_cap(1 << 20)
        T.set_device_memory_cap(0)
    finally:
        T.set_device_memory_cap(0)


def _write(folder, beg, adj):
    os.makedirs(folder, exist_ok=True)
    np.asarray(beg, np.int64).tofile(os.path.join(folder, "begin.bin"))
    np.asarray(adj, np.int32).tofile(os.path.join(folder, "adjacent.bin"))


def _code(fn):
    with pytest.raises(T.TrustError) as e:
        fn()
    return e.value.code


def test_load_missing_folder(tmp_path):
    assert _code(lambda: T.load_folder(str(tmp_path / "nope"))) == T.ERR_IO


def test_load_truncated_begin(tmp_path):
    d = str(tmp_path)
    _write(d, [0, 1, 2], [1, 2])
    with open(os.path.join(d, "begin.bin"), "ab") as f:
        f.write(b"\x00\x01\x02")
    assert _code(lambda: T.load_folder(d)) == T.ERR_FORMAT


def test_load_truncated_adjacent(tmp_path):
    d = str(tmp_path)
    _write(d, [0, 1, 1], [1])
    with open(os.path.join(d, "adjacent.bin"), "ab") as f:
        f.write(b"\x00")
    assert _code(lambda: T.load_folder(d)) == T.ERR_FORMAT


@pytest.mark.parametrize("beg,adj", [
    ([0, 1, 3], [1, 0]),        # last offset != E
    ([1, 1, 2], [0, 1]),        # first offset != 0
    ([0, 2, 1, 2], [1, 2]),     # offsets decrease
    ([0, 1, 2], [1, 2]),        # id >= V
    ([0, 1, 2], [1, -1]),       # negative id
])
def test_validation_rejects(beg, adj, tmp_path):
    assert _code(lambda: T.validate(beg, adj)) == T.ERR_FORMAT
    d = str(tmp_path)
    _write(d, beg, adj)
    assert _code(lambda: T.load_folder(d)) == T.ERR_FORMAT
    # the counting entry point fails the same way before touching a device
    assert _code(lambda: T.triangle_count(0, d)) == T.ERR_FORMAT


def test_drop_in_missing_folder(tmp_path):
    assert _code(lambda: T.triangle_count(0, str(tmp_path / "x"))) == T.ERR_IO
    assert _code(lambda: T.triangle_count_partition(0, str(tmp_path), 8, 0, 0, 0, 2)) == T.ERR_IO


def test_drop_in_partition_rank_layout(tmp_path):
    # total_process must be a multiple of n^3 (one job per (i,j,k) triple)
    assert _code(lambda: T.triangle_count_partition(0, str(tmp_path), 7, 0, 0, 0, 2)) == T.ERR_ARG
    assert _code(lambda: T.triangle_count_partition(9, str(tmp_path), 8, 0, 0, 0, 2)) == T.ERR_ARG


def test_preprocess_rejects_negative_ids():
    assert _code(lambda: T.preprocess_edges([0, -3], [1, 2])) == T.ERR_FORMAT


def test_empty_edge_list():
    beg, adj = T.preprocess_edges(np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert len(beg) == 1 and beg[0] == 0 and len(adj) == 0


def test_save_load_roundtrip(tmp_path):
    beg, adj = T.rmat_csr(9, 8, seed=5)
    T.save_folder(str(tmp_path), beg, adj)
    b2, a2 = T.load_folder(str(tmp_path))
    assert np.array_equal(beg, b2) and np.array_equal(adj, a2)


def test_cli_usage_and_errors(tmp_path):
    exe = os.path.join(T.BIN_DIR, "trianglecounting")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([exe, str(tmp_path / "missing"), "1", "1024", "1024", "1"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "missing" in r.stderr
