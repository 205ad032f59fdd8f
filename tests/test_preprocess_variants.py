"""The reference's other ingest tools: the binary fast path
(Preprocess/speedupIO/TransformTxt2Binary.cpp + preprocess_out.cpp) and the
alternative row orders (Preprocess/otherPreprocess/preprocess.cpp,
preprocess_indegree.cpp).

Golden files under tests/golden/<case>/{speedup,plain,indegree}/ were written
by those reference tools (built from source into oracle/_ref by
tests/golden/make_golden.py); when oracle/_ref is present the tools are also
run on fresh inputs here.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
import trust_amd as T
from conftest import GOLDEN, golden_cases, read_edges

TOOL = os.path.join(T.BIN_DIR, "trust_tool")


def _read(d, name, dtype):
    return np.fromfile(os.path.join(d, name), dtype)


def _bytes(path):
    with open(path, "rb") as f:
        return f.read()


# ------------------------------------------------------------ binary fast path
@pytest.mark.parametrize("name", golden_cases())
def test_txt2bin_and_preprocess_bin_match_golden(name, tmp_path):
    g = os.path.join(GOLDEN, name, "speedup")
    # edges.txt starts with a "#" line: txt2bin skips it (the reference tool
    # cannot, so its golden input had none)
    subprocess.run([TOOL, "txt2bin", os.path.join(GOLDEN, name, "edges.txt")], cwd=tmp_path,
                   check=True, capture_output=True)
    for f in ("prop.txt", "degree", "BinEdgelist"):
        assert _bytes(tmp_path / f) == _bytes(os.path.join(g, f)), f
    subprocess.run([TOOL, "preprocess-bin", str(tmp_path)], check=True, capture_output=True)
    for f in ("begin.bin", "adjacent.bin"):
        assert _bytes(tmp_path / f) == _bytes(os.path.join(g, f)), f


@pytest.mark.parametrize("name", golden_cases())
def test_self_loop_once_api_matches_golden(name):
    src, dst = read_edges(name)
    beg, adj = T.preprocess_edges(src, dst, self_loop_once=True)
    g = os.path.join(GOLDEN, name, "speedup")
    assert beg.tobytes() == _bytes(os.path.join(g, "begin.bin"))
    assert adj.tobytes() == _bytes(os.path.join(g, "adjacent.bin"))


def test_self_loop_rule_keeps_the_count():
    # K4 on 0..3 plus a self-loop on 0: the loop weighs 2 or 1 in vertex 0's
    # sort key and leaves no entry either way (that the weight changes the
    # bytes is pinned by the golden cases with self-loops: tiny, rmat10, er700)
    src = np.array([0, 0, 1, 0, 1, 2, 3], np.int32)
    dst = np.array([1, 2, 2, 0, 3, 3, 0], np.int32)
    b2, a2 = T.preprocess_edges(src, dst)
    b1, a1 = T.preprocess_edges(src, dst, self_loop_once=True)
    assert oracle.count_hash(b1, a1) == oracle.count_hash(b2, a2) == 4
    T.validate(b1, a1)


def test_txt2bin_stream_pairs_across_lines(tmp_path):
    # fscanf("%d%d") pairs tokens regardless of line ends
    (tmp_path / "e.txt").write_text("5 7 7\n9\n5 9 % comment 1 2\n")
    subprocess.run([TOOL, "txt2bin", "e.txt"], cwd=tmp_path, check=True, capture_output=True)
    assert (tmp_path / "prop.txt").read_text() == "3 3"
    assert _read(tmp_path, "degree", np.int32).tolist() == [2, 2, 2]
    assert _read(tmp_path, "BinEdgelist", np.int32).tolist() == [0, 1, 1, 2, 0, 2]


@pytest.mark.parametrize("text", ["1 2 3\n", "1 -2\n", "1 99999999999\n", "1 -\n"])
def test_txt2bin_rejects_bad_input(tmp_path, text):
    (tmp_path / "e.txt").write_text(text)
    r = subprocess.run([TOOL, "txt2bin", "e.txt"], cwd=tmp_path, capture_output=True)
    assert r.returncode != 0 and not (tmp_path / "prop.txt").exists()


def test_preprocess_bin_rejects_short_and_bad_files(tmp_path):
    (tmp_path / "prop.txt").write_text("3 4")
    np.array([0, 1, 1, 2], np.int32).tofile(tmp_path / "BinEdgelist")  # 2 of 3 pairs
    assert subprocess.run([TOOL, "preprocess-bin", str(tmp_path)], capture_output=True).returncode
    np.array([0, 1, 1, 2, 2, 4], np.int32).tofile(tmp_path / "BinEdgelist")  # id 4 >= 4
    assert subprocess.run([TOOL, "preprocess-bin", str(tmp_path)], capture_output=True).returncode
    assert not (tmp_path / "begin.bin").exists()


def test_txt2bin_empty_input(tmp_path):
    (tmp_path / "e.txt").write_text("")
    subprocess.run([TOOL, "txt2bin", "e.txt"], cwd=tmp_path, check=True, capture_output=True)
    assert (tmp_path / "prop.txt").read_text() == "0 0"
    subprocess.run([TOOL, "preprocess-bin", str(tmp_path)], check=True, capture_output=True)
    assert _read(tmp_path, "begin.bin", np.int64).tolist() == [0]
    assert os.path.getsize(tmp_path / "adjacent.bin") == 0


# ------------------------------------------------------------- other orders
@pytest.mark.parametrize("order", ["plain", "indegree"])
@pytest.mark.parametrize("name", golden_cases())
def test_order_api_matches_golden(name, order):
    src, dst = read_edges(name)
    beg, adj = T.preprocess_edges(src, dst, order=order)
    g = os.path.join(GOLDEN, name, order)
    assert beg.tobytes() == _bytes(os.path.join(g, "begin.bin"))
    assert adj.tobytes() == _bytes(os.path.join(g, "adjacent.bin"))


@pytest.mark.parametrize("order", ["plain", "indegree"])
@pytest.mark.parametrize("name", ["tiny", "rmat10", "k120"])
def test_order_tool_matches_golden(name, order, tmp_path):
    subprocess.run([TOOL, "undirect", os.path.join(GOLDEN, name, "edges.txt"),
                    str(tmp_path / "1.mmio")], check=True)
    subprocess.run([TOOL, "preprocess", str(tmp_path / "1.mmio"), str(tmp_path), "--order", order],
                   check=True)
    g = os.path.join(GOLDEN, name, order)
    for f in ("begin.bin", "adjacent.bin", "edge"):
        assert _bytes(tmp_path / f) == _bytes(os.path.join(g, f)), f


@pytest.mark.parametrize("seed", range(3))
def test_orders_are_valid_and_count_alike(seed):
    src, dst = T.gen_rmat(11, 8, seed=seed)
    counts = []
    for order in ("outdegree", "plain", "indegree"):
        beg, adj = T.preprocess_edges(src, dst, order=order)
        T.validate(beg, adj)
        d = np.diff(beg)
        cls = np.where(d > 100, 0, np.where(d >= 2, 1, 2))
        assert np.all(np.diff(cls) >= 0)  # class ranges contiguous in every order
        if order == "plain":
            assert np.all(np.diff(d) <= 0)  # rows by descending out-degree
        counts.append(oracle.count_hash(beg, adj))
    assert counts[0] == counts[1] == counts[2]
    ids, inv = np.unique(np.stack([src, dst], 1), return_inverse=True)
    inv = inv.reshape(-1, 2)
    assert counts[0] == oracle.count_brute(len(ids), inv[:, 0], inv[:, 1])


def test_indegree_keeps_ascending_degree_inside_classes():
    # inside each class the rows keep the ascending-degree order, so the
    # undirected degree (in + out) is non-decreasing within a class; without
    # self-loops (which weigh 2 in the sort key and leave no entry) that
    # degree is exactly the sort key
    src, dst = T.gen_barabasi_albert(3000, 5, seed=4)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    beg, adj = T.preprocess_edges(src, dst, order="indegree")
    V = len(beg) - 1
    out = np.diff(beg)
    deg = out + np.bincount(adj, minlength=V)
    cls = np.where(out > 100, 0, np.where(out >= 2, 1, 2))
    for c in range(3):
        assert np.all(np.diff(deg[cls == c]) >= 0)


def test_unknown_order_and_flags_rejected():
    src = np.array([0, 1], np.int32)
    dst = np.array([1, 2], np.int32)
    with pytest.raises(ValueError):
        T.preprocess_edges(src, dst, order="bogus")
    import ctypes as C
    c = T.Csr()
    ps = src.ctypes.data_as(C.POINTER(C.c_int32))
    pd = dst.ctypes.data_as(C.POINTER(C.c_int32))
    assert T.lib().trust_preprocess_edges_ex(ps, pd, 2, 7, 0, C.byref(c)) == T.ERR_ARG
    assert T.lib().trust_preprocess_edges_ex(ps, pd, 2, 0, 2, C.byref(c)) == T.ERR_ARG


# ------------------------------------------------ fresh inputs vs the reference
needs_ref = pytest.mark.skipif(oracle.ref_tool("preprocess_out") is None,
                               reason="reference tools not built (no reference tree)")


def _fresh(kind, seed):
    if kind == "rmat":
        return T.gen_rmat(11, 8, seed=seed)
    if kind == "er":
        return T.gen_erdos_renyi(3000, 20000, seed=seed)
    rng = np.random.default_rng(seed)  # dense: many rows above the class boundary
    return (rng.integers(0, 400, 60000).astype(np.int32),
            rng.integers(0, 400, 60000).astype(np.int32))


def _ref_run(cmd, cwd, outputs):
    # otherPreprocess/preprocess.cpp can abort at exit (heap damage in its
    # averagedegree(), :136-152) after its files are closed
    r = subprocess.run(cmd, cwd=cwd, capture_output=True)
    assert r.returncode == 0 or all(os.path.exists(os.path.join(cwd, f)) for f in outputs)


@needs_ref
@pytest.mark.parametrize("kind,seed", [("rmat", 11), ("er", 12), ("dense", 13)])
def test_fresh_inputs_match_reference_tools(kind, seed):
    src, dst = _fresh(kind, seed)
    ref, ours = tempfile.mkdtemp(), tempfile.mkdtemp()
    text = "".join(f"{a} {b}\n" for a, b in zip(src.tolist(), dst.tolist()))
    for d in (ref, ours):
        with open(os.path.join(d, "e.txt"), "w") as f:
            f.write(text)
    subprocess.run([oracle.ref_tool("TransformTxt2Binary"), "e.txt"], cwd=ref, check=True,
                   capture_output=True)
    subprocess.run([oracle.ref_tool("preprocess_out")], cwd=ref, check=True, capture_output=True)
    subprocess.run([TOOL, "txt2bin", "e.txt"], cwd=ours, check=True, capture_output=True)
    subprocess.run([TOOL, "preprocess-bin"], cwd=ours, check=True, capture_output=True)
    for f in ("prop.txt", "degree", "BinEdgelist", "begin.bin", "adjacent.bin"):
        assert _bytes(os.path.join(ref, f)) == _bytes(os.path.join(ours, f)), f
    subprocess.run([oracle.ref_tool("fromDirectToUndirect"), "e.txt"], cwd=ref, check=True,
                   capture_output=True)
    for tool, order in (("preprocess_plain", "plain"), ("preprocess_indegree", "indegree")):
        r = tempfile.mkdtemp()
        os.rename(os.path.join(ref, "1.mmio"), os.path.join(r, "1.mmio"))
        _ref_run([oracle.ref_tool(tool)], r, ("begin.bin", "adjacent.bin", "edge"))
        o = tempfile.mkdtemp()
        subprocess.run([TOOL, "preprocess", os.path.join(r, "1.mmio"), o, "--order", order],
                       check=True, capture_output=True)
        for f in ("begin.bin", "adjacent.bin", "edge"):
            assert _bytes(os.path.join(r, f)) == _bytes(os.path.join(o, f)), (order, f)
        os.rename(os.path.join(r, "1.mmio"), os.path.join(ref, "1.mmio"))


def test_txt2bin_many_blocks_against_restatement(tmp_path):
    # more pairs than one storage block (2^22): ids compacted in ascending
    # order, degrees and remapped pairs as TransformTxt2Binary.cpp:97-152 write them
    rng = np.random.default_rng(5)
    m = (1 << 22) + 1_000_003
    ids = rng.integers(0, 3_000_000, 2 * m) * 3 + 7  # sparse ids, some never used
    with open(tmp_path / "e.txt", "w") as f:
        f.write("\n".join(f"{ids[2 * i]} {ids[2 * i + 1]}" for i in range(m)))
    subprocess.run([TOOL, "txt2bin", "e.txt"], cwd=tmp_path, check=True, capture_output=True,
                   timeout=300)
    uniq, inv, cnt = np.unique(ids, return_inverse=True, return_counts=True)
    assert (tmp_path / "prop.txt").read_text().split() == [str(m), str(len(uniq))]
    assert np.array_equal(np.fromfile(tmp_path / "degree", np.int32), cnt.astype(np.int32))
    assert np.array_equal(np.fromfile(tmp_path / "BinEdgelist", np.int32), inv.astype(np.int32))


def test_undirect_many_blocks_against_restatement(tmp_path):
    # more pairs than one storage block (2^22), with repeats in both directions
    # and self-loops: 1.mmio = "k k m", then the unique pairs (min, max) of the
    # 1-based compacted ids in ascending order
    rng = np.random.default_rng(6)
    m = (1 << 22) + 777_777
    ids = rng.integers(0, 400_000, 2 * m) * 2 + 1
    with open(tmp_path / "e.txt", "w") as f:
        f.write("\n".join(f"{ids[2 * i]} {ids[2 * i + 1]}" for i in range(m)))
    subprocess.run([TOOL, "undirect", "e.txt", "1.mmio"], cwd=tmp_path, check=True,
                   capture_output=True, timeout=300)
    uniq, inv = np.unique(ids, return_inverse=True)
    p = (inv + 1).reshape(-1, 2)
    p = np.unique(np.stack([p.min(axis=1), p.max(axis=1)], axis=1), axis=0)
    lines = (tmp_path / "1.mmio").read_text().split("\n")
    assert lines[0].split() == [str(len(uniq)), str(len(uniq)), str(len(p))]
    got = np.array(" ".join(lines[1:]).split(), np.int64).reshape(-1, 2)
    assert np.array_equal(got, p)


@pytest.mark.parametrize("text", ["3 -5\n", "3 -2000000000\n", "1 2\n4 -7\n"])
def test_undirect_rejects_negative_ids(tmp_path, text):
    # the reference's fromDirectToUndirect indexes its arrays with the id
    # (undefined behaviour); the tool must stop with an error and no output
    (tmp_path / "e.txt").write_text(text)
    r = subprocess.run([TOOL, "undirect", "e.txt", "1.mmio"], cwd=tmp_path, capture_output=True,
                       timeout=60)
    assert r.returncode != 0
    assert b"negative vertex id" in r.stderr
    assert not (tmp_path / "1.mmio").exists()


@pytest.mark.parametrize("body", ["1 0 1\n", "0 2 1\n", "2 -3 1\n", "1 -2147483648 1\n"])
def test_preprocess_rejects_ids_below_one(tmp_path, body):
    # 1.mmio ids are 1-based (preprocess_outdegree.cpp:73-74 decrements them)
    (tmp_path / "1.mmio").write_text("3 3 1\n" + body)
    r = subprocess.run([TOOL, "preprocess", "1.mmio", "."], cwd=tmp_path, capture_output=True,
                       timeout=60)
    assert r.returncode != 0
    assert b"ids start at 1" in r.stderr
    assert not (tmp_path / "begin.bin").exists()
