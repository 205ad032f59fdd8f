"""Input contract: edge list -> oriented CSR, byte-identical to the reference's
Preprocess/fromDirectToUndirect + Preprocess/preprocess_outdegree.

Golden files were written by the reference tools themselves (compiled from
source into oracle/_ref by tests/golden/make_golden.py); when oracle/_ref is
present the tools are also run on fresh inputs here.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
import trust_amd as T
from conftest import GOLDEN, golden_cases, load_golden, read_edges


@pytest.mark.parametrize("name", golden_cases())
def test_byte_identical_to_reference_golden(name):
    beg_ref, adj_ref, meta = load_golden(name)
    src, dst = read_edges(name)
    beg, adj = T.preprocess_edges(src, dst)
    assert beg.tobytes() == beg_ref.tobytes()
    assert adj.tobytes() == adj_ref.tobytes()


def _run_ref_tools(src, dst):
    und, pre = oracle.ref_tool("fromDirectToUndirect"), oracle.ref_tool("preprocess")
    d = tempfile.mkdtemp()
    with open(os.path.join(d, "e.txt"), "w") as f:
        f.write("# header line\n")
        f.writelines(f"{a} {b}\n" for a, b in zip(src.tolist(), dst.tolist()))
    subprocess.run([und, "e.txt"], cwd=d, check=True, capture_output=True)
    subprocess.run([pre], cwd=d, check=True, capture_output=True)
    return (np.fromfile(os.path.join(d, "begin.bin"), np.int64),
            np.fromfile(os.path.join(d, "adjacent.bin"), np.int32), d)


needs_ref = pytest.mark.skipif(oracle.ref_tool("preprocess") is None,
                               reason="reference tools not built (no reference tree)")


@needs_ref
@pytest.mark.parametrize("kind,seed", [("rmat", 1), ("rmat", 2), ("er", 3), ("ba", 4),
                                       ("chunglu", 5), ("dense", 6)])
def test_byte_identical_to_reference_tools(kind, seed):
    if kind == "rmat":
        src, dst = T.gen_rmat(11, 8, seed=seed)
    elif kind == "er":
        src, dst = T.gen_erdos_renyi(3000, 20000, seed=seed)
    elif kind == "ba":
        src, dst = T.gen_barabasi_albert(3000, 6, seed=seed)
    elif kind == "chunglu":
        src, dst = T.gen_chung_lu(5000, 20000, 2.5, 300.0, seed=seed)
    else:  # dense: many vertices above the 100 class boundary
        rng = np.random.default_rng(seed)
        src = rng.integers(0, 400, 60000).astype(np.int32)
        dst = rng.integers(0, 400, 60000).astype(np.int32)
    beg, adj = T.preprocess_edges(src, dst)
    rb, ra, _ = _run_ref_tools(src, dst)
    assert beg.tobytes() == rb.tobytes()
    assert adj.tobytes() == ra.tobytes()


@pytest.mark.parametrize("seed", range(4))
def test_contract_invariants(seed):
    rng = np.random.default_rng(seed)
    n = 300
    src = rng.integers(0, n, 4000).astype(np.int32)
    dst = rng.integers(0, n, 4000).astype(np.int32)
    beg, adj = T.preprocess_edges(src, dst)
    V = len(beg) - 1
    T.validate(beg, adj)
    present = np.unique(np.concatenate([src, dst]))
    assert V == len(present)  # every id that appears (incl. self-loop-only) is kept
    d = np.diff(beg)
    u = np.repeat(np.arange(V), d)
    assert np.all(u != adj)  # no self-loops
    pairs = np.sort(np.stack([u, adj], 1), axis=1)
    keys = pairs[:, 0] * V + pairs[:, 1]
    assert len(np.unique(keys)) == len(keys)  # each undirected edge once
    und = {(min(a, b), max(a, b)) for a, b in zip(src.tolist(), dst.tolist()) if a != b}
    assert len(keys) == len(und)
    # classes are contiguous: d+ > 100, then 2..100, then < 2
    cls = np.where(d > 100, 0, np.where(d >= 2, 1, 2))
    assert np.all(np.diff(cls) >= 0)
    # count preserved
    ids, inv = np.unique(np.stack([src, dst], 1), return_inverse=True)
    inv = inv.reshape(-1, 2)
    assert oracle.count_hash(beg, adj) == oracle.count_brute(len(ids), inv[:, 0], inv[:, 1])


def test_generators_deterministic():
    a = T.gen_rmat(10, 4, seed=9)
    b = T.gen_rmat(10, 4, seed=9)
    c = T.gen_rmat(10, 4, seed=10)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])
    for gen in (lambda: T.gen_chung_lu(1000, 5000, seed=2), lambda: T.gen_erdos_renyi(500, 3000, 2),
                lambda: T.gen_barabasi_albert(500, 3, 2)):
        x, y = gen(), gen()
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])


def test_rmat_quadrant_frequencies():
    # first-level quadrant frequencies match (a, b, c, d) = (.57,.19,.19,.05)
    s, d = T.gen_rmat(8, 64, seed=1)
    top_u, top_v = s >> 7, d >> 7
    q = np.bincount(top_u * 2 + top_v, minlength=4) / len(s)
    assert np.allclose(q, [0.57, 0.19, 0.19, 0.05], atol=0.01)


def test_chung_lu_max_degree_target():
    s, d = T.gen_chung_lu(200000, 800000, 2.5, 800.0, seed=1)
    deg = np.bincount(np.concatenate([s, d]), minlength=200000)
    # vertex 0 has expected degree 800 (before dedup)
    assert 650 < deg[0] < 950
    assert deg.max() < 1000


def test_barabasi_albert_shape():
    s, d = T.gen_barabasi_albert(5000, 4, seed=1)
    assert len(s) == 5000 * 4
    assert np.all(d <= s)  # links only to earlier (or same) vertices


def test_tool_pipeline_matches_golden(tmp_path):
    tool = os.path.join(T.BIN_DIR, "trust_tool")
    src = os.path.join(GOLDEN, "rmat10", "edges.txt")
    subprocess.run([tool, "undirect", src, str(tmp_path / "1.mmio")], check=True)
    subprocess.run([tool, "preprocess", str(tmp_path / "1.mmio"), str(tmp_path)], check=True)
    beg_ref, adj_ref, _ = load_golden("rmat10")
    assert (tmp_path / "begin.bin").read_bytes() == beg_ref.tobytes()
    assert (tmp_path / "adjacent.bin").read_bytes() == adj_ref.tobytes()


@needs_ref
def test_tool_undirect_matches_reference_mmio(tmp_path):
    tool = os.path.join(T.BIN_DIR, "trust_tool")
    src = os.path.join(GOLDEN, "tiny", "edges.txt")
    subprocess.run([tool, "undirect", src, str(tmp_path / "ours.mmio")], check=True)
    subprocess.run([oracle.ref_tool("fromDirectToUndirect"), src, str(tmp_path / "ref.mmio")],
                   check=True, capture_output=True)
    assert (tmp_path / "ours.mmio").read_text() == (tmp_path / "ref.mmio").read_text()
