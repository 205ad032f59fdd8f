"""Writes the golden fixtures under tests/golden/<case>/ (run in a container
that has the reference tree; the fixtures are committed and the tests only
read them).

For each case:
  edges.txt          input edge list (reference text format: "u v" lines)
  begin.bin,
  adjacent.bin       output of the reference's own Preprocess/fromDirectToUndirect
                     + Preprocess/preprocess_outdegree, built from source into
                     oracle/_ref by oracle/Makefile
  part{n}/partition{i}_{j}/{begin.bin,adjacent.bin,map,division}
                     output of the reference's Preprocess/partition for n = 2, 3
  speedup/{prop.txt,degree,BinEdgelist,begin.bin,adjacent.bin}
                     output of Preprocess/speedupIO/TransformTxt2Binary (fed the
                     pairs without the comment line, which that tool cannot
                     skip) and speedupIO/preprocess_out
  plain/{begin.bin,adjacent.bin,edge}, indegree/{begin.bin,adjacent.bin,edge}
                     output of Preprocess/otherPreprocess/preprocess.cpp and
                     preprocess_indegree.cpp on the same 1.mmio
  expect.json        V, E, and the triangle count from two sources independent
                     of this project: networkx (sum(triangles)/3) and the
                     brute-force bitset counter on the raw edge list

Usage:  python tests/golden/make_golden.py
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
import trust_amd as T  # noqa: E402


def cases():
    rng = np.random.default_rng(2024)
    out = {}
    # hand-made: self-loops, duplicates, both directions, id gaps, isolated ids
    tiny = [(1, 2), (2, 1), (2, 3), (3, 1), (3, 3), (5, 5), (10, 11), (11, 12), (12, 10),
            (10, 12), (1, 10), (2, 10), (3, 10), (20, 21), (40, 40), (7, 2), (7, 3), (7, 1)]
    out["tiny"] = np.array(tiny, np.int64)
    s, d = T.gen_rmat(10, 8, seed=7)
    out["rmat10"] = np.stack([s, d], 1)
    n = 120
    out["k120"] = np.array([(i, j) for i in range(n) for j in range(i + 1, n)], np.int64)
    # wheel (50 rim vertices + hub 0) plus a friendship graph F_20 on other ids
    wheel = [(0, i) for i in range(1, 51)] + [(i, i % 50 + 1) for i in range(1, 51)]
    fr = []
    for t in range(20):
        a, b = 1000 + 2 * t, 1001 + 2 * t
        fr += [(999, a), (999, b), (a, b)]
    out["wheel50_f20"] = np.array(wheel + fr, np.int64)
    s, d = T.gen_barabasi_albert(2000, 4, seed=3)
    out["ba2k"] = np.stack([s, d], 1)
    # Erdos-Renyi with shuffled ids and reversed duplicates
    s, d = T.gen_erdos_renyi(700, 6000, seed=5)
    perm = rng.permutation(5000)[:700]
    e = np.stack([perm[s], perm[d]], 1)
    out["er700"] = np.concatenate([e, e[:500, ::-1]])
    return out


def run_ref(edges, workdir):
    ref = oracle.ref_tool("preprocess")
    und = oracle.ref_tool("fromDirectToUndirect")
    part = oracle.ref_tool("partition")
    if not (ref and und and part):
        raise SystemExit("oracle/_ref tools missing: run `make -C oracle` with the reference present")
    with open(os.path.join(workdir, "edges.txt"), "w") as f:
        f.write("# FromNodeId\tToNodeId\n")
        for a, b in edges:
            f.write(f"{a}\t{b}\n")
    subprocess.run([und, "edges.txt"], cwd=workdir, check=True, capture_output=True)
    subprocess.run([ref], cwd=workdir, check=True, capture_output=True)
    for n in (2, 3):
        pd = os.path.join(workdir, f"part{n}")
        os.makedirs(pd)
        shutil.copy(os.path.join(workdir, "begin.bin"), pd)
        shutil.copy(os.path.join(workdir, "adjacent.bin"), pd)
        for i in range(n):
            for j in range(n):
                os.makedirs(os.path.join(pd, f"partition{i}_{j}"))
        subprocess.run([part, str(n)], cwd=pd, check=True, capture_output=True)
        os.remove(os.path.join(pd, "begin.bin"))
        os.remove(os.path.join(pd, "adjacent.bin"))


def _run_allow_exit_abort(cmd, cwd, outputs):
    """otherPreprocess/preprocess.cpp corrupts its heap in averagedegree()
    (memset of sizeof(pointer) bytes, then unbounded reads of the
    uninitialised counts, :136-152) and can abort at exit, after computeCSR
    has written and closed its files.  Accept that abort when every output
    exists."""
    r = subprocess.run(cmd, cwd=cwd, capture_output=True)
    if r.returncode != 0 and not all(os.path.exists(os.path.join(cwd, f)) for f in outputs):
        raise SystemExit(f"{cmd} failed with {r.returncode}")


def run_variants(edges, workdir):
    t2b, pout = oracle.ref_tool("TransformTxt2Binary"), oracle.ref_tool("preprocess_out")
    plain, indeg = oracle.ref_tool("preprocess_plain"), oracle.ref_tool("preprocess_indegree")
    if not (t2b and pout and plain and indeg):
        raise SystemExit("oracle/_ref tools missing: run `make -C oracle` with the reference present")
    sd = os.path.join(workdir, "speedup")
    os.makedirs(sd)
    with open(os.path.join(sd, "pairs.txt"), "w") as f:
        for a, b in edges:
            f.write(f"{a}\t{b}\n")
    subprocess.run([t2b, "pairs.txt", str(len(edges))], cwd=sd, check=True, capture_output=True)
    subprocess.run([pout], cwd=sd, check=True, capture_output=True)
    os.remove(os.path.join(sd, "pairs.txt"))
    for name, tool in (("plain", plain), ("indegree", indeg)):
        d = os.path.join(workdir, name)
        os.makedirs(d)
        shutil.copy(os.path.join(workdir, "1.mmio"), d)
        _run_allow_exit_abort([tool], d, ("begin.bin", "adjacent.bin", "edge"))
        os.remove(os.path.join(d, "1.mmio"))


def main():
    import networkx as nx

    for name, edges in cases().items():
        dst = os.path.join(HERE, name)
        if os.path.exists(dst):
            shutil.rmtree(dst)
        work = tempfile.mkdtemp()
        run_ref(edges, work)
        run_variants(edges, work)
        os.makedirs(dst)
        for fn in ("edges.txt", "begin.bin", "adjacent.bin"):
            shutil.copy(os.path.join(work, fn), dst)
        for sub in ("part2", "part3", "speedup", "plain", "indegree"):
            shutil.copytree(os.path.join(work, sub), os.path.join(dst, sub))
        shutil.rmtree(work)
        g = nx.Graph()
        g.add_edges_from((int(a), int(b)) for a, b in edges if a != b)
        nx_count = sum(nx.triangles(g).values()) // 3
        ids, inv = np.unique(edges, return_inverse=True)
        inv = inv.reshape(-1, 2)
        brute = oracle.count_brute(len(ids), inv[:, 0], inv[:, 1])
        beg = np.fromfile(os.path.join(dst, "begin.bin"), np.int64)
        adj = np.fromfile(os.path.join(dst, "adjacent.bin"), np.int32)
        meta = {"V": int(len(beg) - 1), "E": int(len(adj)), "triangles_networkx": int(nx_count),
                "triangles_brute": int(brute)}
        with open(os.path.join(dst, "expect.json"), "w") as f:
            json.dump(meta, f, indent=1)
        print(name, meta)


if __name__ == "__main__":
    main()
