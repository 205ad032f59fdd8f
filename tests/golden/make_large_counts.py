"""Writes the oracle's counts of the full-size workloads to fixtures.

    python tests/golden/make_large_counts.py [--only NAME ...] [--threads N]

Runs on the host only (no GPU).  Each workload is generated with the
package's seeded generators and preprocessed with its reference-contract
preprocessing, exactly as the GPU tests and bench.py build it; the CPU
oracle (oracle/tc_oracle.c, the C/OpenMP restatement of the reference's
hashed loop) then counts it.  Written:

  large_counts.json      per workload: V, E, a fingerprint of the CSR
                         (xxh64 of adjacent.bin and of begin.bin), the
                         oracle's triangle count, and its probe totals for
                         the reference walk and the short-side walk
  batch10000_counts.npy  uint64 per-graph oracle counts of the 10,000-graph
                         ER/BA batch (graph order of batch_generate), made
                         one graph at a time (run this step with
                         OMP_NUM_THREADS=1: graphs run in parallel threads)

The fingerprint lets a test tell a changed generator or preprocess (which
would make the committed count meaningless) from a wrong GPU count.  The
R-MAT scale-26 count takes about 25 minutes on 8 host threads.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OUT_JSON = os.path.join(HERE, "large_counts.json")
OUT_BATCH = os.path.join(HERE, "batch10000_counts.npy")


def fingerprint(beg, adj):
    import xxhash

    return {"adj_xxh64": xxhash.xxh64(memoryview(np.ascontiguousarray(adj))).hexdigest(),
            "beg_xxh64": xxhash.xxh64(memoryview(np.ascontiguousarray(beg))).hexdigest()}


def build(name):
    """The CSR of a named workload (the same calls the GPU tests make)."""
    import trust_amd as T

    if name.startswith("rmat"):
        scale = int(name[4:6])
        return T.rmat_csr(scale, 16, seed=1)
    if name == "chunglu3.8M":
        return T.preprocess_edges(*T.gen_chung_lu(3_800_000, 16_500_000, 2.5, 800.0, seed=1))
    raise ValueError(name)


def mix64(x):
    """The batch generator's 64-bit mixer (trust_amd/csrc/batch.cpp)."""
    m = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & m
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
    return x ^ (x >> 31)


def batch_counts(threads, ngraphs=10000, seed=1):
    """Per-graph oracle counts of trust_batch_generate(ngraphs, seed, 1000,
    200000, 8, 64), one graph at a time: graph g is generated from the seed
    batch.cpp derives for it, preprocessed and counted alone, so the whole
    batch (about 30 GB) is never held.  The fingerprints are those of the
    batch's concatenated CSR (global ids and offsets), fed graph by graph."""
    from concurrent.futures import ThreadPoolExecutor

    import xxhash

    import oracle
    import trust_amd as T

    n, d, kind = T.batch_shape(ngraphs, seed, 1000, 200000, 8.0, 64.0)

    def one(g):
        gs = mix64(seed ^ mix64(g + 0x51ED2701))
        if kind[g] == 0:
            src, dst = T.gen_erdos_renyi(int(n[g]), max(1, int(round(n[g] * d[g] / 2.0))), gs)
        else:
            src, dst = T.gen_barabasi_albert(int(n[g]), max(1, int(round(d[g] / 2.0))), gs)
        beg, adj = T.preprocess_edges(src, dst)
        return beg, adj, oracle.count_hash(beg, adj, 1)

    h_adj, h_beg, h_vb = xxhash.xxh64(), xxhash.xxh64(), xxhash.xxh64()
    counts = np.zeros(ngraphs, np.uint64)
    v0 = e0 = 0
    h_vb.update(np.int64(0).tobytes())
    workers = threads if threads > 0 else os.cpu_count()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(workers) as pool:
        for c0 in range(0, ngraphs, 200):
            for g, (beg, adj, c) in zip(range(c0, min(ngraphs, c0 + 200)),
                                        pool.map(one, range(c0, min(ngraphs, c0 + 200)))):
                counts[g] = c
                h_beg.update(np.ascontiguousarray(beg[:-1] + e0).tobytes())
                h_adj.update(np.ascontiguousarray(adj + np.int32(v0)).tobytes())
                v0 += len(beg) - 1
                e0 += len(adj)
                h_vb.update(np.int64(v0).tobytes())
            print(f"batch10000: graphs [0, {min(ngraphs, c0 + 200)}) "
                  f"{time.perf_counter() - t0:.0f}s", flush=True)
    h_beg.update(np.int64(e0).tobytes())
    meta = {
        "workload": "10,000 graphs, half Erdos-Renyi G(n, n*d/2), half Barabasi-Albert "
                    "k = round(d/2); n log-uniform [1000, 200000], d uniform [8, 64], seed 1",
        "V": int(v0), "E": int(e0), "adj_xxh64": h_adj.hexdigest(), "beg_xxh64": h_beg.hexdigest(),
        "vbase_xxh64": h_vb.hexdigest(), "triangles": int(counts.sum(dtype=np.uint64)),
        "per_graph_counts": os.path.basename(OUT_BATCH),
    }
    return counts, meta


WORKLOADS = {
    "rmat16": "R-MAT scale 16, edge factor 16, (a,b,c) = (.57,.19,.19), seed 1",
    "chunglu3.8M": "Chung-Lu n = 3.8M, 16.5M pairs, gamma 2.5, max expected degree 800, seed 1",
    "rmat24": "R-MAT scale 24, edge factor 16, (a,b,c) = (.57,.19,.19), seed 1",
    "rmat26": "R-MAT scale 26, edge factor 16, (a,b,c) = (.57,.19,.19), seed 1",
}


def main():
    import oracle

    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*", default=None,
                    help="workloads to (re)compute: " + " ".join(list(WORKLOADS) + ["batch10000"]))
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    todo = args.only if args.only else list(WORKLOADS) + ["batch10000"]
    data = json.load(open(OUT_JSON)) if os.path.exists(OUT_JSON) else {}
    for name in todo:
        t0 = time.perf_counter()
        if name == "batch10000":
            counts, meta = batch_counts(args.threads)
            np.save(OUT_BATCH, counts)
            data[name] = dict(meta, generate_and_oracle_s=round(time.perf_counter() - t0, 1))
        else:
            beg, adj = build(name)
            t1 = time.perf_counter()
            tri = oracle.count_hash(beg, adj, args.threads)
            t2 = time.perf_counter()
            data[name] = {
                "workload": WORKLOADS[name],
                "V": int(len(beg) - 1), "E": int(len(adj)), **fingerprint(beg, adj),
                "triangles": int(tri),
                "probes_reference_walk": int(oracle.probes(beg, adj)),
                "probes_short_walk": int(oracle.probes_short(beg, adj)),
                "generate_s": round(t1 - t0, 1), "oracle_s": round(t2 - t1, 1),
                "oracle_threads": oracle.max_threads() if args.threads <= 0 else args.threads,
            }
            del beg, adj
        print(name, json.dumps(data[name]), flush=True)
        with open(OUT_JSON, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
            f.write("\n")


if __name__ == "__main__":
    main()
