import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP device 0)")


def golden_cases():
    return sorted(d for d in os.listdir(GOLDEN)
                  if os.path.isfile(os.path.join(GOLDEN, d, "expect.json")))


def load_golden(name):
    d = os.path.join(GOLDEN, name)
    beg = np.fromfile(os.path.join(d, "begin.bin"), np.int64)
    adj = np.fromfile(os.path.join(d, "adjacent.bin"), np.int32)
    with open(os.path.join(d, "expect.json")) as f:
        meta = json.load(f)
    return beg, adj, meta


def read_edges(name):
    src, dst = [], []
    with open(os.path.join(GOLDEN, name, "edges.txt")) as f:
        for line in f:
            if not line or not line[0].isdigit():
                continue
            a, b = line.split()[:2]
            src.append(int(a))
            dst.append(int(b))
    return np.array(src, np.int32), np.array(dst, np.int32)


# ---- closed-form graphs (undirected edge lists) --------------------------
def complete(n, base=0):
    return [(base + i, base + j) for i in range(n) for j in range(i + 1, n)]


def complete_bipartite(a, b, base=0):
    return [(base + i, base + a + j) for i in range(a) for j in range(b)]


def cycle(n, base=0):
    return [(base + i, base + (i + 1) % n) for i in range(n)]


def wheel(n, base=0):
    return [(base, base + 1 + i) for i in range(n)] + [(base + 1 + i, base + 1 + (i + 1) % n)
                                                       for i in range(n)]


def friendship(k, base=0):
    e = []
    for t in range(k):
        a, b = base + 1 + 2 * t, base + 2 + 2 * t
        e += [(base, a), (base, b), (a, b)]
    return e


def complete_tripartite(a, b, c, base=0):
    A = range(base, base + a)
    B = range(base + a, base + a + b)
    Cc = range(base + a + b, base + a + b + c)
    return ([(x, y) for x in A for y in B] + [(x, y) for x in A for y in Cc] +
            [(x, y) for x in B for y in Cc])


CLOSED_FORMS = {
    "K3": (complete(3), 1),
    "K10": (complete(10), 120),
    "K50": (complete(50), 19600),
    "K150": (complete(150), 551300),
    "K40_60": (complete_bipartite(40, 60), 0),
    "C3": (cycle(3), 1),
    "C9": (cycle(9), 0),
    "W30": (wheel(30), 30),
    "F25": (friendship(25), 25),
    "K5_6_7": (complete_tripartite(5, 6, 7), 210),
    "union": (complete(12) + wheel(20, base=100) + friendship(7, base=300) +
              complete_tripartite(3, 4, 5, base=500), 220 + 20 + 7 + 60),
}


# Share of the reference walk's probes at or below which a count takes the
# short-side walk by default (trust_amd/csrc/tc_kernels.h kShortWalkShare).
SHORT_WALK_SHARE = 0.45


def expected_probes(beg, adj, flags=0, per_vertex=False):
    """Probes the GPU count makes: the short-side walk's when forced, or when
    it makes at most SHORT_WALK_SHARE of the reference walk's; else the
    reference walk's."""
    import oracle
    import trust_amd as T
    ref, short = oracle.probes(beg, adj), oracle.probes_short(beg, adj)
    use_short = bool(flags & T.FLAG_SHORT_WALK) or (
        not flags & T.FLAG_REFERENCE_WALK and short <= SHORT_WALK_SHARE * ref)
    if per_vertex:
        return (oracle.probes_short if use_short else oracle.probes)(beg, adj, per_vertex=True)
    return short if use_short else ref


def closed_form_csr(name):
    import trust_amd as T

    e, want = CLOSED_FORMS[name]
    e = np.array(e, np.int32)
    beg, adj = T.preprocess_edges(e[:, 0], e[:, 1])
    return beg, adj, want


@pytest.fixture(scope="session")
def gpu():
    import trust_amd as T

    n = T.device_count()
    if n < 1:
        pytest.fail("no HIP device: gpu tests need an MI355X")
    return 0
