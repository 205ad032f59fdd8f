"""GPU tests of the launch autotuner: the parameters it returns count
exactly and are no slower than the defaults it started from."""
import pytest

import oracle
import trust_amd as T

pytestmark = pytest.mark.gpu


def test_autotune_rmat(gpu):
    beg, adj = T.rmat_csr(16, 16, seed=1)
    want = oracle.count_hash(beg, adj)
    with T.Context(gpu) as c:
        c.load(beg, adj)
        base = min(c.count().seconds for _ in range(3))
        tuned, sec, n = c.autotune(reps=2, max_seconds=30)
        print(f"[gpu] autotune rmat16: {tuned} {sec * 1e3:.3f} ms "
              f"(default {base * 1e3:.3f} ms, {n} candidates)", flush=True)
        assert set(tuned) == set(T.TUNED_FIELDS) and n >= 2
        r = min((c.count(**tuned) for _ in range(3)), key=lambda x: x.seconds)
        assert r.triangles == want
        assert r.seconds <= 1.25 * base + 20e-6


def test_autotune_respects_budget_and_base(gpu):
    beg, adj = T.rmat_csr(12, 8, seed=5)
    with T.Context(gpu) as c:
        c.load(beg, adj)
        tuned, sec, n = c.autotune(reps=1, max_seconds=1e-9, threads=512, chunk=3)
        # no time for any candidate but the starting point
        assert n == 1 and tuned["threads"] == 512 and tuned["chunk"] == 3
        with pytest.raises(T.TrustError):
            c.autotune(reps=0)
