"""GPU tests of the learned cost scheduler: the MLP kernels against the float64
oracle, GPU training, the feature kernel, and counts under the learned
schedule (which must equal the degree split's, bit for bit)."""
import numpy as np
import pytest

import oracle
import trust_amd as T
from conftest import expected_probes

pytestmark = pytest.mark.gpu


def test_param_count_and_init():
    P = T.mlp_param_count()
    assert P == 32 * 32 + 6 * 32 + 1
    a, b = T.mlp_init(7), T.mlp_init(7)
    assert np.array_equal(a, b) and not np.array_equal(a, T.mlp_init(8))


def test_forward_matches_oracle(gpu):
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, T.mlp_param_count()).astype(np.float32)
    x = rng.uniform(0, 1, (5000, 3)).astype(np.float32)
    y = T.mlp_forward(p, x, device=gpu)
    want = oracle.mlp_forward(p.astype(np.float64), x.astype(np.float64))
    assert np.allclose(y, want, rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("n", [1, 63, 64, 300, 4096])
def test_gradient_matches_oracle(gpu, n):
    rng = np.random.default_rng(n)
    p = rng.normal(0, 0.3, T.mlp_param_count()).astype(np.float32)
    x = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    t = rng.uniform(0, 1, n).astype(np.float32)
    g, loss = T.mlp_loss_grad(p, x, t, device=gpu)
    gw, lw = oracle.mlp_loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    scale = np.abs(gw).max()
    assert np.max(np.abs(g - gw)) <= 1e-4 * scale + 1e-6
    assert np.isclose(loss, lw, rtol=1e-4)


def test_gradient_deterministic(gpu):
    rng = np.random.default_rng(9)
    p = rng.normal(0, 0.3, T.mlp_param_count()).astype(np.float32)
    x = rng.uniform(0, 1, (3000, 3)).astype(np.float32)
    t = rng.uniform(0, 1, 3000).astype(np.float32)
    g1, _ = T.mlp_loss_grad(p, x, t, device=gpu)
    g2, _ = T.mlp_loss_grad(p, x, t, device=gpu)
    assert np.array_equal(g1, g2)


def test_training_reduces_loss(gpu):
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 1, (20000, 3)).astype(np.float32)
    t = (0.2 + 0.5 * x[:, 1] + 0.1 * np.maximum(x[:, 0] - 0.3, 0) ** 2).astype(np.float32)
    p0 = T.mlp_init(1)
    _, l0 = oracle.mlp_loss_grad(p0.astype(np.float64), x, t)
    p1, l1 = T.mlp_train(p0, x, t, epochs=15, batch=1024, lr=3e-3, device=gpu)
    assert l1 < l0 / 20
    assert np.isclose(l1, oracle.mlp_loss_grad(p1.astype(np.float64), x, t)[1], rtol=1e-3)


@pytest.fixture(scope="module")
def rmat16():
    beg, adj = T.rmat_csr(16, 16, seed=1)
    return beg, adj, oracle.count_hash(beg, adj)


def test_features_match_oracle(gpu, rmat16):
    beg, adj, _ = rmat16
    with T.Context(gpu) as c:
        c.load(beg, adj)
        f = c.features()
    want = oracle.mlp_scaled_features(beg, adj)
    assert np.allclose(f, want, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("case", ["K150", "union", "tiny_graphs", "rmat12", "K300"])
def test_features_edge_cases(gpu, case):
    # lists longer than a window span, empty lists, V not a multiple of 64
    from conftest import closed_form_csr
    if case in ("K150", "union"):
        beg, adj, _ = closed_form_csr(case)
    elif case == "tiny_graphs":
        rng = np.random.default_rng(2)
        beg, adj = T.preprocess_edges(rng.integers(0, 5000, 3000).astype(np.int32),
                                      rng.integers(0, 5000, 3000).astype(np.int32))
    elif case == "rmat12":
        beg, adj = T.rmat_csr(12, 8, seed=7)
    else:  # transitive tournament: out-lists of up to 299 entries (> 4 windows)
        i, j = np.triu_indices(300, 1)
        beg, adj = T.preprocess_edges(i.astype(np.int32), j.astype(np.int32))
    with T.Context(gpu) as c:
        c.load(beg, adj)
        f = c.features()
    assert np.allclose(f, oracle.mlp_scaled_features(beg, adj), rtol=1e-5, atol=1e-6), case


@pytest.mark.parametrize("scale,ef,seed", [(14, 16, 1), (16, 16, 1), (12, 32, 4)])
def test_walk_features_match_oracle(gpu, scale, ef, seed):
    # the flat segmented kernel over the walk index's per-entry rows and d+
    beg, adj = T.rmat_csr(scale, ef, seed=seed)
    with T.Context(gpu) as c:
        c.load(beg, adj)
        f = c.walk_features()
    assert np.allclose(f, oracle.mlp_scaled_features_short(beg, adj), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("seed", [1, 2])
def test_cost_predictions_match_oracle(gpu, rmat16, seed):
    # the plan's inference (hidden layer on the matrix cores in f16 with f32
    # accumulation) against the float64 oracle forward on the same features;
    # vertices with d+ < 2 are not counted and cost 0
    beg, adj, _ = rmat16
    rng = np.random.default_rng(seed)
    H = 32
    p = rng.normal(0, 0.3, T.mlp_param_count()).astype(np.float32)
    # last layer small around 0.5, so that outputs stay inside the [0, 1] clamp
    p[4 * H + H * H + H: 4 * H + H * H + 2 * H] = rng.normal(0, 0.02, H)
    p[-1] = 0.5
    with T.Context(gpu) as c:
        c.load(beg, adj)
        c.set_cost_model(p)
        feat = c.features()
        cost = c.predict_costs().astype(np.float64)
    d = np.diff(beg)
    y = np.clip(oracle.mlp_forward(p.astype(np.float64), feat.astype(np.float64)), 0.0, 1.0)
    counted = d >= 2
    assert np.all(cost[~counted] == 0)
    assert np.mean((y[counted] > 0) & (y[counted] < 1)) > 0.9  # the outputs vary
    got = np.log2(1.0 + cost[counted]) / 32.0
    # f16 operands (relative rounding 2^-11) on the 32-term hidden layer:
    # about 1e-3 on y; a wrong operand layout would be off by O(1)
    err = np.abs(got - y[counted])
    print(f"[gpu] cost model inference vs oracle: max {err.max():.2e}, mean {err.mean():.2e}",
          flush=True)
    assert err.max() < 1.5e-2 and err.mean() < 2e-3
    assert np.corrcoef(got, y[counted])[0, 1] > 0.999


def test_learned_without_model_falls_back(gpu, rmat16):
    beg, adj, want = rmat16
    with T.Context(gpu) as c:
        c.load(beg, adj)
        r = c.count(schedule=T.SCHED_LEARNED)
        assert r.triangles == want and r.schedule_used == T.SCHED_DEGREE


def test_learned_schedule_counts_and_predicts(gpu, rmat16):
    beg, adj, want = rmat16
    with T.Context(gpu) as c:
        c.load(beg, adj)
        loss, sec = c.train_cost_model(epochs=20, max_samples=50000, seed=1)
        assert loss < 0.01  # MSE of log2(1+cycles)/32
        for kw in ({}, dict(threads=64), dict(threads=1024), dict(chunk=5),
                   dict(flags=T.FLAG_FORCE_SPILL), dict(rank=1, total_process=2)):
            r = c.count(schedule=T.SCHED_LEARNED, **kw)
            assert r.schedule_used == T.SCHED_LEARNED
            if "rank" in kw:
                r0 = c.count(schedule=T.SCHED_LEARNED, rank=0, total_process=2)
                assert r0.triangles + r.triangles == want
            else:
                assert r.triangles == want
                assert r.probes == expected_probes(beg, adj, kw.get("flags", 0))
        # The model ranks vertices by cost.  Cycle counts of small vertices are
        # noisy, so the bar is the measurement's own repeatability: two
        # instrumented runs against each other.
        c1, _, _ = c.profile_vertices(flags=T.FLAG_WAVE_ONLY)
        c2, _, _ = c.profile_vertices(flags=T.FLAG_WAVE_ONLY)
        pred = c.predict_costs()
        m = (c1 > 0) & (c2 > 0) & (pred > 0)
        lg = lambda a: np.log2(1 + a[m].astype(np.float64))
        rho_noise = np.corrcoef(lg(c1), lg(c2))[0, 1]
        rho = np.corrcoef(lg(pred), lg(c2))[0, 1]
        heavy = c2[m] > np.quantile(c2[m], 0.9)
        rho_heavy = np.corrcoef(lg(pred)[heavy], lg(c2)[heavy])[0, 1]
        print(f"[gpu] cost model: mse {loss:.5f}, log-cycle correlation {rho:.3f} "
              f"(run-to-run {rho_noise:.3f}; top decile {rho_heavy:.3f})", flush=True)
        assert rho > 0.5 and rho >= rho_noise - 0.15
        # set/get round trip drives the same schedule
        params = c.get_cost_model()
        with T.Context(gpu) as c2:
            c2.load(beg, adj)
            c2.set_cost_model(params)
            r2 = c2.count(schedule=T.SCHED_LEARNED)
            assert r2.triangles == want and r2.schedule_used == T.SCHED_LEARNED


def test_judge_and_auto_schedule(gpu, rmat16):
    beg, adj, want = rmat16
    with T.Context(gpu) as c:
        c.load(beg, adj)
        assert c.schedule_judgement()[0] == -1
        r = c.count(schedule=T.SCHED_AUTO)  # no model: the degree split
        assert r.triangles == want and r.schedule_used == T.SCHED_DEGREE
        c.train_cost_model(epochs=5, max_samples=20000, seed=1)
        decision, ls, ds = c.schedule_judgement()
        assert decision in (T.SCHED_LEARNED, T.SCHED_DEGREE) and ls > 0 and ds > 0
        assert decision == (T.SCHED_LEARNED if ls < 0.99 * ds else T.SCHED_DEGREE)
        r = c.count(schedule=T.SCHED_AUTO)
        assert r.triangles == want and r.schedule_used == decision
        # forced learned still runs the model
        assert c.count(schedule=T.SCHED_LEARNED).schedule_used == T.SCHED_LEARNED
        # a model set by hand has no verdict: AUTO runs it
        c.set_cost_model(c.get_cost_model())
        assert c.schedule_judgement()[0] == -1
        assert c.count(schedule=T.SCHED_AUTO).schedule_used == T.SCHED_LEARNED
        # a new graph drops the verdict but keeps the model
        c.load(beg, adj)
        assert c.schedule_judgement()[0] == -1
        assert c.count(schedule=T.SCHED_AUTO).schedule_used == T.SCHED_LEARNED


def test_learned_features_are_a_kept_index(gpu, rmat16):
    # the learned plan's features depend on the graph and the walk alone: the
    # first learned count of a graph makes them (index time), later ones reuse
    # them; TRUST_FLAG_FRESH_INDEX and a new load drop them, and each walk
    # has its own
    beg, adj, want = rmat16
    with T.Context(gpu) as c:
        c.load(beg, adj)
        c.train_cost_model(epochs=2, max_samples=5000, seed=4)
        for walk in (T.FLAG_REFERENCE_WALK, T.FLAG_SHORT_WALK):
            kw = dict(schedule=T.SCHED_LEARNED, flags=walk)
            first = c.count(**dict(kw, flags=walk | T.FLAG_FRESH_INDEX))
            again = c.count(**kw)
            assert first.triangles == again.triangles == want
            assert first.schedule_used == again.schedule_used == T.SCHED_LEARNED
            assert first.index_seconds > 0
            assert again.index_seconds == 0, walk
            assert again.seconds == pytest.approx(again.plan_seconds + again.count_seconds)
        # the other walk's features were made last: the reference walk remakes its own
        assert c.count(schedule=T.SCHED_LEARNED, flags=T.FLAG_REFERENCE_WALK).index_seconds > 0
        c.load(beg, adj)
        r = c.count(schedule=T.SCHED_LEARNED, flags=T.FLAG_REFERENCE_WALK)
        assert r.triangles == want and r.index_seconds > 0


def test_learned_order_counts_exactly_rmat20(gpu):
    # the trained network's order on a mid-size R-MAT counts exactly what the
    # degree split counts (times are printed, not asserted)
    beg, adj = T.rmat_csr(20, 16, seed=1)
    with T.Context(gpu) as c:
        c.load(beg, adj)
        c.train_cost_model(epochs=15, max_samples=200000, seed=2)
        d = min((c.count() for _ in range(3)), key=lambda r: r.seconds)
        l = min((c.count(schedule=T.SCHED_LEARNED) for _ in range(3)), key=lambda r: r.seconds)
        assert d.triangles == l.triangles
        print(f"[gpu] rmat20 degree {d.seconds * 1e3:.2f} ms, learned {l.seconds * 1e3:.2f} ms "
              f"(plan {l.plan_seconds * 1e3:.2f} ms), gain {d.seconds / l.seconds:.3f}x", flush=True)
