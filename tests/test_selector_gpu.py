"""Selector network on the GPU: HIP forward/backward against the numpy
reference, Adam training, and regret on held-out tori against exhaustive
simulation and the hand-written selectors."""
import numpy as np
import pytest

from bine_trees_fugaku_amd import selector as sel
from tests import mlp_ref

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DIMS = [sel.N_FEATURES, 128, 128, 128, len(sel.OPTIONS)]


def batch(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, DIMS[0])).astype(np.float32)
    t = rng.standard_normal((n, DIMS[-1])).astype(np.float32)
    m = (rng.random((n, DIMS[-1])) > 0.25).astype(np.float32)
    return x, t, m


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("n", [1, 63, 300, 4096])
def test_forward_and_gradients_match_numpy(gpu, n):
    """Tolerances, fp32 products accumulated on the matrix cores vs float64:
    outputs within 2e-4 of their largest magnitude; every gradient entry within
    1e-4 of its condition scale (the same sum with every term taken in absolute
    value, so cancellation in a large sum is allowed for), while a wrong term
    or index shifts entries by about the gradient's own size."""
    net = sel.MLP(DIMS, max_batch=4096, seed=n)
    p = net.params().astype(np.float64)
    x, t, m = batch(n, n)
    y = net.forward(cuda(x)).cpu().numpy()
    y_ref = mlp_ref.forward(p, DIMS, x.astype(np.float64))[-1]
    assert np.abs(y - y_ref).max() <= 2e-4 * np.abs(y_ref).max()
    m[mlp_ref.near_kinks(p, DIMS, x)] = 0  # rows where fp32 may take the other side of a ReLU
    loss, g = net.gradients(cuda(x), cuda(t), cuda(m))
    l_ref, g_ref, g_scale = mlp_ref.loss_and_grads(p, DIMS, x, t, m, scale=True)
    assert float(loss.item()) == pytest.approx(l_ref, rel=1e-4)
    g = g.cpu().numpy().astype(np.float64)
    err = np.abs(g - g_ref)
    assert np.all(err <= 1e-4 * g_scale + 1e-12), float((err / (g_scale + 1e-30)).max())
    # the check has teeth: the gradients are not small against their scale
    live = g_scale > 0  # entries behind a dead ReLU are exactly 0 on both sides
    assert live.mean() > 0.1 and np.median(np.abs(g_ref[live]) / g_scale[live]) > 1e-3
    net.close()


def test_gradients_are_deterministic(gpu):
    net = sel.MLP(DIMS, max_batch=1024, seed=3)
    x, t, m = (cuda(a) for a in batch(777, 3))
    _, g1 = net.gradients(x, t, m)
    g1 = g1.clone()
    _, g2 = net.gradients(x, t, m)
    assert torch.equal(g1, g2)
    net.close()


def test_adam_training_fits_a_smooth_function(gpu):
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (4096, DIMS[0])).astype(np.float32)
    t = np.stack([np.sin(3 * x[:, i % DIMS[0]]) * x[:, (i + 1) % DIMS[0]] for i in range(DIMS[-1])], 1)
    net = sel.MLP(DIMS, max_batch=4096, seed=0)
    X, T, M = cuda(x), cuda(t.astype(np.float32)), cuda(np.ones_like(t, np.float32))
    loss = torch.zeros(1, device="cuda")
    net.train_step(X, T, M, 1e-3, loss)
    first = float(loss.item())
    for _ in range(1200):
        net.train_step(X, T, M, 1e-3, loss)
    assert float(loss.item()) < 0.1 * first
    net.close()


def test_errors_are_reported(gpu):
    net = sel.MLP(DIMS, max_batch=16, seed=0)
    x, t, m = (cuda(a) for a in batch(32, 0))
    with pytest.raises(sel._lib.BineError):
        net.gradients(x, t, m)  # batch larger than max_batch
    net.close()


@pytest.fixture(scope="module")
def trained(gpu):
    train = sel.simulate(sel.training_tori(), sel.default_sizes(2))
    test = sel.simulate(sel.HELD_OUT, sel.default_sizes(4))
    s = sel.Selector(seed=0)
    losses = s.fit(train)
    print(f"selector: {len(train.X)} rows, {train.samples} samples, final loss {losses[-1]:.5f}", flush=True)
    yield s, train, test, losses
    s.close()


def test_selector_regret_on_held_out_tori(trained):
    s, train, test, losses = trained
    assert losses[-1] < 0.01 * losses[0]
    choice, _ = s.choose_rows(test)
    r = sel.regret(test, choice)
    print(f"held-out mean regret {r.mean():.4f}, p95 {np.percentile(r, 95):.4f}", flush=True)
    assert r.mean() <= 0.05
    for dims in sel.HELD_OUT:  # every held-out torus on its own
        rows = np.array([k[0] == tuple(dims) for k in test.keys])
        assert r[rows].mean() <= 0.05, dims


def test_selector_beats_hand_written_selectors(trained):
    s, train, test, _ = trained
    r = sel.regret(test, s.choose_rows(test)[0]).mean()
    base = {name: sel.regret(test, sel.rule_choice_rows(test, rule)).mean()
            for name, rule in (("reference defaults", sel.reference_default_choice), ("fjmpi", sel.fjmpi_choice))}
    # strongest table without the network: the single best option per collective on the training tori
    best_static = {}
    for coll in sel.COLLS:
        rows = np.array([k[1] == coll for k in train.keys])
        T = train.T[rows]
        rel = T / T.min(axis=1, keepdims=True)
        best_static[coll] = int(np.argmin(np.where(np.isfinite(rel), rel, np.inf).mean(axis=0)))
    static = np.array([best_static[k[1]] for k in test.keys])
    base["best static per collective"] = sel.regret(test, static).mean()
    print("baseline mean regrets:", {k: round(float(v), 4) for k, v in base.items()}, flush=True)
    for v in base.values():
        assert r < v


def test_fallback_to_exhaustive_simulation(trained):
    s, _, test, _ = trained
    plain = sel.regret(test, s.choose_rows(test)[0]).mean()
    choice, unsure = s.choose_rows(test, margin=0.05)
    assert sel.regret(test, choice).mean() <= plain
    assert 0 < unsure.mean() < 1
    algo, k, used = s.choose([8, 8, 8], "bcast", 1 << 20, margin=10.0)  # always unsure -> simulation
    T = sel.exhaustive([8, 8, 8], "bcast", 1 << 20)
    assert used and sel.OPTIONS.index((algo, k)) == int(np.argmin(T))


# ---- learned launch shapes (trained on measured kernel timings) ----------------

@pytest.fixture(scope="module")
def shape_timings(gpu):
    from bine_trees_fugaku_amd import shapes as S
    w = S.default_workloads()
    data = S.measure(w, iters=5, log=lambda i, k: print(f"shapes: {i}/{len(w)} {k}", flush=True) if i % 100 == 0 else None)
    return S.split(data, frac=0.3, seed=0)


def test_shape_model_regret_against_autotuner(shape_timings):
    """Held-out workloads: the shape the network picks vs the fastest measured
    shape (what the exhaustive autotuner finds), against the fixed default
    shape and the best single shape fitted on the training workloads."""
    from bine_trees_fugaku_amd import shapes as S
    train, test = shape_timings
    m = S.ShapeModel(len(train.shapes), seed=0)
    m.fit(train)
    choice = m.choose(test.X)
    m.close()
    r = S.regret(test, choice)
    rows = np.arange(len(choice))
    tw = test.T[rows, choice].sum() / test.T.min(axis=1).sum() - 1
    default = [i for i, s in enumerate(test.shapes) if tuple(s) == (1, 256, 0, 0)][0]
    bs = S.best_static(train)
    r_def = S.regret(test, np.full(len(rows), default))
    r_bs = S.regret(test, np.full(len(rows), bs))
    tw_bs = test.T[:, bs].sum() / test.T.min(axis=1).sum() - 1
    print(f"shape model: {len(train.X)} training / {len(test.X)} held-out workloads; mean regret {r.mean():.4f} "
          f"(time-weighted {tw:.4f}); default shape {r_def.mean():.4f}; best static {r_bs.mean():.4f} "
          f"(time-weighted {tw_bs:.4f})", flush=True)
    assert r.mean() <= 0.08 and tw <= 0.08
    assert r.mean() < 0.5 * r_bs.mean() and r.mean() < 0.5 * r_def.mean() and tw < 0.5 * tw_bs
