"""Selector host logic: cost model closed forms, features, hand-written rules,
and the numpy network reference pinned by finite differences."""
import numpy as np
import pytest

from bine_trees_fugaku_amd import linksim as ls
from bine_trees_fugaku_amd import selector as sel
from tests import mlp_ref


def stats(num, den, rep, mult, hops):
    a = lambda v: np.asarray(v, np.uint32)  # noqa: E731
    return ls.ClassStats(a(num), a(den), a(rep), a(mult), np.zeros(len(num), np.uint64), a(hops))


def test_cost_single_class_closed_form():
    st = stats([1], [2], [1], [3], [2])
    S = 1000
    T = sel.case_times(st, [S])[0]
    for j, k in enumerate(sel.SEGMENTS):
        seg = -(-S // k)
        stage = sel.ALPHA_US + sel.BETA_US * 2 + 3 * (seg // 2) / sel.BW_BYTES_PER_US
        assert T[j] == pytest.approx(stage + (k - 1) * stage, rel=1e-15)


def test_cost_pipeline_bottleneck_and_repetitions():
    st = stats([1, 1], [1, 4], [1, 5], [1, 2], [1, 3])
    S = 2 ** 20
    T = sel.case_times(st, [S])[0]
    for j, k in enumerate(sel.SEGMENTS):
        seg = -(-S // k)
        s0 = sel.ALPHA_US + sel.BETA_US * 1 + 1 * seg / sel.BW_BYTES_PER_US
        s1 = sel.ALPHA_US + sel.BETA_US * 3 + 2 * (seg // 4) / sel.BW_BYTES_PER_US
        assert T[j] == pytest.approx(s0 + 5 * s1 + (k - 1) * max(s0, s1), rel=1e-12)


def test_shares_are_exact_integers_at_large_sizes():
    """floor(seg * num / den) in uint64: no float rounding up to 2^28 bytes x 2^14 denominators."""
    st = stats([3], [16384], [1], [1], [0])
    S = 2 ** 28 + 12345
    T = sel.case_times(st, [S])[0]
    want = sel.ALPHA_US + (S * 3 // 16384) / sel.BW_BYTES_PER_US
    assert T[0] == want


def test_features_shape_and_onehots():
    f = sel.features([8, 8, 8], "alltoall", [8, 2 ** 28])
    assert f.shape == (2, sel.N_FEATURES)
    assert f[0, sel.COLLS.index("alltoall")] == 1 and f[:, :8].sum() == 2
    assert f[0, 13] == 1 and f[0, 8] == pytest.approx(3 / 28) and f[1, 8] == pytest.approx(1.0)


def test_training_tori_exclude_held_out_shapes():
    tr = sel.training_tori()
    held = {tuple(sorted(t)) for t in sel.HELD_OUT}
    assert all(tuple(sorted(t)) not in held for t in tr)
    assert [256, 32] not in tr and [8, 8, 8] not in tr and [4, 8, 16] in tr
    assert all(8 <= np.prod(t) <= 16384 for t in tr)


def test_option_validity():
    assert sel.option_valid("bcast").sum() == 2 * len(sel.SEGMENTS)
    assert sel.option_valid("alltoall").sum() == 3 * len(sel.SEGMENTS)
    assert sel.option_valid("allreduce").sum() == 3 * len(sel.SEGMENTS)


def test_rules_return_defined_options():
    for coll in sel.COLLS:
        for P in (8, 64, 512, 4096, 8192):
            for S in (8, 1000, 65536, 2 ** 20, 2 ** 28):
                for rule in (sel.fjmpi_choice, sel.reference_default_choice):
                    a, k = rule(coll, P, S)
                    assert a in ls.DEFINED[coll] and k in sel.SEGMENTS


def test_fjmpi_allreduce_thresholds():
    """fjmpi-dectree.conf:110-120 for 33..3840 processes."""
    assert sel.fjmpi_choice("allreduce", 512, 98304) == ("recdoub", 1)
    assert sel.fjmpi_choice("allreduce", 512, 98305)[0] == "ring"
    assert sel.fjmpi_choice("allreduce", 512, 393216) == ("ring", sel._seg(393216, 4096))
    assert sel.fjmpi_choice("allreduce", 8, 8 * 8191) == ("recdoub", 1)


DIMS = [7, 33, 20, 5]


def test_numpy_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    n = sum(fo * fi + fo for fi, fo in zip(DIMS[:-1], DIMS[1:]))
    p = rng.standard_normal(n) * 0.3
    x = rng.standard_normal((11, DIMS[0]))
    t = rng.standard_normal((11, DIMS[-1]))
    mask = (rng.random((11, DIMS[-1])) > 0.3).astype(np.float64)
    loss, g = mlp_ref.loss_and_grads(p, DIMS, x, t, mask)
    eps = 1e-6
    for i in rng.choice(n, 60, replace=False):
        e = np.zeros(n)
        e[i] = eps
        lp, _ = mlp_ref.loss_and_grads(p + e, DIMS, x, t, mask)
        lm, _ = mlp_ref.loss_and_grads(p - e, DIMS, x, t, mask)
        assert g[i] == pytest.approx((lp - lm) / (2 * eps), rel=1e-5, abs=1e-8)


def test_numpy_reference_forward_is_the_textbook_network():
    rng = np.random.default_rng(1)
    n = sum(fo * fi + fo for fi, fo in zip(DIMS[:-1], DIMS[1:]))
    p = rng.standard_normal(n)
    x = rng.standard_normal((4, DIMS[0]))
    layers = mlp_ref.unpack(p, DIMS)
    h = x
    for i, (W, b) in enumerate(layers):
        h = np.array([[sum(W[o, j] * row[j] for j in range(W.shape[1])) + b[o] for o in range(W.shape[0])]
                      for row in h])
        if i < len(layers) - 1:
            h = np.maximum(h, 0)
    assert np.allclose(mlp_ref.forward(p, DIMS, x)[-1], h, rtol=1e-12)
