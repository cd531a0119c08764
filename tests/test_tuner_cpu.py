"""CPU checks of the learned tuner: the network's CPU restatement against
finite differences and closed forms, the sweep's candidate enumeration
against the reference's primes sweep, features, and the regret arithmetic."""
from __future__ import annotations

import itertools
import math
import os

import numpy as np
import pytest

import oracle
from oracle import mlp
from threadcoarsenedstencilgenerator_amd import tuner
from threadcoarsenedstencilgenerator_amd.spec import ALIGN_WRITE, STAR, VARIABLE, StencilDesc


def _net(in_dim, hidden, B, seed):
    rng = np.random.default_rng(seed)
    p = mlp.init_params(in_dim, hidden, seed)
    # non-zero biases so that some ReLUs are inactive and some active
    W1, b1, W2, b2, W3, b3 = mlp.split(p, in_dim, hidden)
    b1[:] = rng.normal(0, 0.3, hidden)
    b2[:] = rng.normal(0, 0.3, hidden)
    b3[:] = 0.1
    x = rng.normal(size=(B, in_dim))
    y = rng.normal(size=B)
    return p, x, y


@pytest.mark.parametrize("in_dim,hidden,B", [(3, 4, 5), (7, 16, 9)])
def test_gradients_match_finite_differences_everywhere(in_dim, hidden, B):
    p, x, y = _net(in_dim, hidden, B, 3)
    g, _ = mlp.gradients(p, x, y, in_dim, hidden)
    fd = mlp.finite_difference(p, x, y, in_dim, hidden, range(p.size))
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-7), np.abs(g - fd).max()


def test_gradients_match_finite_differences_at_tuner_size():
    in_dim, hidden, B = len(tuner.FEATURES), 128, 32
    p, x, y = _net(in_dim, hidden, B, 5)
    g, _ = mlp.gradients(p, x, y, in_dim, hidden)
    rng = np.random.default_rng(0)
    o = np.cumsum([0, hidden * in_dim, hidden, hidden * hidden, hidden, hidden, 1])
    # indices from every parameter block, incl. both bias vectors and b3
    idx = np.concatenate([rng.integers(o[i], o[i + 1], size=min(40, o[i + 1] - o[i])) for i in range(6)])
    fd = mlp.finite_difference(p, x, y, in_dim, hidden, idx)
    assert np.allclose(g[idx], fd, rtol=1e-4, atol=1e-8)


def test_all_active_network_is_linear():
    """With every hidden unit active the network is the affine map
    W3 (W2 (W1 x + b1) + b2) + b3; its gradient in b3 is the mean residual x2."""
    in_dim, hidden, B = 4, 6, 7
    rng = np.random.default_rng(2)
    p = np.zeros(mlp.num_params(in_dim, hidden))
    W1, b1, W2, b2, W3, b3 = mlp.split(p, in_dim, hidden)
    W1[:] = rng.uniform(0, 1, W1.shape)
    W2[:] = rng.uniform(0, 1, W2.shape)
    W3[:] = rng.normal(size=W3.shape)
    b1[:] = 5.0
    b2[:] = 5.0
    b3[:] = 0.25
    x = rng.uniform(0, 1, (B, in_dim))
    y = rng.normal(size=B)
    pred, _ = mlp.forward(p, x, in_dim, hidden)
    lin = ((x @ W1.T + b1) @ W2.T + b2) @ W3.T + b3
    assert np.allclose(pred, lin[:, 0], rtol=1e-13)
    g, L = mlp.gradients(p, x, y, in_dim, hidden)
    gW1, gb1, gW2, gb2, gW3, gb3 = mlp.split(g, in_dim, hidden)
    assert gb3[0] == pytest.approx(2 * np.mean(pred - y), rel=1e-12)
    assert L == pytest.approx(np.mean((pred - y) ** 2), rel=1e-12)
    # linear chain rule for W1: dL/dW1 = (W2^T W3^T dout^T) x
    dout = 2 * (pred - y) / B
    expect = (W2.T @ W3.T @ dout[None, :]) @ x
    assert np.allclose(gW1, expect, rtol=1e-12)


def test_tiny_network_by_hand():
    """in=1, hidden=1: y = w3 relu(w2 relu(w1 x + b1) + b2) + b3."""
    p = np.array([2.0, -1.0, 3.0, 0.5, -2.0, 1.0])   # w1 b1 w2 b2 w3 b3
    x = np.array([[1.0], [0.25]])
    y = np.array([0.0, 1.0])
    # sample 1: h1 = relu(1) = 1, h2 = relu(3.5) = 3.5, pred = -6
    # sample 2: h1 = relu(-0.5) = 0, h2 = relu(0.5) = 0.5, pred = 0
    pred, _ = mlp.forward(p, x, 1, 1)
    assert pred.tolist() == [-6.0, 0.0]
    g, L = mlp.gradients(p, x, y, 1, 1)
    assert L == (36 + 1) / 2
    d = np.array([-6.0, -1.0])             # dL/dpred = 2 (pred - y) / 2
    assert g[5] == d.sum()                   # b3
    assert g[4] == -6.0 * 3.5 + -1.0 * 0.5   # w3: sum d h2
    # layer 2: dh2 = d w3 = [12, 2], both h2 > 0
    assert g[3] == 14.0                      # b2
    assert g[2] == 12.0 * 1.0 + 2.0 * 0.0    # w2: sum dh2 h1
    # layer 1: dh1 = dh2 w2 = [36, 6], masked by h1 > 0 -> [36, 0]
    assert g[1] == 36.0
    assert g[0] == 36.0 * 1.0


def test_adam_first_step_closed_form():
    """After one step m/c1 = g and v/c2 = g^2, so p1 = p0 - lr g / (|g| + eps)."""
    g = np.array([1e-3, -2.0, 0.0, 5e-9, -7.5])
    p0 = np.array([0.5, 0.5, 0.5, 0.5, 0.5])
    opt = mlp.Adam(g.size, lr=0.01)
    p1 = opt.step(p0, g)
    assert np.allclose(p1, p0 - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)


def test_init_params_uniform_bounds():
    p = mlp.init_params(33, 128, 7)
    W1, b1, W2, b2, W3, b3 = mlp.split(p, 33, 128)
    assert np.all(np.abs(W1) <= 1 / math.sqrt(33)) and np.all(np.abs(W2) <= 1 / math.sqrt(128))
    assert not b1.any() and not b2.any() and not b3.any()
    assert abs(W2.mean()) < 0.01 and W2.std() == pytest.approx(1 / math.sqrt(128) / math.sqrt(3), rel=0.02)
    assert np.array_equal(p, p.astype(np.float32).astype(np.float64))


# ---- sweep candidates ----
@pytest.mark.parametrize("dim,mt", [(2, 1024), (3, 1024), (3, 256), (2, 512)])
def test_sweep_shapes_match_the_reference_sweep(dim, mt):
    """The oracle restates hip/primes.cpp; the tuner's enumeration is the same list in the same order."""
    assert tuner.sweep_shapes(dim, mt) == oracle.block_shapes(dim, mt)


def test_sweep_shape_counts():
    assert len(tuner.sweep_shapes(3, 1024)) == 1529
    assert len(tuner.sweep_shapes(2, 1024)) == 230
    assert len(tuner.sweep_shapes(3, 256)) == 193


def test_divisors_complete():
    for n in range(1, 2000):
        d = tuner.divisors_in_sweep_order(n)
        assert sorted(d) == [k for k in range(1, n + 1) if n % k == 0]


# ---- features ----
def test_features_of_the_headline_case():
    d = StencilDesc(dimension=3, neighbourhood=STAR, coeff=VARIABLE, range=4, xc=2, yc=4, zc=128,
                    alignment=ALIGN_WRITE, cl_size=16)
    f = dict(zip(tuner.FEATURES, tuner.features(d, (520, 520, 520), (128, 2, 1))))
    assert len(f) == len(tuner.FEATURES)
    assert f["log2_nx"] == 9 and f["log2_tile_x"] == 8 and f["log2_tile_z"] == 7
    assert f["log2_workgroups"] == math.log2(2 * 64 * 4) and f["log2_cover"] == 0.0
    assert f["log2_points"] == math.log2(25) and f["bytes_per_point"] == 8.0
    f2 = tuner.features(d, (520, 520, 520), (64, 2, 1))
    assert not np.array_equal(tuner.features(d, (520, 520, 520), (128, 2, 1)), f2)


# ---- regret arithmetic ----
def test_random_k_regret_matches_enumeration():
    rng = np.random.default_rng(4)
    v = rng.uniform(1, 10, 11)
    for k in (1, 3, 8, 11):
        brute = np.mean([v.max() / v[list(c)].max() - 1 for c in itertools.combinations(range(11), k)])
        assert tuner.random_k_regret(v, k) == pytest.approx(brute, rel=1e-12)


def test_top_k_outcome():
    g = np.array([10.0, 40.0, 20.0, 30.0])
    ms = np.array([4.0, 1.0, 2.0, 3.0])
    reg, frac = tuner.top_k_outcome(g, ms, np.array([2, 3, 1, 0]), 2)
    assert reg == pytest.approx(40 / 30 - 1) and frac == pytest.approx(5 / 10)
    reg, _ = tuner.top_k_outcome(g, ms, np.array([1, 0, 2, 3]), 1)
    assert reg == 0.0


def test_mixed_order():
    """The k shapes tune() measures: the model's first k - mix, then the model-free ranking's
    first not already chosen; the rest follows in the model's order, every shape once."""
    pred, fixed = np.array([3, 1, 2, 0, 4]), np.array([1, 4, 0, 2, 3])
    assert list(tuner.mixed_order(pred, fixed, 4, 0)) == [3, 1, 2, 0, 4]
    assert list(tuner.mixed_order(pred, fixed, 4, 1)) == [3, 1, 2, 4, 0]
    assert list(tuner.mixed_order(pred, fixed, 4, 2)) == [3, 1, 4, 0, 2]
    assert list(tuner.mixed_order(pred, fixed, 4, 4)) == [1, 4, 0, 2, 3]
    assert list(tuner.mixed_order(pred, fixed, 8, 1)) == [3, 1, 2, 0, 4]   # k past the candidates


def test_live_configurations_are_fresh():
    """The live check's configurations: distinct kernels in neither measured corpus's
    configuration set, drawn from the same workload space."""
    live = tuner.live_configs()
    assert len(live) == tuner.LIVE_CONFIGS
    keys = [d.key() for d, _ in live]
    assert len(set(keys)) == len(keys)
    assert not set(keys) & {d.key() for d, _ in tuner.train_configs() + tuner.heldout_configs()}


def test_corpus_files_match_their_configuration_sets():
    """The committed corpora: the held-out file holds one whole sweep per
    held-out configuration (every launchable shape once), the training file
    a sample of shapes per training configuration; no kernel is in both."""
    if not (os.path.exists(tuner.TRAIN_DATA) and os.path.exists(tuner.HELDOUT_DATA)):
        pytest.skip("no measured corpus in this tree")
    train = tuner.Corpus(tuner.read_rows(tuner.TRAIN_DATA))
    held = tuner.Corpus(tuner.read_rows(tuner.HELDOUT_DATA))
    assert not set(train.descs) & set(held.descs)
    for cp, cfgs in ((train, tuner.train_configs()), (held, tuner.heldout_configs())):
        assert np.all(cp.glups > 0) and np.all(np.isfinite(cp.X))
        want = {d.key(): tuple(s) + (1,) * (3 - len(s)) for d, s in cfgs}
        for gk, idx in cp.groups.items():
            assert want[gk[0]] == gk[1:]
            d = cp.descs[gk[0]]
            shapes = [cp.shape[i] for i in idx]
            assert len(set(shapes)) == len(shapes)
            assert set(shapes) <= set(tuner.sweep_shapes(d.dimension, d.max_threads))
            if d.xext == "dppw" and d.variant == 1:
                assert all(s[0] % 64 == 0 for s in shapes)
            if cp is train:
                assert len(shapes) <= tuner.TRAIN_SHAPES_PER_CONFIG
    assert len(held.groups) >= 0.9 * tuner.HELDOUT_CONFIGS


def test_configuration_sampler_stays_in_the_workload_space():
    """Sampled configurations: 2^26-2^27 interior points, ranges 1-12,
    coarsening 1-16 per axis, a bounded statement count, distinct kernels,
    and held-out kernels disjoint from the training kernels."""
    tr, ho = tuner.train_configs(), tuner.heldout_configs()
    assert len(tr) == tuner.TRAIN_CONFIGS and len(ho) == tuner.HELDOUT_CONFIGS
    assert not {d.key() for d, _ in tr} & {d.key() for d, _ in ho}
    for d, s in tr + ho:
        pts = math.prod(s[i] - 2 * d.range for i in range(d.dimension))
        assert 2 ** 26 * 0.98 <= pts <= 2 ** 27 * 1.02, (d, s)
        assert 1 <= d.range <= 12 and all(c in (1, 2, 4, 8, 16) for c in d.coarsening)
        assert d.xc * d.yc * d.zc * d.n <= tuner.MAX_TERMS
        for a in range(d.dimension):
            assert (s[a] - 2 * d.range) % d.coarsening[a] == 0
    dims = {d.dimension for d, _ in ho}
    assert dims == {2, 3} and {d.variant for d, _ in ho} == {0, 1}
