"""GPU checks of the learned tuner: the HIP network (csrc/tcsg_mlp.hip)
against its CPU restatement, training on the measured corpus, and the
top-k ranking against the exhaustive sweep and model-free strategies."""
from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu

IN, H = 41, 128


@pytest.fixture(scope="module")
def tuner():
    import torch
    torch.cuda.set_device(0)
    from threadcoarsenedstencilgenerator_amd import tuner as T
    assert len(T.FEATURES) == IN
    return T


def _case(seed, B):
    rng = np.random.default_rng(seed)
    p = mlp.init_params(IN, H, seed)
    W1, b1, W2, b2, W3, b3 = mlp.split(p, IN, H)
    b1[:] = rng.normal(0, 0.3, H).astype(np.float32)
    b2[:] = rng.normal(0, 0.3, H).astype(np.float32)
    b3[:] = np.float32(0.1)
    x = rng.normal(size=(B, IN)).astype(np.float32)
    y = rng.normal(size=B).astype(np.float32)
    return p, x, y


def test_hip_init_is_the_oracle_init(tuner):
    m = tuner.CostModel(in_dim=IN, hidden=H, seed=7)
    print("case: mlp init seed 7", flush=True)
    assert np.array_equal(m.params().astype(np.float64), mlp.init_params(IN, H, 7))


@pytest.mark.parametrize("B", [1, 37, 256])
def test_hip_forward_and_gradients_match_oracle(tuner, B):
    print(f"case: mlp forward/gradients B={B}", flush=True)
    p, x, y = _case(11 + B, B)
    m = tuner.CostModel(in_dim=IN, hidden=H, max_batch=256)
    m.set_params(p.astype(np.float32))
    got = m.forward_raw(x)
    ref, _ = mlp.forward(p, x, IN, H)
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
    g, L = m.gradients_raw(x, y)
    gr, Lr = mlp.gradients(p, x, y, IN, H)
    # FP32 sums of at most max(B, 128) products against FP64
    assert np.abs(g - gr).max() <= 1e-4 * np.abs(gr).max() + 1e-7
    assert L == pytest.approx(Lr, rel=1e-5)


def test_hip_adam_steps_match_oracle(tuner):
    """Two Adam steps: each HIP step against the CPU Adam applied to the HIP
    gradient at the same parameters (so only the update arithmetic differs)."""
    print("case: mlp adam 2 steps", flush=True)
    p, x, y = _case(5, 64)
    m = tuner.CostModel(in_dim=IN, hidden=H, max_batch=64)
    m.set_params(p.astype(np.float32))
    opt = mlp.Adam(p.size, lr=1e-3)
    cur = p.copy()
    for _ in range(2):
        g, _ = m.gradients_raw(x, y)
        loss_before = m.train_raw(x, y, 1e-3)
        assert loss_before == pytest.approx(mlp.loss(cur, x, y, IN, H), rel=1e-5)
        cur = opt.step(cur, g.astype(np.float64))
        got = m.params().astype(np.float64)
        assert np.abs(got - cur).max() <= 1e-6
        cur = got


@pytest.fixture(scope="module")
def corpora(tuner):
    if not (os.path.exists(tuner.TRAIN_DATA) and os.path.exists(tuner.HELDOUT_DATA)):
        pytest.skip("no measured corpus in this tree")
    return tuner.Corpus(tuner.read_rows(tuner.TRAIN_DATA)), tuner.Corpus(tuner.read_rows(tuner.HELDOUT_DATA))


def test_training_fits_the_corpus(tuner, corpora):
    train, _ = corpora
    print(f"case: fit training corpus rows={len(train.rows)}", flush=True)
    m = tuner.CostModel()
    losses = m.fit(train.X, train.y, epochs=30)
    var = float(np.var(train.y))
    assert losses[-1] < 0.1 * var and losses[-1] < losses[0]
    pred = m.predict(train.X)
    assert np.corrcoef(pred, train.y)[0, 1] > 0.95


def test_top8_on_heldout_configurations(tuner, corpora):
    """The network trained by the HIP kernels on the training configurations
    ranks the block shapes of held-out configurations (kernels it has not
    seen); timing only its top 8 must land within 5% of the exhaustive
    sweep's best on average in at most 10% of the sweeps' GPU time, and lose
    less than the model-free strategies with the same 8 measurements: 8
    shapes at random, and the 8 shapes best on average over the training
    configurations."""
    train, held = corpora
    print(f"case: top-8 on {len(held.groups)} held-out configurations", flush=True)
    rep = tuner.evaluate(train, held, k=8, epochs=60)
    mr, rr, fr = rep["model_regret"], rep["random_regret"], rep["fixed_regret"]
    print({k: rep[k] for k in ("model_regret", "random_regret", "fixed_regret", "model_time_frac",
                               "model_time_frac_total", "within_5pct")}, flush=True)
    assert mr["mean"] <= 0.05
    # GPU time of tuning every held-out configuration (top 8 measured plus the
    # network's inference) over the time of their exhaustive sweeps
    assert rep["model_time_frac_total"] <= 0.10
    assert mr["mean"] < 0.5 * rr["mean"]
    assert mr["mean"] < fr["mean"]


def test_tune_with_the_committed_model_on_unseen_kernels(tuner):
    """tune() with the committed model on four kernels outside both corpora
    (the reference's code shape with launch bound 1024: 2D Box Jacobi r=1
    1x1, 3D Star Variable r=4 1x1x1 and 2x4x8, 3D Box Jacobi r=2 2x2x4) on
    2^26-point grids: it times exactly 8 shapes, and its picks lose on
    average at most 10% against the full sweeps of the same kernels
    measured here."""
    if not os.path.exists(tuner.DEFAULT_MODEL):
        pytest.skip("no trained cost model in this tree")
    from threadcoarsenedstencilgenerator_amd import manifest
    from threadcoarsenedstencilgenerator_amd import runtime as rt
    seen = {r["key"] for p in (tuner.TRAIN_DATA, tuner.HELDOUT_DATA) if os.path.exists(p) for r in tuner.read_rows(p)}
    cases = [(manifest.cfg1_desc("plain"), (8194, 8194)), (manifest.cfg2_desc(1, 1, 1), (520, 520, 264)),
             (manifest.cfg2_desc(2, 4, 8), (520, 264, 520)),
             (manifest.cfg4_desc(shape="plain", max_threads=1024), (516, 260, 516))]
    regrets = []
    for d, size in cases:
        assert d.key() not in seen
        print(f"case: tune {d.key()} {size}", flush=True)
        k = rt.Kernel(d)
        if d.coeff == 1:
            k.set_mask(rt.mask_seeded(2, d.n))
        block, glups, tried = tuner.tune(k, size, k=8)
        assert len(tried) == 8 and block in tuner.candidates(k, d.dimension)
        rows = tuner.measure_sweep(d, size)
        best = max(r["glups"] for r in rows)
        regrets.append(best / glups - 1.0)
        print(f"tuned {block} {glups:.1f} GLup/s vs sweep best {best:.1f} over {len(rows)} shapes", flush=True)
        k.close()
    print(f"regrets {[round(r, 4) for r in regrets]}", flush=True)
    assert np.mean(regrets) <= 0.10


def test_tune_command_line():
    """`tuner tune` on a kernel of the default build set: the model's top 8
    shapes are timed and the fastest is reported as one JSON line."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    argv = [sys.executable, "-m", "threadcoarsenedstencilgenerator_amd.tuner", "tune", "3D",
            "Star Row Variable 4 1 4 16 Write 16 --max-threads=256", "--grid", "136", "72", "72"]
    print(" ".join(argv[3:]), flush=True)
    r = subprocess.run(argv, capture_output=True, text=True, timeout=240, cwd=root)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads([ln for ln in r.stdout.splitlines() if ln.startswith("{")][-1])
    assert out["model"] and len(out["tried"]) == 8 and out["glups"] > 0
    assert out["block"] in [b for b, _ in out["tried"]]
