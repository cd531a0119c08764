"""The learned tuner's live check on the GPU: the committed model against exhaustive sweeps
of configurations that neither measured corpus contains, swept now on the GPU that runs the
test.  (It sorts just before the allocation-free test: its sweeps take about a minute, and its
bound, like every timing comparison, depends on this GPU's measurements.)"""
from __future__ import annotations

import os

import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tuner():
    import torch
    torch.cuda.set_device(0)
    from threadcoarsenedstencilgenerator_amd import tuner as T
    return T


def test_live_top8_on_fresh_configurations(tuner):
    """The committed model on 40 configurations whose kernels neither corpus contains,
    swept exhaustively on this GPU now (nothing about them was measured before): timing
    only the 8 shapes tune() picks of each configuration's launchable shapes lands within 5%
    of the sweep's best on average, in at most 10% of the sweeps' GPU time, and loses less
    than half what 8 random shapes would."""
    if not os.path.exists(tuner.DEFAULT_MODEL):
        pytest.skip("no trained cost model in this tree")
    print(f"case: live top-8 on {tuner.LIVE_CONFIGS} fresh configurations", flush=True)
    rep = tuner.live_evaluate(k=8, log=lambda m: print(m, flush=True))
    mr, rr = rep["model_regret"], rep["random_regret"]
    print({k: rep[k] for k in ("model_regret", "random_regret", "fixed_regret", "model_time_frac_total",
                               "within_5pct", "wall_seconds_sweeps")}, flush=True)
    assert rep["heldout_groups"] == tuner.LIVE_CONFIGS
    assert mr["mean"] <= 0.05
    assert rep["model_time_frac_total"] <= 0.10
    assert mr["mean"] < 0.5 * rr["mean"]
