"""The shipped learned selector's HIP inference kernels (hmlMlpPredict) on
its held-out evaluation set (tests/golden/selector_heldout.npz): their
picks reach at least 0.95 of the autotune-best throughput on average, more
than the built-in heuristic, and agree with the oracle's float64 forward
pass of the same parameters."""
import os

import numpy as np
import pytest

from test_selector_heldout import (FIXTURE, MODEL, fixture_inputs, load_model, oracle_picks,
                                   score)

pytestmark = pytest.mark.gpu


def test_hip_inference_on_heldout(hml):
    fx = dict(np.load(FIXTURE))
    dims, params = load_model(hml)
    X, masks = fixture_inputs(hml, fx)
    net = hml.Mlp.load(MODEL)
    P = net.predict(X.astype(np.float32), masks.astype(np.float32))
    picks = np.argmax(np.where(masks > 0, P, -1.0), axis=1)
    f = score(fx, picks)
    f_heur = fx["best_ms"] / fx["heuristic_ms"]
    print("HIP inference: mean %.4f, within 5%% %.3f; heuristic mean %.4f"
          % (f.mean(), (f >= 0.95).mean(), f_heur.mean()))
    assert f.mean() >= 0.95
    assert f.mean() >= f_heur.mean() + 0.05
    want = oracle_picks(dims, params, X, masks)
    assert (picks == want).mean() >= 0.995
