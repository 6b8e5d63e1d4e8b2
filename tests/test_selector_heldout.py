"""The shipped learned selector (config/hml_selector_mi355x.bin) scored on
its held-out evaluation set (tests/golden/selector_heldout.npz: 2,000
random shapes autotuned exhaustively on an MI355X, none of them in the
training set).  Each shape's pick is scored as the fraction of the
autotune-best throughput it reaches, t_best / t_pick, and compared with
the built-in heuristic's fraction on the same shapes.

Host only: the picks come from the oracle's float64 forward pass
(oracle/mlp_oracle.py) of the shipped parameters, masked by the library's
runnable classes, and the library's own host-side selection (loaded
through HML_SELECTOR in a child process) must make the same picks."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import mlp_oracle as mo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "config", "hml_selector_mi355x.bin")
FIXTURE = os.path.join(ROOT, "tests", "golden", "selector_heldout.npz")


def load_model(hml):
    hdr = np.fromfile(MODEL, np.int32, 5)
    assert hdr[0] == 0x504C4D48
    n_in, hidden, n_hidden, n_out = (int(x) for x in hdr[1:])
    assert n_in == hml.lib.hmlSelectorFeatureCount()
    assert n_out == hml.lib.hmlSelectorClassCount(), "shipped model does not match the class space"
    dims = [n_in] + [hidden] * n_hidden + [n_out]
    params = np.fromfile(MODEL, np.float32, offset=20)
    assert params.size == sum(dims[i - 1] * dims[i] + dims[i] for i in range(1, len(dims)))
    return dims, params.astype(np.float64)


def fixture_inputs(hml, fx):
    X, masks = [], []
    for M, N, K, ta, tb, b0 in fx["shapes"]:
        kw = dict(transA=bool(ta), transB=bool(tb), beta0=bool(b0))
        X.append(hml.selector_features(int(M), int(N), int(K), **kw))
        masks.append(hml.selector_mask(int(M), int(N), int(K), **kw))
    return np.array(X, np.float64), np.array(masks)


def score(fx, picks):
    t = fx["class_ms"][np.arange(len(picks)), picks].astype(np.float64)
    return np.where(np.isfinite(t), fx["best_ms"] / t, 0.0)


def oracle_picks(dims, params, X, masks):
    Z = mo.forward(params, dims, X)[-1]
    return np.argmax(np.where(masks > 0, Z, -np.inf), axis=1)


@pytest.fixture(scope="module")
def heldout():
    import hiperml_amd as hml
    fx = dict(np.load(FIXTURE))
    dims, params = load_model(hml)
    X, masks = fixture_inputs(hml, fx)
    return hml, fx, dims, params, X, masks


def test_fixture_is_held_out_scale(heldout):
    hml, fx, *_ = heldout
    n = len(fx["shapes"])
    assert n >= 1900
    assert fx["class_ms"].shape == (n, hml.lib.hmlSelectorClassCount())
    # every shape's best time is a measured time and no class beats it
    finite = np.where(np.isfinite(fx["class_ms"]), fx["class_ms"], np.inf)
    assert (finite.min(axis=1) >= fx["best_ms"] * (1 - 1e-6)).all()
    assert (fx["heuristic_ms"] >= fx["best_ms"] * (1 - 1e-6)).all()


def test_selector_gain_over_heuristic(heldout):
    """The trained network against the method without it: mean fraction of
    autotune-best throughput at least 0.95, and above the heuristic's."""
    hml, fx, dims, params, X, masks = heldout
    f_mlp = score(fx, oracle_picks(dims, params, X, masks))
    f_heur = fx["best_ms"] / fx["heuristic_ms"]
    print("held-out shapes %d: learned mean %.4f (within 5%%: %.3f, min %.3f), heuristic "
          "mean %.4f (within 5%%: %.3f)" % (len(f_mlp), f_mlp.mean(), (f_mlp >= 0.95).mean(),
                                           f_mlp.min(), f_heur.mean(), (f_heur >= 0.95).mean()))
    assert f_mlp.mean() >= 0.95
    assert f_mlp.mean() >= f_heur.mean() + 0.05


def test_library_selection_matches_oracle(heldout):
    """The library's host-side selection of the shipped model picks the same
    classes as the oracle's forward pass (first 300 held-out shapes)."""
    hml, fx, dims, params, X, masks = heldout
    n = 300
    want = oracle_picks(dims, params, X[:n], masks[:n])
    shapes = [tuple(int(v) for v in s) for s in fx["shapes"][:n]]
    code = ("import sys, json; sys.path.insert(0, %r); import hiperml_amd as h; "
            "out = [h.sgemm_select(m, n, k, transA=bool(a), transB=bool(b), beta=0.0 if z else 1.0)"
            " for m, n, k, a, b, z in %r]; print(json.dumps([[list(v.astuple()), s] for v, s in out]))"
            % (ROOT, shapes))
    env = dict(os.environ, HML_SELECTOR=MODEL)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    got = json.loads(r.stdout.strip().splitlines()[-1])
    agree = 0
    for i, ((v, src), s) in enumerate(zip(got, shapes)):
        M, N, K, ta, tb, b0 = s
        assert src == 2
        # compared as variants: classes the split rule makes identical map
        # to one variant, whose class_of is the lowest of them
        w = hml.selector_variant(M, N, K, int(want[i]), transA=bool(ta), transB=bool(tb),
                                 beta0=bool(b0))
        agree += tuple(v) == w.astuple()
    # float64 on both sides; only an exact-tie reordering could differ
    assert agree >= n - 1, (agree, n)
