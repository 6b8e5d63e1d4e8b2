"""GPU tests of the learned selector's training/inference kernels against
the numpy oracle (oracle/mlp_oracle.py, itself pinned by finite
differences), and of its use as the SGEMM selector."""
import os

import numpy as np
import pytest

import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def batch(rng, B, n_in, n_out):
    X = rng.normal(0, 1, (B, n_in)).astype(np.float32)
    mask = (rng.random((B, n_out)) < 0.6).astype(np.float32)
    mask[:, 0] = 1
    T = (rng.random((B, n_out)) * mask).astype(np.float32)
    T /= T.sum(axis=1, keepdims=True)
    return X, T, mask


def test_gradients_match_oracle(hml):
    rng = np.random.default_rng(0)
    net = hml.Mlp(6, 64, 3, 86, seed=5)
    dims = mo.layer_dims(6, 64, 3, 86)
    X, T, mask = batch(rng, 200, 6, 86)
    p = net.get_params().astype(np.float64)
    loss, g = net.loss_grad(X, T, mask)
    L, g64 = mo.loss_grad(p, dims, X.astype(np.float64), T.astype(np.float64), mask)
    assert abs(loss - L) <= 1e-5 * max(1.0, abs(L))
    scale = np.abs(g64).max()
    np.testing.assert_allclose(g, g64, rtol=1e-3, atol=1e-5 * scale)
    # forward probabilities
    P = net.predict(X, mask)
    P64 = mo.masked_softmax(mo.forward(p, dims, X.astype(np.float64))[-1], mask)
    np.testing.assert_allclose(P, P64, rtol=1e-4, atol=1e-6)
    assert np.all(P[mask == 0] == 0)


def test_adam_step_matches_oracle(hml):
    """One full-batch epoch = one Adam step: w1 = w0 - lr*m/(sqrt(v)+eps)
    with bias-corrected moments of the oracle gradient."""
    rng = np.random.default_rng(1)
    net = hml.Mlp(6, 16, 2, 10, seed=3)
    dims = mo.layer_dims(6, 16, 2, 10)
    X, T, mask = batch(rng, 64, 6, 10)
    w0 = net.get_params().astype(np.float64)
    _, g = mo.loss_grad(w0, dims, X.astype(np.float64), T.astype(np.float64), mask)
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    m = (1 - b1) * g
    v = (1 - b2) * g * g
    lr_t = lr * np.sqrt(1 - b2) / (1 - b1)
    w1 = w0 - lr_t * m / (np.sqrt(v) + eps)
    net.train(X, T, mask, epochs=1, batch=64, lr=lr, seed=0)
    got = net.get_params()
    # where |g| is tiny the Adam step direction is ill-conditioned; compare
    # the step on coordinates with a well-defined sign
    sel = np.abs(g) > 1e-6 * np.abs(g).max()
    np.testing.assert_allclose(got[sel], w1[sel], rtol=1e-4, atol=1e-6)


def test_training_learns_a_rule(hml):
    """A learnable target: the best class is a fixed function of the
    features; training must drive the argmax accuracy up."""
    rng = np.random.default_rng(2)
    n, C = 3000, 12
    X = rng.uniform(0, 1, (n, 6)).astype(np.float32)
    label = (np.floor(X[:, 0] * 3) * 4 + np.floor(X[:, 2] * 4)).astype(int)
    T = np.zeros((n, C), np.float32)
    T[np.arange(n), label] = 1
    mask = np.ones((n, C), np.float32)
    net = hml.Mlp(6, 64, 3, C, seed=1)
    losses = net.train(X, T, mask, epochs=60, batch=128, lr=3e-3, seed=4)
    assert losses[-1] < 0.5 * losses[0]
    acc = (net.predict(X, mask).argmax(axis=1) == label).mean()
    print("final loss %.4f acc %.3f" % (losses[-1], acc))
    assert acc > 0.85


def test_selector_install_and_fallback(hml, tmp_path):
    M, N, K = 1 << 16, 16, 4096
    mask = hml.selector_mask(M, N, K)
    target_v = (1, 2, 1, 4, 4)
    cls = hml.selector_class_of(M, N, K, target_v)
    assert cls >= 0 and mask[cls] == 1
    assert hml.selector_variant(M, N, K, cls).astuple() == target_v
    # teach a net to always answer `cls`
    rng = np.random.default_rng(3)
    X = np.stack([hml.selector_features(int(m), int(n), int(k))
                  for m, n, k in zip(2 ** rng.integers(12, 20, 400),
                                     2 ** rng.integers(0, 6, 400),
                                     2 ** rng.integers(8, 14, 400))])
    T = np.zeros((400, hml.lib.hmlSelectorClassCount()), np.float32)
    T[:, cls] = 1
    net = hml.Mlp(seed=2)
    net.train(X, T, np.ones_like(T), epochs=40, batch=100, lr=3e-3)
    hml.selector_install(net)
    try:
        v, src = hml.sgemm_select(M, N, K)
        assert src == 2 and v.astuple() == target_v
        # the pick runs and is exact on integer data
        A = rng.integers(-2, 3, (M, K)).astype(np.float32)
        B = rng.integers(-2, 3, (K, N)).astype(np.float32)
        dA, dB = hml.DeviceBuffer.from_numpy(A), hml.DeviceBuffer.from_numpy(B)
        dC = hml.DeviceBuffer(4 * M * N)
        hml.sgemm_ex(M, N, K, dA, dB, dC)
        np.testing.assert_array_equal(dC.to_numpy((M, N)),
                                      (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32))
        # save / load round trip keeps the decision
        path = str(tmp_path / "sel.bin")
        net.save(path)
        net2 = hml.Mlp.load(path)
        np.testing.assert_array_equal(net2.get_params(), net.get_params())
    finally:
        hml.selector_install(None)
    v, src = hml.sgemm_select(M, N, K)
    assert src == 0


def test_shipped_selector_autoload(hml, tmp_path):
    """With HML_SELECTOR naming a model file, `shipped`, or unset (the model
    under config/, when it matches the library's class space), the first
    selection in a fresh process installs it (source 2 = learned);
    HML_SELECTOR=off or 0 keeps the heuristic.  Runs in child processes so
    this process keeps its state."""
    import subprocess
    import sys
    import os
    net = hml.Mlp(seed=5)
    path = str(tmp_path / "sel.bin")
    net.save(path)
    code = ("import sys; sys.path.insert(0, %r); import hiperml_amd as h; "
            "h.check(h.lib.hmlInit()); v, s = h.sgemm_select(1 << 16, 16, 4096); "
            "print('SRC', s)" % os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    shipped_path = os.path.join(root, "config", "hml_selector_mi355x.bin")
    shipped = False
    if os.path.exists(shipped_path):
        hdr = np.fromfile(shipped_path, np.int32, 5)
        shipped = int(hdr[4]) == hml.lib.hmlSelectorClassCount()
    cases = [(path, 2), ("off", 0), ("0", 0), (None, 2 if shipped else 0),
             ("shipped", 2 if shipped else 0)]
    for setting, want in cases:
        env = dict(os.environ)
        env.pop("HML_SELECTOR", None)
        if setting is not None:
            env["HML_SELECTOR"] = setting
        r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                           text=True, timeout=120)
        print(r.stdout, r.stderr[-1000:])
        assert r.returncode == 0
        assert ("SRC %d" % want) in r.stdout
