"""Pins the numpy MLP oracle (oracle/mlp_oracle.py) that the GPU training
kernels are checked against: analytic gradients equal central finite
differences of the oracle's own loss; closed forms of the loss."""
import numpy as np

import mlp_oracle as mo


def problem(seed=0, B=7, dims=(6, 5, 4, 9)):
    rng = np.random.default_rng(seed)
    n = sum(dims[i - 1] * dims[i] + dims[i] for i in range(1, len(dims)))
    params = rng.normal(0, 0.7, n)
    X = rng.normal(0, 1, (B, dims[0]))
    mask = (rng.random((B, dims[-1])) < 0.7).astype(np.float64)
    mask[:, 0] = 1.0
    T = rng.random((B, dims[-1])) * mask
    T /= T.sum(axis=1, keepdims=True)
    return list(dims), params, X, T, mask


def test_gradient_matches_finite_differences():
    dims, params, X, T, mask = problem()
    L, g = mo.loss_grad(params, dims, X, T, mask)
    assert abs(L - mo.loss(params, dims, X, T, mask)) < 1e-12
    h = 1e-6
    fd = np.empty_like(params)
    for i in range(len(params)):
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd[i] = (mo.loss(p1, dims, X, T, mask) - mo.loss(p2, dims, X, T, mask)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)


def test_closed_forms():
    dims = [6, 4, 4, 5]
    n = sum(dims[i - 1] * dims[i] + dims[i] for i in range(1, len(dims)))
    X = np.ones((3, 6))
    mask = np.array([[1, 1, 1, 1, 1], [1, 0, 1, 0, 0], [0, 0, 0, 1, 0]], float)
    T = mask / mask.sum(axis=1, keepdims=True)
    P = mo.masked_softmax(mo.forward(np.zeros(n), dims, X)[-1], mask)
    np.testing.assert_allclose(P, T)                       # zero net: uniform over the mask
    onehot = np.zeros((3, 5)); onehot[0, 2] = onehot[1, 0] = onehot[2, 3] = 1
    want = -(np.log(1 / 5) + np.log(1 / 2) + np.log(1.0)) / 3
    assert abs(mo.loss(np.zeros(n), dims, X, onehot, mask) - want) < 1e-12


def test_forward_hand_computed():
    """A 2-2-2 net with nonzero weights, worked by hand: fixes the [in][out]
    weight layout, the bias order and ReLU on the hidden layer only.
    Hidden: z = [1*1 + 2*3 + 0.5, 1*(-2) + 2*0.5 - 1] = [7.5, -2] -> [7.5, 0];
    output: [7.5*2 + 0*(-1) + 0, 7.5*(-1) + 0*4 + 2] = [15, -5.5] (no ReLU)."""
    dims = [2, 2, 2]
    W1 = [1.0, -2.0, 3.0, 0.5]      # W1[in][out]
    b1 = [0.5, -1.0]
    W2 = [2.0, -1.0, -1.0, 4.0]
    b2 = [0.0, 2.0]
    params = np.array(W1 + b1 + W2 + b2)
    acts = mo.forward(params, dims, np.array([[1.0, 2.0]]))
    np.testing.assert_array_equal(acts[1], [[7.5, 0.0]])
    np.testing.assert_array_equal(acts[2], [[15.0, -5.5]])
    P = mo.masked_softmax(acts[2], np.ones((1, 2)))
    np.testing.assert_allclose(P, [[1 / (1 + np.exp(-20.5)), 1 / (1 + np.exp(20.5))]])


def test_library_host_selection_matches_oracle(tmp_path):
    """The library runs the installed selector's forward pass on the host
    (selection never touches the device).  A random network saved in the
    library's file format and loaded through HML_SELECTOR must pick, for
    every shape, the allowed class of largest logit under the oracle's
    forward pass.  Runs in a child process (the selector loads once per
    process); needs no GPU."""
    import json
    import os
    import subprocess
    import sys
    import hiperml_amd as hml
    C = hml.lib.hmlSelectorClassCount()
    dims = [6, 64, 64, 64, C]
    rng = np.random.default_rng(17)
    n = sum(dims[i - 1] * dims[i] + dims[i] for i in range(1, len(dims)))
    params = rng.normal(0, 0.5, n).astype(np.float32)
    path = str(tmp_path / "net.bin")
    with open(path, "wb") as f:
        np.array([0x504C4D48, 6, 64, 3, C], np.int32).tofile(f)
        params.tofile(f)
    shapes = []
    for _ in range(120):
        M, N, K = (int(2 ** rng.uniform(4, 16)), int(2 ** rng.uniform(0, 9)),
                   int(2 ** rng.uniform(4, 16)))
        shapes.append((M, N, K, bool(rng.integers(2)), bool(rng.integers(2))))
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys, json; sys.path.insert(0, %r); import hiperml_amd as h; "
            "out = [h.sgemm_select(m, n, k, transA=a, transB=b) for m, n, k, a, b in %r]; "
            "print(json.dumps([[v.astuple(), s] for v, s in out]))" % (root, shapes))
    env = dict(os.environ, HML_SELECTOR=path)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    picks = json.loads(r.stdout.strip().splitlines()[-1])
    for (M, N, K, ta, tb), (v, src) in zip(shapes, picks):
        assert src == 2
        feats = hml.selector_features(M, N, K, transA=ta, transB=tb)
        mask = hml.selector_mask(M, N, K, transA=ta, transB=tb)
        z = mo.forward(params.astype(np.float64), dims, np.array([feats]))[-1][0]
        z = np.where(np.asarray(mask) > 0, z, -np.inf)
        want = int(np.argmax(z))
        got = hml.selector_class_of(M, N, K, tuple(v), transA=ta, transB=tb)
        assert got == want, ((M, N, K, ta, tb), got, want)
