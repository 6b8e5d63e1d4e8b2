"""CPU oracle of the learned selector's MLP (TEST INFRASTRUCTURE ONLY).

numpy restatement, in float64, of what the HIP kernels of
hiperml_amd/csrc/hml_k_mlp.hip compute: dense layers with weights stored
[in][out] followed by the bias, ReLU between layers, a masked softmax over
the output logits, the soft-target cross-entropy averaged over the batch,
and its gradient.  Parameter vector layout: for each layer l, W_l (in x out,
row-major) then b_l (out).

Pinned by central finite differences in tests/test_mlp_oracle.py (the
analytic gradient of this restatement against its own loss), and by closed
forms (a zero network gives a uniform softmax; the loss of a one-hot target
is -log p).
"""
import numpy as np


def layer_dims(n_in, hidden, n_hidden, n_out):
    return [n_in] + [hidden] * n_hidden + [n_out]


def unpack(params, dims):
    out, off = [], 0
    for l in range(1, len(dims)):
        i, o = dims[l - 1], dims[l]
        W = params[off:off + i * o].reshape(i, o)
        off += i * o
        b = params[off:off + o]
        off += o
        out.append((W, b))
    assert off == len(params)
    return out


def forward(params, dims, X):
    acts = [np.asarray(X, np.float64)]
    layers = unpack(np.asarray(params, np.float64), dims)
    for l, (W, b) in enumerate(layers):
        z = acts[-1] @ W + b
        acts.append(np.maximum(z, 0.0) if l < len(layers) - 1 else z)
    return acts


def masked_softmax(Z, mask):
    Z = np.where(mask > 0, Z, -np.inf)
    mx = Z.max(axis=1, keepdims=True)
    e = np.where(mask > 0, np.exp(Z - mx), 0.0)
    return e / e.sum(axis=1, keepdims=True)


def loss(params, dims, X, T, mask):
    Z = forward(params, dims, X)[-1]
    P = masked_softmax(Z, mask)
    lp = np.log(np.where(mask > 0, P, 1.0))
    return float(-(T * lp * (mask > 0)).sum() / X.shape[0])


def loss_grad(params, dims, X, T, mask):
    """Mean loss and d(mean loss)/d(params) in the library's layout."""
    B = X.shape[0]
    acts = forward(params, dims, X)
    layers = unpack(np.asarray(params, np.float64), dims)
    P = masked_softmax(acts[-1], mask)
    lp = np.log(np.where(mask > 0, P, 1.0))
    L = float(-(T * lp * (mask > 0)).sum() / B)
    dz = np.where(mask > 0, P - T, 0.0) / B
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        grads.append((acts[l].T @ dz, dz.sum(axis=0)))
        if l > 0:
            dz = (dz @ W.T) * (acts[l] > 0)
    grads.reverse()
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return L, flat
