"""CPU restatement of the learned support proposer's network (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
forward/backward kernels (cuda-k-svd-denoising_amd/csrc/mlp.hip), never by the
product path.  The proposer has no counterpart in the reference; this module
states the network the GPU trains:

  h1 = ReLU(W1 [x, 1]),  h2 = ReLU(W2 [h1, 1]),  z = W3 [h2, 1]
  loss = mean over the batch of sum over atoms of softplus(z) - t z

with each layer's bias stored as the weight column that multiplies the
constant-1 entry.  `rounding="bf16"` rounds at the points where the GPU
stores BF16 (weights, inputs, h1, h2, dz, dh2, dh1) and accumulates in
float64; `rounding=None` is the plain float64 network, whose gradients are
pinned by central finite differences in the tests.
"""
import numpy as np


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(a)
    out[nan] = np.nan
    return out


def _r(a, rounding):
    return bf16(a) if rounding == "bf16" else np.asarray(a, dtype=np.float64)


def forward_backward(W1, W2, W3, X, T, n_atoms, hidden=512, rounding="bf16"):
    """Loss and gradients of one batch.

    W1 [hidden x kin], W2 [hidden x hp], W3 [ko x hp] with the bias in
    column `dim` of W1 (the column where X holds 1) and in column `hidden`
    of W2 and W3; X [nb x kin] (the constant-1 column included), T [nb x ko]
    0/1 targets; only the first n_atoms logits enter the loss.
    Returns (loss, dW1, dW2, dW3) with the weight layouts.
    """
    nb = X.shape[0]
    hp = W2.shape[1]
    w1, w2, w3 = (_r(w, rounding) for w in (W1, W2, W3))
    x = _r(X, rounding)

    def aug(h):
        a = np.zeros((nb, hp))
        a[:, :hidden] = h
        a[:, hidden] = 1.0
        return a

    h1 = _r(np.maximum(x @ w1.T, 0.0), rounding)
    h1a = aug(h1)
    h2 = _r(np.maximum(h1a @ w2.T, 0.0), rounding)
    h2a = aug(h2)
    z = h2a @ w3.T
    if rounding == "bf16":
        z = z.astype(np.float32).astype(np.float64)
    t = np.asarray(T, dtype=np.float64)
    live = np.zeros(z.shape[1], bool)
    live[:n_atoms] = True
    sp = np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z)))
    loss = float(np.sum((sp - t * z)[:, live]) / nb)
    dz = (1.0 / (1.0 + np.exp(-z)) - t) / nb
    dz[:, ~live] = 0.0
    dz = _r(dz, rounding)
    dW3 = dz.T @ h2a
    dh2 = _r((dz @ w3)[:, :hidden] * (h2 > 0), rounding)
    dW2 = dh2.T @ h1a
    dh1 = _r((dh2 @ w2)[:, :hidden] * (h1 > 0), rounding)
    dW1 = dh1.T @ x
    return loss, dW1, dW2, dW3


def loss_only(W1, W2, W3, X, T, n_atoms, hidden):
    return forward_backward(W1, W2, W3, X, T, n_atoms, hidden, rounding=None)[0]
