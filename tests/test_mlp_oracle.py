"""The learned proposer's CPU oracle (oracle/mlp_oracle.py), pinned without a GPU:
float64 gradients against central finite differences, BF16 rounding against
torch's bfloat16 conversion."""
import numpy as np
import pytest
import torch

import mlp_oracle as M


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 10000),
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -1.0 - 2 ** -8, 0.0, -0.0, 65504.0],
                                 np.float32)])
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(M.bf16(a), ref)


def small_net(rng, dim=5, hidden=8, n_atoms=5, nb=4):
    kin, hp, ko = dim + 3, hidden + 2, n_atoms + 2
    W1 = np.zeros((hidden, kin)); W1[:, :dim + 1] = rng.standard_normal((hidden, dim + 1))
    W2 = np.zeros((hidden, hp)); W2[:, :hidden + 1] = rng.standard_normal((hidden, hidden + 1)) * 0.5
    W3 = np.zeros((ko, hp)); W3[:n_atoms, :hidden + 1] = rng.standard_normal((n_atoms, hidden + 1)) * 0.5
    X = np.zeros((nb, kin)); X[:, :dim] = rng.standard_normal((nb, dim)); X[:, dim] = 1.0
    T = np.zeros((nb, ko)); T[:, :n_atoms] = rng.random((nb, n_atoms)) < 0.4
    return W1, W2, W3, X, T, n_atoms, hidden


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    W1, W2, W3, X, T, na, hid = small_net(rng)
    loss, g1, g2, g3 = M.forward_backward(W1, W2, W3, X, T, na, hid, rounding=None)
    h = 1e-6
    for W, g in ((W1, g1), (W2, g2), (W3, g3)):
        fd = np.zeros_like(W)
        for idx in np.ndindex(W.shape):
            old = W[idx]
            W[idx] = old + h
            lp = M.loss_only(W1, W2, W3, X, T, na, hid)
            W[idx] = old - h
            lm = M.loss_only(W1, W2, W3, X, T, na, hid)
            W[idx] = old
            fd[idx] = (lp - lm) / (2 * h)
        # ReLU kinks are measure-zero for random inputs
        np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-7)


def test_loss_closed_form_and_padding():
    # zero network: every logit 0, loss = n_atoms * log 2, gradient dz = (1/2 - t)/nb
    rng = np.random.default_rng(2)
    W1, W2, W3, X, T, na, hid = small_net(rng)
    W3[:] = 0.0
    loss, _, _, g3 = M.forward_backward(W1, W2, W3, X, T, na, hid, rounding=None)
    assert loss == pytest.approx(na * np.log(2.0))
    # padding outputs (rows past n_atoms) get no gradient
    assert np.all(g3[na:] == 0.0)
