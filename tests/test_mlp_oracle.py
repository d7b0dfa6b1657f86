"""The FP64 restatement of the Newton-start network (oracle/mlp.py) pinned
by what the mathematics fixes: central finite differences of the loss for
the backward pass, a one-neuron closed form for the forward pass, and
Adam's first step (lr * sign(g)) for the optimizer."""
import numpy as np

from oracle import mlp as omlp


def _problem(n=40, seed=1):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.4, omlp.N_PARAMS)
    feat = rng.normal(0, 1, (n, omlp.N_IN)) * [3.0, 1.0, 0.5, 0.2, 0.2, 1.0] + [1, 0, 0, 0, 0, -1]
    norm = np.concatenate([[1, 0, 0, 0, 0, -1], 1 / np.array([3.0, 1.0, 0.5, 0.2, 0.2, 1.0])])
    targ = rng.normal(0, 1, n)
    return p, feat, targ, norm


def _fd(p, feat, targ, norm, idx, h=1e-6):
    out = []
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        out.append((omlp.loss(pp, feat, targ, norm) - omlp.loss(pm, feat, targ, norm)) / (2 * h))
    return np.array(out)


def _probe_indices(seed=2):
    rng = np.random.default_rng(seed)
    blocks = [(omlp.OFF_W1, omlp.OFF_B1), (omlp.OFF_B1, omlp.OFF_W2), (omlp.OFF_W2, omlp.OFF_B2),
              (omlp.OFF_B2, omlp.OFF_W3), (omlp.OFF_W3, omlp.OFF_B3), (omlp.OFF_B3, omlp.N_PARAMS)]
    idx = []
    for a, b in blocks:
        idx += list(rng.integers(a, b, min(12, b - a)))
    return np.array(idx)


def test_gradient_matches_finite_differences():
    p, feat, targ, norm = _problem()
    loss, g = omlp.loss_grad(p, feat, targ, norm)
    assert abs(loss - omlp.loss(p, feat, targ, norm)) <= 1e-14 * loss
    idx = _probe_indices()
    fd = _fd(p, feat, targ, norm, idx)
    assert np.abs(g[idx] - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())


def test_finite_differences_catch_a_transposed_weight_gradient():
    """A planted mistake in the oracle (dW2 transposed) fails the check."""
    p, feat, targ, norm = _problem(seed=3)
    _, g = omlp.loss_grad(p, feat, targ, norm)
    bad = g.copy()
    w2 = bad[omlp.OFF_W2:omlp.OFF_B2].reshape(omlp.N_HID, omlp.N_HID)
    bad[omlp.OFF_W2:omlp.OFF_B2] = w2.T.ravel()
    idx = np.arange(omlp.OFF_W2, omlp.OFF_W2 + 200)
    fd = _fd(p, feat, targ, norm, idx)
    assert np.abs(g[idx] - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())
    assert np.abs(bad[idx] - fd).max() > 1e-3 * np.abs(fd).max()


def test_forward_one_neuron_closed_form():
    p = np.zeros(omlp.N_PARAMS)
    W1 = p[omlp.OFF_W1:omlp.OFF_B1].reshape(omlp.N_HID, omlp.N_IN)
    W1[5, 2] = 2.0                           # neuron 5 reads feature 2
    p[omlp.OFF_B1 + 5] = -0.5
    W2 = p[omlp.OFF_W2:omlp.OFF_B2].reshape(omlp.N_HID, omlp.N_HID)
    W2[9, 5] = 3.0                           # neuron 9 of layer 2 reads neuron 5
    p[omlp.OFF_W3 + 9] = 0.25
    p[omlp.OFF_B3] = 0.125
    x = np.linspace(-2, 2, 41)
    feat = np.zeros((41, omlp.N_IN))
    feat[:, 2] = x
    norm = np.concatenate([np.zeros(6), np.ones(6)])
    want = 0.25 * 3.0 * np.maximum(2.0 * x - 0.5, 0.0) + 0.125
    np.testing.assert_allclose(omlp.forward(p, feat, norm), want, rtol=0, atol=1e-15)
    # normalisation is (x - mean) * inv_std
    norm2 = norm.copy()
    norm2[2], norm2[8] = 1.0, 0.5
    want2 = 0.25 * 3.0 * np.maximum(2.0 * (x - 1.0) * 0.5 - 0.5, 0.0) + 0.125
    np.testing.assert_allclose(omlp.forward(p, feat, norm2), want2, rtol=0, atol=1e-15)


def test_adam_first_step_is_signed_learning_rate():
    rng = np.random.default_rng(4)
    g = rng.normal(0, 1, 100)
    p = rng.normal(0, 1, 100)
    p1, m, v = omlp.adam(p, g, np.zeros(100), np.zeros(100), 1, lr=1e-2, eps=0.0)
    np.testing.assert_allclose(p1, p - 1e-2 * np.sign(g), rtol=0, atol=1e-15)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)
