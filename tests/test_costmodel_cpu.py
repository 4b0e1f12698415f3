"""The cost-model oracle (oracle/mlp_oracle.c) pinned before it checks the HIP
kernels: its chain-rule gradient must equal central finite differences of its
own loss, and its forward pass must equal a numpy restatement of the layer
equations.  Also the block-feature function (host code in libqsv.so)."""
import numpy as np

import oracle as O
from qpanda_amd import engine as E

NP = 1505


def params(seed):
    g = np.random.default_rng(seed)
    p = g.normal(0, 0.4, NP)
    return p


def data(seed, batch=37):
    g = np.random.default_rng(seed)
    return g.uniform(0, 1, (batch, 12)), g.normal(0, 1, batch)


def test_forward_matches_numpy_layers():
    p = params(1)
    x, _ = data(2)
    W1 = p[:384].reshape(32, 12); b1 = p[384:416]
    W2 = p[416:1440].reshape(32, 32); b2 = p[1440:1472]
    W3 = p[1472:1504]; b3 = p[1504]
    h1 = np.maximum(x @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    y = h2 @ W3 + b3
    assert np.abs(O.mlp_forward(p, x) - y).max() < 1e-12


def test_gradient_matches_finite_differences():
    p = params(3)
    x, t = data(4)
    g, loss = O.mlp_grad(p, x, t)
    assert abs(loss - O.mlp_loss(p, x, t)) < 1e-14
    rng = np.random.default_rng(5)
    idx = np.concatenate([[0, 383, 384, 415, 416, 1439, 1440, 1471, 1472, 1503, 1504],
                          rng.integers(0, NP, 60)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (O.mlp_loss(pp, x, t) - O.mlp_loss(pm, x, t)) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_block_features():
    # single-block launch of qubits {3, 10, 20, 29} at tile columns 7 (11 tile bits)
    f = E.block_features(30, 4, 1, 7, [3, 10, 20, 29])
    assert f.shape == (12,)
    assert abs(f[0] - 30 / 34) < 1e-7 and f[1] == 1 and f[2] == 1 / 8
    assert f[3] == 16 / 128 and f[4] == 1 and abs(f[5] - 1 / 5) < 1e-7
    # L = {3,10,20,29} + 7 lowest others {0,1,2,4,5,6,7}: contiguous run 0..7 = 8 bits
    assert abs(f[6] - 8 / 11) < 1e-7
    assert abs(f[7] - 4 / 14) < 1e-7 and f[8] == 0
    assert abs(f[9] - 3 / 30) < 1e-7 and abs(f[10] - 29 / 30) < 1e-7
    assert abs(f[11] - 0.25) < 1e-7           # one of the top three qubits
    f5 = E.block_features(28, 5, 0, 8, [0, 1, 2, 3, 4])
    assert f5[3] == 32 / 128 and f5[4] == 0 and abs(f5[5] - 3 / 5) < 1e-7
    assert f5[6] == 1.0 and f5[8] == 1.0 and f5[11] == 0


def test_adam_closed_forms():
    # with a constant gradient the bias-corrected moments are exactly g and g^2,
    # so every step moves each parameter by -lr * g / (|g| + eps); a dropped
    # bias correction or a swapped moment fails this
    g = np.array([3.0, -2e-3, 0.0, 7e-9, -1.5])
    p0 = np.array([0.1, -0.2, 0.3, 0.4, 0.5])
    p, m, v = p0.copy(), np.zeros(5), np.zeros(5)
    lr = 3e-3
    for step in range(1, 6):
        p, m, v = O.adam_update(p, g, m, v, step, lr)
        expect = p0 - step * lr * g / (np.abs(g) + 1e-8)
        assert np.allclose(p, expect, rtol=0, atol=1e-15), step
    # and the first step of a real training run is p - lr * g / (|g| + eps)
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 1, (40, 12))
    t = rng.normal(size=40)
    params = rng.normal(scale=0.2, size=NP)
    g, _ = O.mlp_grad(params, x, t)
    p1 = O.mlp_train(params, x, t, 1, lr)
    assert np.allclose(p1, params - lr * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-14)
