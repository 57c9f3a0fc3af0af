"""Pins of the oracle's FP64 correction network (no reference counterpart):
closed-form single-path networks for the forward pass and central finite
differences for the gradient."""
import math

import numpy as np

NP = 8705
W1, B1, W2, B2, W3, B3, W4, B4 = 0, 256, 320, 4416, 4480, 8576, 8640, 8704


def gelu(z):
    return 0.5 * z * (1.0 + math.erf(z / math.sqrt(2.0)))


def test_gelu_known_values(oracle):
    p = np.zeros(NP)
    p[W4] = 1.0
    p[W3] = 1.0              # W3[0][0]
    p[W2] = 1.0              # W2[0][0]
    p[W1] = 1.0              # W1[0][0] (input 0)
    x = np.array([[1.0, 0, 0, 0], [0.0, 0, 0, 0], [-3.0, 0, 0, 0]])
    y = oracle.mlp_forward(p, x)
    # gelu(1) = Phi(1) = 0.841344746...
    assert abs(gelu(1.0) - 0.8413447460685429) < 1e-15
    assert abs(y[0] - gelu(gelu(gelu(1.0)))) < 1e-14
    assert y[1] == 0.0
    assert abs(y[2] - gelu(gelu(gelu(-3.0)))) < 1e-14


def test_single_path_closed_form(oracle):
    rng = np.random.default_rng(1)
    p = np.zeros(NP)
    a, b, c, d = rng.uniform(-2, 2, 4)
    b1, b2, b3, b4 = rng.uniform(-1, 1, 4)
    p[W1 + 4 * 5 + 2] = a          # feature 5 reads input 2
    p[B1 + 5] = b1
    p[W2 + 64 * 9 + 5] = b         # feature 9 of layer 2 reads feature 5
    p[B2 + 9] = b2
    p[W3 + 64 * 17 + 9] = c
    p[B3 + 17] = b3
    p[W4 + 17] = d
    p[B4] = b4
    # the other hidden features carry their biases only
    p[B3 + 3] = 0.7
    p[W4 + 3] = -0.4
    x = rng.uniform(-1, 1, (5, 4))
    y = oracle.mlp_forward(p, x)
    for m in range(5):
        h1 = gelu(a * x[m, 2] + b1)
        h2 = gelu(b * h1 + b2)
        ref = d * gelu(c * h2 + b3) + b4 + (-0.4) * gelu(0.7)
        # features with zero weights and biases contribute gelu(0) = 0
        assert abs(y[m] - ref) < 1e-13


def test_backward_matches_finite_differences(oracle):
    rng = np.random.default_rng(2)
    p = rng.standard_normal(NP) * 0.3
    x = rng.uniform(-1, 1, (7, 4))
    up = rng.standard_normal(7)
    g = oracle.mlp_backward(p, x, up)
    idx = np.concatenate([rng.choice(256, 6, replace=False) + W1, [B1 + 3, B1 + 60],
                          rng.choice(4096, 8, replace=False) + W2, [B2 + 11],
                          rng.choice(4096, 8, replace=False) + W3, [B3 + 40], [W4 + 7, W4 + 50], [B4]])
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        fp = up @ oracle.mlp_forward(q, x)
        q[i] -= 2 * h
        fm = up @ oracle.mlp_forward(q, x)
        fd = (fp - fm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(fd)), (i, fd, g[i])
