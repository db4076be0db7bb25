"""CPU oracle of the learned-guess network (test infrastructure only).

numpy restatement of the network of include/pk_mlp.h: n_in -> hidden -> hidden
-> n_out with tanh, parameters packed W1 [n_in][hidden], b1, W2, b2, W3
[hidden][n_out], b3; loss L = mean over rows and outputs of (y - t)^2; Adam with
bias correction.  Computed in float64.  Pinned in tests/test_guess_cpu.py
(test_mlp_oracle_gradient_by_finite_differences) by central finite
differences of the loss.  No counterpart in the reference.
"""
import numpy as np


def unpack(p, n_in, hid, n_out):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    out = []
    for shape in [(n_in, hid), (hid,), (hid, hid), (hid,), (hid, n_out), (n_out,)]:
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    assert o == p.size
    return out


def forward(p, X, n_in, hid, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, hid, n_out)
    X = np.asarray(X, dtype=np.float64)
    h1 = np.tanh(X @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    return h2 @ W3 + b3, (X, h1, h2)


def loss_grad(p, X, T, n_in, hid, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, hid, n_out)
    y, (X, h1, h2) = forward(p, X, n_in, hid, n_out)
    d = y - np.asarray(T, dtype=np.float64)
    n = d.size
    loss = float((d * d).sum() / n)
    dy = 2.0 * d / n
    gW3, gb3 = h2.T @ dy, dy.sum(0)
    d2 = (dy @ W3.T) * (1.0 - h2 * h2)
    gW2, gb2 = h1.T @ d2, d2.sum(0)
    d1 = (d2 @ W2.T) * (1.0 - h1 * h1)
    gW1, gb1 = X.T @ d1, d1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, g


def adam(p, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    p = p - lr * (m1 / (1 - b1 ** step)) / (np.sqrt(m2 / (1 - b2 ** step)) + eps)
    return p, m1, m2
