"""ORACLE — test infrastructure only.  float64 numpy restatement of the MLP of
include/op2amd_costmodel.h (nin -> H -> H -> 1, ReLU, mean squared error),
with the same flat parameter layout, used to check the HIP kernels' forward
pass and gradients.  Its own backward pass is pinned by central finite
differences (tests/test_costmodel.py)."""
import numpy as np


def unpack(p, nin, H):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + nin * H].reshape(nin, H); o += nin * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, x, nin, H):
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, H)
    z1 = x @ W1 + b1
    a1 = np.maximum(z1, 0)
    z2 = a1 @ W2 + b2
    a2 = np.maximum(z2, 0)
    return a2 @ W3 + b3, (z1, a1, z2, a2)


def loss_grad(p, x, t, nin, H):
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, H)
    y, (z1, a1, z2, a2) = forward(p, x, nin, H)
    n = x.shape[0]
    loss = np.mean((y - t) ** 2)
    dy = 2 * (y - t) / n
    dW3 = a2.T @ dy
    db3 = dy.sum()
    d2 = np.outer(dy, W3) * (z2 > 0)
    dW2 = a1.T @ d2
    db2 = d2.sum(axis=0)
    d1 = (d2 @ W2.T) * (z1 > 0)
    dW1 = x.T @ d1
    db1 = d1.sum(axis=0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3, [db3]])
    return loss, g
