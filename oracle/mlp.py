"""FP64 numpy restatement of the cost-model MLP (TEST INFRASTRUCTURE ONLY).

The network has no counterpart in the reference; it is the learned part of the
block-op scheduler.  Layout of the flat parameter vector (same as the HIP
library): W1[H, nf], b1[H], W2[H, H], b2[H], W3[H], b3[1].
Network: h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 . h2 + b3;
loss = mean((y - t)^2).
Pinned by central finite differences (tests/test_mlp_oracle.py).
"""
import numpy as np


def unpack(p, H, nf):
    o = 0
    W1 = p[o:o + H * nf].reshape(H, nf); o += H * nf
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o:o + 1]; o += 1
    return W1, b1, W2, b2, W3, b3


def forward(p, X, H, nf):
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float64), H, nf)
    Z1 = X @ W1.T + b1
    H1 = np.maximum(Z1, 0)
    Z2 = H1 @ W2.T + b2
    H2 = np.maximum(Z2, 0)
    return H2 @ W3 + b3[0], (Z1, H1, Z2, H2)


def mlp_forward_backward(p, X, t, H, nf):
    """Returns (predictions, loss, flat gradient) in FP64."""
    p = np.asarray(p, np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p, H, nf)
    y, (Z1, H1, Z2, H2) = forward(p, X, H, nf)
    n = X.shape[0]
    loss = float(np.mean((y - t) ** 2))
    G3 = 2.0 * (y - t) / n
    dW3 = G3 @ H2
    db3 = np.array([G3.sum()])
    G2 = np.outer(G3, W3) * (Z2 > 0)
    dW2 = G2.T @ H1
    db2 = G2.sum(0)
    G1 = (G2 @ W2) * (Z1 > 0)
    dW1 = G1.T @ X
    db1 = G1.sum(0)
    g = np.concatenate([dW1.reshape(-1), db1, dW2.reshape(-1), db2, dW3, db3])
    return y, loss, g


def loss_only(p, X, t, H, nf):
    y, _ = forward(p, X, H, nf)
    return float(np.mean((y - t) ** 2))
