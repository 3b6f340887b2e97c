"""ORACLE -- test infrastructure only.

CPU restatement (numpy, float64) of the order-selector network of the field
engine's learned per-pair M2L orders, which has no counterpart in the
reference.  Same parameter layout as the HIP kernels: W1[64][4], b1[64],
W2[64][64], b2[64], W3[64], b3; inputs standardised as (x - mean) * inv_std.

    y^ = W3 . relu(W2 relu(W1 xn + b1) + b2) + b3
    loss = mean_k w_k (y^_k - y_k)^2,  w_k = asym where y^_k < y_k, else 1

The gradient is pinned by central finite differences in float64
(tests/test_oracle_mlp.py), not by a restatement of itself.
"""
import numpy as np

NIN, NH = 4, 64
NPARAM = NH * NIN + NH + NH * NH + NH + NH + 1


def unpack(W):
    W = np.asarray(W, dtype=np.float64)
    o = 0
    W1 = W[o:o + NH * NIN].reshape(NH, NIN); o += NH * NIN
    b1 = W[o:o + NH]; o += NH
    W2 = W[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = W[o:o + NH]; o += NH
    W3 = W[o:o + NH]; o += NH
    b3 = W[o]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3, [b3]])


def standardise(X, norm):
    norm = np.asarray(norm, dtype=np.float64)
    return (np.asarray(X, dtype=np.float64) - norm[:NIN]) * norm[NIN:]


def forward(W, norm, X):
    W1, b1, W2, b2, W3, b3 = unpack(W)
    xn = standardise(X, norm)
    h1 = np.maximum(xn @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3 + b3


def loss_grad(W, norm, X, Y, asym=4.0):
    """Mean asymmetric squared loss and its gradient (NPARAM) at W."""
    W1, b1, W2, b2, W3, b3 = unpack(W)
    Y = np.asarray(Y, dtype=np.float64)
    B = Y.shape[0]
    xn = standardise(X, norm)
    z1 = xn @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    yh = h2 @ W3 + b3
    d = yh - Y
    w = np.where(d < 0.0, asym, 1.0)
    loss = float(np.mean(w * d * d))
    g = 2.0 * w * d / B
    gW3 = h2.T @ g
    gb3 = g.sum()
    dz2 = np.outer(g, W3) * (z2 > 0.0)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * (z1 > 0.0)
    gW1 = dz1.T @ xn
    gb1 = dz1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def loss_only(W, norm, X, Y, asym=4.0):
    d = forward(W, norm, X) - np.asarray(Y, dtype=np.float64)
    w = np.where(d < 0.0, asym, 1.0)
    return float(np.mean(w * d * d))


def init(rng, ymean=0.0):
    """He-uniform weights, zero biases, output bias at the mean label."""
    def u(fan_in, size):
        lim = np.sqrt(6.0 / fan_in)
        return rng.uniform(-lim, lim, size)
    return pack(u(NIN, (NH, NIN)), np.zeros(NH), u(NH, (NH, NH)), np.zeros(NH), u(NH, NH), ymean)
