"""CPU restatement of the cardinality-estimator MLP -- TEST INFRASTRUCTURE ONLY.

float64 forward and backward of the 40 -> 64 -> 64 -> 1 ReLU network with the device's
BF16 rounding (round to nearest even) applied where operands enter its MFMA products, so the
device's FP32-accumulated results agree to accumulation-order rounding.  Parameter layout:
W1[64][48], b1[64], W2[64][64], b2[64], w3[64], b3.
"""
import numpy as np

IN, K1, H = 40, 48, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * K1
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NUM_PARAMS = OFF_B3 + 1


def bf16(x, rounding=True):
    """Round float32 values to BF16 (nearest even) and return them as float64."""
    if not rounding:
        return np.asarray(x, dtype=np.float64)
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float32)
    return (p[OFF_W1:OFF_B1].reshape(H, K1), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, X, rounding=True):
    """rounding=False gives the exact float64 network (for finite-difference checks)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    n = X.shape[0]
    Xp = np.zeros((n, K1), np.float64)
    Xp[:, :IN] = X
    xb = bf16(Xp.astype(np.float32), rounding) if rounding else Xp
    z1 = xb @ bf16(W1, rounding).T + b1
    h1 = np.maximum(z1, 0.0)
    h1 = bf16(h1.astype(np.float32), rounding) if rounding else h1
    z2 = h1 @ bf16(W2, rounding).T + b2
    h2 = np.maximum(z2, 0.0)
    pred = h2 @ w3.astype(np.float64) + float(b3)
    return pred, (xb, z1, h1, z2, h2)


def gradients(p, X, y, rounding=True):
    """Gradients of mean((pred - y)^2) in the parameter layout, and the loss."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    pred, (xb, z1, h1, z2, h2) = forward(p, X, rounding)
    n = X.shape[0]
    e = pred - np.asarray(y, dtype=np.float64)
    dpred = 2.0 * e / n
    g = np.zeros(NUM_PARAMS, np.float64)
    g[OFF_W3:OFF_B3] = dpred @ h2
    g[OFF_B3] = dpred.sum()
    dz2 = dpred[:, None] * w3[None, :].astype(np.float64) * (h2 > 0)
    dz2 = bf16(dz2.astype(np.float32)) if rounding else dz2
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W2:OFF_B2] = (dz2.T @ h1).ravel()
    dh1 = dz2 @ bf16(W2, rounding)
    dz1 = dh1 * (z1 > 0)
    dz1 = bf16(dz1.astype(np.float32)) if rounding else dz1
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    gw1 = dz1.T @ xb
    gw1[:, IN:] = 0.0
    g[OFF_W1:OFF_B1] = gw1.ravel()
    return g, float(np.mean(e * e))
