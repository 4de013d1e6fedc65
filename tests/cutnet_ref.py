"""NumPy restatement of the learned cut's MLP, for checking the HIP kernels.

Same parameter vector as include/arc.h (ARC_CUTNET_PARAMS floats) and the
same roundings as the device code: inputs normalised in FP32, then rounded
to BF16 (nearest even); weights rounded to BF16; layer-1 activations rounded
to BF16 before layer 2; layer-2 activations and w3 in FP32.  Products are
summed in FP64 here and in FP32 on the matrix cores, so results agree to the
FP32 summation error (about 1e-6 relative), except where that error moves a
BF16 rounding of a hidden activation by one unit.
"""
import numpy as np

IN, HID = 24, 64
MEAN, ISTD, W1 = 0, 24, 48
B1 = W1 + 32 * HID
W2 = B1 + HID
B2 = W2 + HID * HID
W3 = B2 + HID
B3 = W3 + HID
PARAMS = 6400


def bf16(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return u.view(np.float32)


def unpack(w):
    w = np.asarray(w, np.float32)
    return dict(mean=w[MEAN:MEAN + IN], istd=w[ISTD:ISTD + IN], W1=w[W1:B1].reshape(32, HID)[:IN],
                b1=w[B1:W2], W2=w[W2:B2].reshape(HID, HID), b2=w[B2:W3], w3=w[W3:B3], b3=w[B3])


def forward(w, X, keep=False):
    p = unpack(w)
    xn = bf16((np.asarray(X, np.float32) - p["mean"]) * p["istd"]).astype(np.float64)
    a1 = xn @ bf16(p["W1"]).astype(np.float64) + p["b1"]
    h1 = bf16(np.maximum(a1, 0.0).astype(np.float32)).astype(np.float64)
    a2 = h1 @ bf16(p["W2"]).astype(np.float64) + p["b2"]
    h2 = np.maximum(a2, 0.0)
    z = h2 @ p["w3"].astype(np.float64) + float(p["b3"])
    prob = 1.0 / (1.0 + np.exp(-z))
    if keep:
        return prob, dict(xn=xn, h1=h1, h2=h2, z=z)
    return prob


def grad(w, X, y, wt=None):
    """Gradient of the mean weighted binary cross-entropy, with the device's
    BF16 roundings of the backward operands (dH2, dH1, H1, X, W2)."""
    p = unpack(w)
    prob, t = forward(w, X, keep=True)
    n = len(y)
    wt = np.ones(n) if wt is None else np.asarray(wt, np.float64)
    dz = (prob - y) * wt / n
    h2, h1, xn = t["h2"], t["h1"], t["xn"]
    g = np.zeros(PARAMS)
    g[W3:B3] = h2.T @ dz
    g[B3] = dz.sum()
    dh2 = np.where(h2 > 0, np.outer(dz, p["w3"]), 0.0)
    g[B2:W3] = dh2.sum(0)
    dh2b = bf16(dh2.astype(np.float32)).astype(np.float64)
    g[W2:B2] = (h1.T @ dh2b).ravel()
    dh1 = np.where(h1 > 0, dh2b @ bf16(p["W2"]).astype(np.float64).T, 0.0)
    g[B1:W2] = dh1.sum(0)
    dh1b = bf16(dh1.astype(np.float32)).astype(np.float64)
    gw1 = np.zeros((32, HID))
    gw1[:IN] = xn.T @ dh1b
    g[W1:B1] = gw1.ravel()
    return g


def random_weights(rng, scale=1.0):
    w = np.zeros(PARAMS, np.float32)
    w[MEAN:MEAN + IN] = rng.normal(0, 0.3, IN)
    w[ISTD:ISTD + IN] = rng.uniform(0.5, 2.0, IN)
    w1 = np.zeros((32, HID), np.float32)
    w1[:IN] = rng.normal(0, scale * np.sqrt(2 / IN), (IN, HID))
    w[W1:B1] = w1.ravel()
    w[B1:W2] = rng.normal(0, 0.1, HID)
    w[W2:B2] = rng.normal(0, scale * np.sqrt(2 / HID), HID * HID)
    w[B2:W3] = rng.normal(0, 0.1, HID)
    w[W3:B3] = rng.normal(0, np.sqrt(1 / HID), HID)
    w[B3] = 0.1
    return w
