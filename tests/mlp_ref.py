"""CPU fp32 reference of the per-pixel error model (fg_mlp.hip), for tests.

The device kernels feed BF16 operands to the matrix instructions and
accumulate in FP32; this reference rounds the same operands to BF16 (round to
nearest even) at the same points and accumulates in float64, so device and
reference differ only by FP32 accumulation order.
"""
import numpy as np

P_W1, P_B1 = 0, 16 * 64
P_W2 = P_B1 + 64
P_B2 = P_W2 + 64 * 64
P_W3 = P_B2 + 64
P_B3 = P_W3 + 64
N_PARAMS = P_B3 + 1


def bf16(x):
    """Round f32 values to the nearest BF16 (ties to even), returned as f32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def bf16_bits_to_f32(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def unpack(P):
    P = np.asarray(P)
    if P.dtype != np.float64:
        P = P.astype(np.float32)
    return (P[P_W1:P_B1].reshape(16, 64), P[P_B1:P_W2], P[P_W2:P_B2].reshape(64, 64), P[P_B2:P_W3],
            P[P_W3:P_B3], P[P_B3])


def forward(P, X):
    """X: (n, 16) float32 holding BF16 values."""
    W1, b1, W2, b2, w3, b3 = unpack(P)
    X = X.astype(np.float64)
    z1 = X @ bf16(W1).astype(np.float64) + b1
    h1 = bf16(np.maximum(z1, 0)).astype(np.float64)
    z2 = h1 @ bf16(W2).astype(np.float64) + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ bf16(w3).astype(np.float64) + b3
    return z1, h1, z2, h2, y


def loss_terms(y, t, loss="squared"):
    """Per-row loss and its derivative in y: squared error (y - t)^2, or the
    Gaussian negative log-likelihood y + t^2 exp(-2y) / 2 of a residual t
    under the predicted log std y (exponent clamped to [-80, 80])."""
    if loss == "nll":
        e = np.exp(np.clip(-2.0 * y, -80.0, 80.0)) * t * t
        return y + 0.5 * e, 1.0 - e
    return (y - t) ** 2, 2.0 * (y - t)


def grad(P, X, t, n_total=None, loss="squared"):
    """Gradient of (1/n_total) sum L(y, t), packed like the parameters, and
    the loss sum."""
    W1, b1, W2, b2, w3, b3 = unpack(P)
    n_total = n_total or X.shape[0]
    z1, h1, z2, h2, y = forward(P, X)
    lv, dl = loss_terms(y, np.asarray(t, np.float64), loss)
    g = dl / n_total
    dw3 = h2.T @ g
    db3 = g.sum()
    dz2 = np.where(z2 > 0, g[:, None] * bf16(w3).astype(np.float64)[None, :], 0.0)
    db2 = dz2.sum(0)
    dz2b = bf16(dz2).astype(np.float64)
    dW2 = h1.T @ dz2b
    dh1 = dz2b @ bf16(W2).astype(np.float64).T
    dz1 = np.where(z1 > 0, dh1, 0.0)
    db1 = dz1.sum(0)
    dW1 = X.astype(np.float64).T @ bf16(dz1).astype(np.float64)
    out = np.zeros(N_PARAMS)
    out[P_W1:P_B1] = dW1.ravel()
    out[P_B1:P_W2] = db1
    out[P_W2:P_B2] = dW2.ravel()
    out[P_B2:P_W3] = db2
    out[P_W3:P_B3] = dw3
    out[P_B3] = db3
    return out, float(lv.sum())


def features(u, out_w, out_h, zoom, radius, sd, sigma, channel):
    """The 16 features per output pixel (f32 values before BF16 rounding)."""
    u = np.asarray(u, dtype=np.float32)
    in_h, in_w = u.shape
    iz = np.float32(1.0) / np.float32(zoom)
    xs = np.clip(np.floor((np.arange(out_w, dtype=np.float32) + np.float32(0.5)) * iz).astype(int), 0, in_w - 1)
    ys = np.clip(np.floor((np.arange(out_h, dtype=np.float32) + np.float32(0.5)) * iz).astype(int), 0, in_h - 1)
    pad = np.pad(u, 1, mode="edge")
    def at(dx, dy):
        return pad[np.ix_(ys + 1 + dy, xs + 1 + dx)]
    win = np.stack([at(dx, dy) for dy in (-1, 0, 1) for dx in (-1, 0, 1)])
    mean = win.mean(0)
    var = np.maximum((win ** 2).mean(0) - mean ** 2, 0)
    gx, gy = 0.5 * (at(1, 0) - at(-1, 0)), 0.5 * (at(0, 1) - at(0, -1))
    c = at(0, 0)
    f = np.zeros((out_h, out_w, 16), dtype=np.float32)
    f[..., 0], f[..., 1], f[..., 2] = c, mean, np.sqrt(var)
    f[..., 3], f[..., 4] = np.sqrt(gx * gx + gy * gy), c * (1 - c)
    f[..., 5], f[..., 6], f[..., 7], f[..., 8] = radius / 0.3, sd / 0.15, sigma / 2.0, zoom / 4.0
    f[..., 9 + channel] = 1.0
    return f.reshape(-1, 16)
