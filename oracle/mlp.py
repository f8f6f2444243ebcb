"""FP64 oracle of the entry-predictor MLP (TEST INFRASTRUCTURE ONLY).

Mirrors the numerics of the HIP kernels: the input, every activation that
feeds a GEMM and every weight are rounded to bf16 (round to nearest even);
everything else is exact FP64.  The remaining GPU/oracle difference is the
GPU's FP32 accumulation, so gradients agree to ~K * 2^-24 relative.
"""
import numpy as np


def bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _exact(x):
    return np.asarray(x, np.float64)


def normalise(X, mean=None, inv=None):
    """(x - mean) * inv_std in FP32, as the input kernel computes it."""
    if mean is None:
        return X
    X = np.asarray(X, np.float32)
    return ((X - np.asarray(mean, np.float32)).astype(np.float32)
            * np.asarray(inv, np.float32)).astype(np.float32)


def forward(W, b, X, q=bf16, norm=(None, None)):
    """q is the rounding applied where the GPU rounds to bf16 (identity for
    the exact network); norm = (mean, inv_std) of the input."""
    Wq = [q(w) for w in W]
    x0 = q(normalise(X, *norm))
    z1 = x0 @ Wq[0] + b[0].astype(np.float64)
    h1 = np.maximum(z1, 0.0)
    h1f = h1.astype(np.float32) if q is bf16 else h1  # stored FP32 on the GPU
    z2 = q(h1f) @ Wq[1] + b[1].astype(np.float64)
    h2 = np.maximum(z2, 0.0)
    h2f = h2.astype(np.float32) if q is bf16 else h2
    z3 = q(h2f) @ Wq[2] + b[2].astype(np.float64)
    return dict(Wq=Wq, x0=x0, h1=h1f, h2=h2f, z3=z3)


def grads(W, b, X, y, q=bf16, norm=(None, None)):
    """Gradients of the mean softmax cross-entropy: (gW list, gb list, loss).
    With q = bf16 the roundings are those of the GPU and the gradient is the
    straight-through one; with q = _exact it is the exact FP64 gradient."""
    f = forward(W, b, X, q, norm)
    n = X.shape[0]
    z = f["z3"].astype(np.float32).astype(np.float64) if q is bf16 else f["z3"]
    m = z.max(1, keepdims=True)
    e = np.exp(z - m)
    p = e / e.sum(1, keepdims=True)
    loss = float(np.mean(np.log(e.sum(1)) + m[:, 0] - z[np.arange(n), y]))
    dz = p.copy()
    dz[np.arange(n), y] -= 1.0
    dz /= n
    rq = (lambda a: bf16(a.astype(np.float32))) if q is bf16 else _exact
    dzq = rq(dz)
    gW3 = q(f["h2"]).T @ dzq
    gb3 = dz.sum(0)
    dh2 = (dzq @ f["Wq"][2].T) * (f["h2"] > 0)
    dh2q = rq(dh2)
    gW2 = q(f["h1"]).T @ dh2q
    gb2 = dh2.sum(0)
    dh1 = (dh2q @ f["Wq"][1].T) * (f["h1"] > 0)
    dh1q = rq(dh1)
    gW1 = f["x0"].T @ dh1q
    gb1 = dh1.sum(0)
    return [gW1, gW2, gW3], [gb1, gb2, gb3], loss


def predict(W, b, X, fallback=None, min_conf=0.5, norm=(None, None)):
    z = forward(W, b, X, norm=norm)["z3"]
    m = z.max(1, keepdims=True)
    p = 1.0 / np.exp(z - m).sum(1)
    leaf = z.argmax(1).astype(np.int32)
    if fallback is not None:
        leaf = np.where(p >= min_conf, leaf, fallback).astype(np.int32)
    return leaf, p
