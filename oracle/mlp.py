"""CPU restatement of the criterion-ranker MLP (TEST INFRASTRUCTURE ONLY).

Float64 numpy forward/backward of the network the GPU trains: features
(centre and log2 width of lambda, r1, r2, r3, normalised and rounded to
BF16), three dense layers with ReLU between them, softmax cross-entropy;
weights used as BF16 (round to nearest even), biases in full precision.
Pinned by a central finite-difference check of its own gradient.
"""
import numpy as np

K_IN, K_H, K_OUT = 8, 64, 11
SHAPES = [("W1", (K_IN, K_H)), ("b1", (K_H,)), ("W2", (K_H, K_H)), ("b2", (K_H,)), ("W3", (K_H, K_OUT)),
          ("b3", (K_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    r = np.where(special, u.astype(np.uint32), r)
    return r.view(np.float32)


def unpack(params):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(params[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def features(soa, idx, norm):
    """(len(idx), 8) float32 features as the device computes them."""
    soa = np.asarray(soa)
    f = np.empty((len(idx), K_IN), dtype=np.float32)
    for d in range(4):
        lo, hi = soa[2 * d, idx], soa[2 * d + 1, idx]
        w = hi - lo
        f[:, 2 * d] = (0.5 * (lo + hi)).astype(np.float32)
        f[:, 2 * d + 1] = np.where(w > 2.0 ** -60, np.log2(np.maximum(w, 2.0 ** -60)), -60.0).astype(np.float32)
    norm = np.asarray(norm, dtype=np.float32)
    return ((f - norm[:8]) / norm[8:16]).astype(np.float32)


def loss_and_grad(params, x, y, quantize=True):
    """Mean softmax cross-entropy over the rows of x (labels y) and its
    gradient with respect to the packed parameters (the BF16-rounded weights
    are the ones differentiated, as on the device)."""
    p = unpack(bf16(params) if quantize else np.asarray(params, dtype=np.float64))
    if quantize:  # biases are used unrounded
        raw = unpack(params)
        for b in ("b1", "b2", "b3"):
            p[b] = raw[b]
    x = bf16(x).astype(np.float64) if quantize else np.asarray(x, dtype=np.float64)
    B = x.shape[0]
    z1 = x @ p["W1"] + p["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ p["W2"] + p["b2"]
    h2 = np.maximum(z2, 0)
    z3 = h2 @ p["W3"] + p["b3"]
    z = z3 - z3.max(axis=1, keepdims=True)
    e = np.exp(z)
    s = e.sum(axis=1, keepdims=True)
    loss = np.log(s[:, 0]) - z[np.arange(B), y]
    d3 = (e / s) / B
    d3[np.arange(B), y] -= 1.0 / B
    d2 = (d3 @ p["W3"].T) * (z2 > 0)
    d1 = (d2 @ p["W2"].T) * (z1 > 0)
    g = {"W1": x.T @ d1, "b1": d1.sum(0), "W2": h1.T @ d2, "b2": d2.sum(0), "W3": h2.T @ d3, "b3": d3.sum(0)}
    return loss.mean(), np.concatenate([g[n].ravel() for n, _ in SHAPES]), loss


def logits(params, x):
    """Inference forward as the MFMA kernel does it: BF16 inputs, weights and
    hidden activations, FP32-class accumulation (float64 here)."""
    p = unpack(bf16(params))
    raw = unpack(params)
    h = bf16(x).astype(np.float64)
    h = bf16(np.maximum(h @ p["W1"] + raw["b1"], 0)).astype(np.float64)
    h = bf16(np.maximum(h @ p["W2"] + raw["b2"], 0)).astype(np.float64)
    return h @ p["W3"] + raw["b3"]
