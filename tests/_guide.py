"""Shared helpers of the guide tests: parameter sets, sample batches, the
oracle's loss/gradient in FP64, and an FP16-emulating forward pass of the
network in numpy (the GPU's rounding points: FP16 encodings, weights and
hidden activations, FP32 accumulation and biases)."""
import math

import numpy as np

from _libs import GUIDE_PARAMS, oracle

W_OFF = [l * (64 * 64 + 64) for l in range(5)]
ROWS = [64, 64, 64, 64, 48]


def random_params(seed, scale_out=1.0, fp16_weights=True):
    rng = np.random.default_rng(seed)
    p = np.zeros(GUIDE_PARAMS, np.float32)
    for l in range(5):
        w = rng.uniform(-1, 1, (ROWS[l], 64)) * math.sqrt(6.0 / 64)
        if l == 4:
            w *= scale_out
        if fp16_weights:
            w = w.astype(np.float16).astype(np.float64)
        o = W_OFF[l]
        p[o:o + ROWS[l] * 64] = w.ravel()
        p[o + ROWS[l] * 64:o + ROWS[l] * 65] = rng.normal(0, 0.1, ROWS[l])
    # lobe means away from zero length
    b4 = W_OFF[4] + 48 * 64
    p[b4 + 16:b4 + 40] += rng.normal(0, 1.0, 24).astype(np.float32)
    return p


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def random_inputs(seed, n):
    """Rows pn[3] n[3] v[3] (float32)."""
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 9), np.float32)
    x[:, 0:3] = rng.uniform(size=(n, 3))
    x[:, 3:6] = unit(rng.normal(size=(n, 3)))
    x[:, 6:9] = unit(rng.normal(size=(n, 3)))
    return x


def random_samples(seed, n):
    """Rows pn n v wi weight (float32)."""
    rng = np.random.default_rng(seed + 1000)
    s = np.zeros((n, 13), np.float32)
    s[:, 0:9] = random_inputs(seed, n)
    s[:, 9:12] = unit(rng.normal(size=(n, 3)))
    s[:, 12] = rng.exponential(size=n)
    return s


def encode_f64(rows9):
    O = oracle()
    enc = np.zeros((len(rows9), 64))
    for j, r in enumerate(rows9):
        O.spo_guide_encode(np.ascontiguousarray(r[0:3]), np.ascontiguousarray(r[3:6]),
                           np.ascontiguousarray(r[6:9]), enc[j])
    return enc


def oracle_forward(params, rows9):
    O = oracle()
    p = params.astype(np.float64)
    enc = encode_f64(rows9)
    out = np.zeros((len(rows9), 48))
    for j in range(len(rows9)):
        O.spo_guide_forward(p, enc[j], out[j], None)
    return out


def oracle_loss_grad(params, samples13, weights=None):
    O = oracle()
    p = params.astype(np.float64)
    enc = encode_f64(samples13[:, 0:9])
    wi = np.ascontiguousarray(samples13[:, 9:12].astype(np.float64).ravel())
    w = np.ascontiguousarray((samples13[:, 12] if weights is None else weights)
                             .astype(np.float64))
    g = np.zeros(GUIDE_PARAMS)
    loss = O.spo_guide_loss(p, enc, wi, w, len(samples13), g.ctypes.data)
    return loss, g


def fp16_forward(params, rows9):
    """The GPU's rounding points in numpy: encodings, weights and hidden
    activations in FP16, FP32 products and biases."""
    enc = encode_f64(rows9).astype(np.float16).astype(np.float32)
    h = enc.T  # features x rays
    for l in range(5):
        o = W_OFF[l]
        W = params[o:o + ROWS[l] * 64].reshape(ROWS[l], 64)
        W = W.astype(np.float16).astype(np.float32)
        b = params[o + ROWS[l] * 64:o + ROWS[l] * 65].astype(np.float32)
        z = W @ h + b[:, None]
        if l < 4:
            h = np.maximum(z, 0).astype(np.float16).astype(np.float32)
        else:
            return z.T
