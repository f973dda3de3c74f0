"""The FP64 guide oracle (oracle/spo_guide.c) pinned by what the mathematics
fixes: the density integrates to one over the sphere, samples follow the
density, kappa -> 0 is uniform, the loss gradient equals finite
differences, Adam's first step is lr * sign(g), and the encoding and the
parameter layout give known values."""
import ctypes as C
import math

import numpy as np
import pytest

from _libs import GUIDE_PARAMS, SpoLobes, oracle


def fib_sphere(n):
    """n nearly equal-area points on the unit sphere (Fibonacci lattice)."""
    k = np.arange(n) + 0.5
    z = 1.0 - 2.0 * k / n
    r = np.sqrt(1.0 - z * z)
    phi = np.pi * (1.0 + 5 ** 0.5) * k
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def random_params(seed, scale_out=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(GUIDE_PARAMS)
    off = 0
    for layer in range(5):
        rows = 64 if layer < 4 else 48
        w = rng.uniform(-1, 1, (rows, 64)) * math.sqrt(6.0 / 64)
        if layer == 4:
            w *= scale_out
        p[off:off + rows * 64] = w.ravel()
        off += rows * 64
        p[off:off + rows] = rng.normal(0, 0.1, rows)
        off += rows
    assert off == GUIDE_PARAMS
    return p


def encode(pn, n, v):
    e = np.zeros(64)
    oracle().spo_guide_encode(np.float32(pn), np.float32(n), np.float32(v), e)
    return e


def lobes_of(out):
    L = SpoLobes()
    oracle().spo_guide_lobes(np.ascontiguousarray(out, np.float64), C.byref(L))
    return L


def test_encoding_known_values():
    e = encode([0.25, 0.5, 0.75], [0, 1, 0], [0.6, 0.0, 0.8])
    h = math.sqrt(0.5)
    assert e[0:6] == pytest.approx([h, h, 1, 0, h, -h], abs=1e-15)
    # 2^1 pi * 0.25 = pi/2
    assert e[6:8] == pytest.approx([1, 0], abs=1e-15)
    assert list(e[48:58]) == pytest.approx([0.25, 0.5, 0.75, 0, 1, 0, 0.6, 0, 0.8, 1],
                                           abs=1e-7)
    assert np.all(e[58:] == 0)


def test_forward_layout_against_matrix_products():
    """W[out][in] row-major, biases after each matrix, ReLU on the four
    hidden layers only."""
    p = random_params(1)
    e = encode([0.1, 0.7, 0.3], [0, 0, 1], [1, 0, 0])
    out = np.zeros(48)
    acts = np.zeros(256)
    oracle().spo_guide_forward(p, e, out, acts.ctypes.data)
    h, off = e, 0
    for layer in range(5):
        rows = 64 if layer < 4 else 48
        W = p[off:off + rows * 64].reshape(rows, 64)
        b = p[off + rows * 64:off + rows * 64 + rows]
        off += rows * 64 + rows
        h = W @ h + b
        if layer < 4:
            h = np.maximum(h, 0)
            assert acts[64 * layer:64 * layer + 64] == pytest.approx(h, rel=1e-12)
    assert out == pytest.approx(h, rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("kappa", [0.14, 1.0, 10.0, 100.0, 400.0])
def test_vmf_density_integrates_to_one(kappa):
    O = oracle()
    mu = np.array([0.36, -0.48, 0.8])
    pts = fib_sphere(60000 if kappa < 50 else 240000)
    s = sum(O.spo_vmf_pdf(kappa, mu, np.ascontiguousarray(w)) for w in pts)
    assert s * 4 * math.pi / len(pts) == pytest.approx(1.0, abs=2e-3)


def test_vmf_small_kappa_is_uniform():
    O = oracle()
    mu = np.array([0.0, 0.0, 1.0])
    for w in ([0, 0, 1], [1, 0, 0], [0, 0, -1]):
        assert O.spo_vmf_pdf(1e-7, mu, np.float64(w)) == pytest.approx(
            1 / (4 * math.pi), rel=1e-6)


def _moments_by_sampling(sample, n=40000):
    """Stratified samples -> means of a few test functions."""
    m = int(math.sqrt(n))
    acc = np.zeros(6)
    w = np.zeros(3)
    for i in range(m):
        for j in range(m):
            sample((i + 0.5) / m, (j + 0.37) / m, w)
            acc += [w[0], w[1], w[2], w[0] * w[0], w[1] * w[2], float(w[2] > 0.5)]
            assert abs(np.dot(w, w) - 1) < 1e-12
    return acc / (m * m)


def _moments_by_quadrature(pdf, n=200000):
    pts = fib_sphere(n)
    q = np.array([pdf(np.ascontiguousarray(w)) for w in pts]) * 4 * math.pi / n
    f = np.stack([pts[:, 0], pts[:, 1], pts[:, 2], pts[:, 0] ** 2,
                  pts[:, 1] * pts[:, 2], (pts[:, 2] > 0.5).astype(float)], 1)
    return q @ f


@pytest.mark.parametrize("kappa", [0.5, 5.0, 60.0])
def test_vmf_samples_follow_density(kappa):
    O = oracle()
    mu = np.array([0.6, 0.0, 0.8])
    a = _moments_by_sampling(lambda u1, u2, w: O.spo_vmf_sample(kappa, mu, u1, u2, w))
    b = _moments_by_quadrature(lambda w: O.spo_vmf_pdf(kappa, mu, w))
    assert a == pytest.approx(b, abs=4e-3)


def test_mixture_samples_follow_density():
    O = oracle()
    p = random_params(3, scale_out=3.0)
    out = np.zeros(48)
    O.spo_guide_forward(p, encode([0.3, 0.2, 0.9], [0, 1, 0], [0, 0.6, 0.8]), out,
                        None)
    L = lobes_of(out)
    assert sum(L.pi) == pytest.approx(1.0, abs=1e-12)
    assert all(math.exp(-2) < k < math.exp(6) for k in L.kappa)
    rng = np.random.default_rng(0)

    def sample(u1, u2, w):
        O.spo_guide_sample(C.byref(L), rng.uniform(), u1, u2, w)

    a = _moments_by_sampling(sample, n=90000)
    b = _moments_by_quadrature(lambda w: O.spo_guide_pdf(C.byref(L), w))
    assert a == pytest.approx(b, abs=1e-2)


def _batch(seed, n):
    rng = np.random.default_rng(seed)
    enc = np.zeros((n, 64))
    wi = np.zeros((n, 3))
    for j in range(n):
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        enc[j] = encode(rng.uniform(size=3), nrm, v)
        d = rng.normal(size=3)
        wi[j] = d / np.linalg.norm(d)
    w = rng.exponential(size=n)
    return enc, wi, w


def test_loss_gradient_matches_finite_differences():
    O = oracle()
    p = random_params(5, scale_out=2.0)
    enc, wi, w = _batch(6, 6)
    g = np.zeros(GUIDE_PARAMS)
    O.spo_guide_loss(p, enc, wi.ravel().copy(), w, 6, g.ctypes.data)
    rng = np.random.default_rng(7)
    # every output-layer parameter and a sample of the hidden ones
    out_start = 4 * (64 * 64 + 64)
    idx = np.concatenate([np.arange(out_start, GUIDE_PARAMS),
                          rng.choice(out_start, 300, replace=False)])
    h = 1e-6
    for k in idx:
        q = p.copy()
        q[k] += h
        lp = O.spo_guide_loss(q, enc, wi.ravel().copy(), w, 6, None)
        q[k] -= 2 * h
        lm = O.spo_guide_loss(q, enc, wi.ravel().copy(), w, 6, None)
        fd = (lp - lm) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-4, abs=1e-7), k


def test_loss_decreases_along_negative_gradient():
    O = oracle()
    p = random_params(8)
    enc, wi, w = _batch(9, 16)
    g = np.zeros(GUIDE_PARAMS)
    l0 = O.spo_guide_loss(p, enc, wi.ravel().copy(), w, 16, g.ctypes.data)
    l1 = O.spo_guide_loss(p - 1e-3 * g / np.abs(g).max(), enc, wi.ravel().copy(), w,
                          16, None)
    assert l1 < l0


def test_adam_first_step_is_lr_times_sign():
    O = oracle()
    p = np.float32([1.0, -2.0, 0.5, 0.0])
    m = np.zeros(4, np.float32)
    v = np.zeros(4, np.float32)
    g = np.array([0.3, -4.0, 1e-3, 0.0])
    O.spo_guide_adam(p, m, v, g, 4, 0.01, 0.9, 0.999, 1e-12, 1)
    assert p == pytest.approx([0.99, -1.99, 0.49, 0.0], abs=1e-6)
