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


# ---- the FP32 restatement of the product's guided scatter and estimator
def _lobes_f32(rng):
    """A random mixture as the product stores it: pi[8] kappa[8] mu[8][3]."""
    pi = rng.uniform(0.05, 1.0, 8)
    pi /= pi.sum()
    kap = np.exp(rng.uniform(np.log(0.5), np.log(300.0), 8))
    mu = rng.normal(size=(8, 3))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    return np.concatenate([pi, kap, mu.ravel()]).astype(np.float32)


def _as_spo(L40):
    L = SpoLobes()
    for k in range(8):
        L.pi[k] = float(L40[k])
        L.kappa[k] = float(L40[8 + k])
        for a in range(3):
            L.mu[k][a] = float(L40[16 + 3 * k + a])
    return L


def _unit(v):
    return (v / np.linalg.norm(v)).astype(np.float32)


def test_fp32_guided_scatter_against_fp64_mixture():
    """spo_guided_scatter (the product's FP32 sampler, one-sample MIS of the
    BSDF lobe and the mixture) against the independent FP64 guide oracle
    (pinned by quadrature and sample histograms): with fb = 0 the direction
    equals spo_guide_sample's for the same uniforms and the density is
    spo_guide_pdf's; with fb = 1 on a rough-0 material the density is
    cos/pi; with fb = 0.5 it is the average of the two."""
    from _libs import SPO_TX_CR

    O = oracle()
    rng = np.random.default_rng(41)
    checked = 0
    for trial in range(400):
        L40 = _lobes_f32(rng)
        n = _unit(rng.normal(size=3))
        v = _unit(rng.normal(size=3))
        if n @ v < 0:
            v = -v
        u0, u1, u2 = rng.uniform(size=3).astype(np.float32)
        cum = np.cumsum(L40[:8].astype(np.float64))
        if np.min(np.abs(cum - u0)) < 1e-5:
            continue  # FP32 and FP64 cumulative sums may pick different lobes
        wi = np.zeros(3, np.float32)
        pdf = C.c_float()
        O.spo_guided_scatter(L40, n, v, 0.0, 0.0, u0, u1, u2, SPO_TX_CR, wi,
                             C.byref(pdf))
        L = _as_spo(L40)
        w64 = np.zeros(3)
        O.spo_guide_sample(L, float(u0), float(u1), float(u2), w64)
        assert np.max(np.abs(wi - w64)) < 2e-5, trial
        q = O.spo_guide_pdf(L, w64)
        if float(n @ wi) > 1e-3:
            assert pdf.value == pytest.approx(q, rel=2e-4, abs=1e-6), trial
            # fb = 0.5: average of cos/pi and q at the same direction
            O.spo_guided_scatter(L40, n, v, 0.0, 0.5, 0.5 + 0.5 * u0, u1, u2,
                                 SPO_TX_CR, wi, C.byref(pdf))
            cos = float(n.astype(np.float64) @ wi)
            want = 0.5 * cos / math.pi + 0.5 * O.spo_guide_pdf(L, wi.astype(np.float64))
            assert pdf.value == pytest.approx(want, rel=2e-4), trial
            checked += 1
        else:
            assert pdf.value == 0.0
        # fb = 1: the cosine lobe, density cos/pi, direction above the surface
        O.spo_guided_scatter(L40, n, v, 0.0, 1.0, u0, u1, u2, SPO_TX_CR, wi,
                             C.byref(pdf))
        cos = float(n.astype(np.float64) @ wi)
        assert cos > -1e-6 and abs(np.linalg.norm(wi) - 1) < 1e-5
        if cos > 1e-6:
            assert pdf.value == pytest.approx(cos / math.pi, rel=1e-5)
    assert checked > 150


def test_analytic_mixture_definition():
    """The verification mixture: weights 0.5/0.3/0.2, unit means toward the
    point (0.3, 2.5, -0.4), along the normal and along the mirror
    direction; concentrations 40, 3, 120."""
    O = oracle()
    rng = np.random.default_rng(5)
    for _ in range(50):
        P = rng.uniform(-2, 2, 3).astype(np.float32)
        n = _unit(rng.normal(size=3))
        v = _unit(rng.normal(size=3))
        L = np.zeros(40, np.float32)
        O.spo_guide_analytic_lobes(P, n, v, L)
        assert L[:8].sum() == pytest.approx(1.0) and np.all(L[3:8] == 0)
        assert list(L[8:11]) == [40.0, 3.0, 120.0]
        mu = L[16:40].reshape(8, 3).astype(np.float64)
        t = np.array([0.3, 2.5, -0.4]) - P
        assert mu[0] @ (t / np.linalg.norm(t)) == pytest.approx(1.0, abs=1e-6)
        assert np.allclose(mu[1], n)
        r = 2 * (n @ v) * n - v
        assert mu[2] @ (r / np.linalg.norm(r)) == pytest.approx(1.0, abs=1e-6)


def _guided_cfg(spp, bounces, seed, est, fb=0.5):
    from _libs import SPO_HIT_BRUTE, SPO_RNG_PCG, SPO_TX_CR, SpoConfig

    return SpoConfig(rng=SPO_RNG_PCG, tx=SPO_TX_CR, hit=SPO_HIT_BRUTE, estimator=est,
                     samplesPerPixel=spp, bounces=bounces, radianceClamp=0.0,
                     jitterYFirst=1, seed=seed, bsdfFraction=fb)


def test_oracle_guided_estimator_closed_form_and_unbiased():
    """The oracle's guided estimator (verification mixture, one-sample MIS)
    is unbiased: inside a closed sphere of albedo a and emission E its mean
    reaches E (1 - q^N) / (1 - q) up to the Fresnel term, as the plain
    estimator's does; on the config-1 scene its mean equals the plain
    estimator's within the Monte Carlo error; every sink weight is positive
    and finite."""
    from _libs import SPO_EST_GUIDED, SPO_EST_UNBIASED
    from _scenes import load_oracle
    import vk_cinematic_amd as V
    from vk_cinematic_amd import scenes

    O = oracle()
    v, i = V.gen_icosphere(3)
    vin = v.copy()
    vin[:, 3:6] *= -1
    iin = i.reshape(-1, 3)[:, ::-1].ravel().copy()
    a, E, N = 0.5, 1.0, 5
    desc = {"meshes": [(vin, iin, True)],
            "objects": [(0, 1, (0, 0, 0), (0, 0, 0, 1), (5, 5, 5))],
            "materials": [(1, (a, a, a), (E, E, E), 0.0, V.SPG_NO_TEXTURE,
                           V.SPG_NO_TEXTURE)],
            "background": 9, "camera": (16, 16, (0, 0, 0), (0, 0, 0, 1), 0.8)}
    s, cam = load_oracle(desc)
    px = np.zeros(16 * 16 * 4, np.float32)
    assert O.spo_path_trace_tile(s, C.byref(cam), C.byref(_guided_cfg(64, N, 3, SPO_EST_GUIDED)),
                                 0, 0, 16, 16, None, px, np.zeros(12, np.uint64)) == 0
    q = a * (1 - 0.04)
    hi = E * (1 - q ** N) / (1 - q)
    mean = px.reshape(-1, 4)[:, 0].mean()
    assert E < mean < hi and mean == pytest.approx(hi, rel=0.05)
    O.spo_scene_destroy(s)

    desc = scenes.config1(24, 24, env="cube", env_size=16)
    s, cam = load_oracle(desc)
    sink = np.zeros((200000, 16), np.float32)
    O.spo_guided_sink(sink.ctypes.data, len(sink))
    means = {}
    for est, seeds in ((SPO_EST_GUIDED, (11, 12, 13)), (SPO_EST_UNBIASED, (21, 22, 23))):
        ms = []
        for sd in seeds:
            px = np.zeros(24 * 24 * 4, np.float32)
            assert O.spo_path_trace_tile(s, C.byref(cam),
                                         C.byref(_guided_cfg(32, 4, sd, est)), 0, 0, 24,
                                         24, None, px, np.zeros(12, np.uint64)) == 0
            ms.append(px.reshape(-1, 4)[:, :3].mean())
        means[est] = np.array(ms)
    n = O.spo_guided_sink_count()
    O.spo_guided_sink(None, 0)
    O.spo_scene_destroy(s)
    g, p = means[SPO_EST_GUIDED], means[SPO_EST_UNBIASED]
    sd = np.sqrt(g.var(ddof=1) / 3 + p.var(ddof=1) / 3) + 1e-3 * p.mean()
    assert abs(g.mean() - p.mean()) < 4 * sd
    assert n > 1000
    w = sink[:n, 12]
    assert np.all(w > 0) and np.all(np.isfinite(w))
