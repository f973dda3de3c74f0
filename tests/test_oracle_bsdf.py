"""BSDF sampling of the unbiased estimator (GGX half-vector sampling,
Walter et al. 2007) pinned by what the mathematics fixes: samples follow the
stated density (moments against quadrature of the density), the density of
directions above the surface never exceeds one in total, and it concentrates
around the mirror direction as the roughness falls."""
import math

import numpy as np
import pytest

from _libs import f32p, oracle


def _sig(O):
    O.spo_ggx_direction.restype = None
    O.spo_ggx_direction.argtypes = [f32p, f32p, np.ctypeslib.ctypes.c_float,
                                    np.ctypeslib.ctypes.c_float,
                                    np.ctypeslib.ctypes.c_float,
                                    np.ctypeslib.ctypes.c_int, f32p]
    O.spo_ggx_pdf.restype = np.ctypeslib.ctypes.c_float
    O.spo_ggx_pdf.argtypes = [f32p, f32p, f32p, np.ctypeslib.ctypes.c_float]
    return O


def fib_sphere(n):
    k = np.arange(n) + 0.5
    z = 1.0 - 2.0 * k / n
    r = np.sqrt(1.0 - z * z)
    phi = np.pi * (1.0 + 5 ** 0.5) * k
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1).astype(np.float32)


N = np.float32([0.0, 0.0, 1.0])


def _view(theta):
    return np.float32([math.sin(theta), 0.0, math.cos(theta)])


@pytest.mark.parametrize("rough,theta", [(0.3, 0.3), (0.6, 0.9), (0.8, 0.2)])
def test_ggx_samples_follow_density(rough, theta):
    O = _sig(oracle())
    V = _view(theta)
    m = 240
    acc = np.zeros(5)
    out = np.zeros(3, np.float32)
    for i in range(m):
        for j in range(m):
            O.spo_ggx_direction(N, V, rough, (i + 0.5) / m, (j + 0.37) / m, 1, out)
            h = out + V
            h = h / np.linalg.norm(h)
            if np.dot(h, V) <= 0.0 or h[2] <= 0.0:
                # the sampled microfacet faced away from V: L lies below
                # the surface and the estimator drops it (density 0)
                continue
            acc += [out[0], out[1], out[2], out[0] * out[2], float(out[2] > 0.5)]
    a = acc / (m * m)
    pts = fib_sphere(400000)
    q = np.array([O.spo_ggx_pdf(N, V, np.ascontiguousarray(p), rough) for p in pts])
    q *= 4 * math.pi / len(pts)
    f = np.stack([pts[:, 0], pts[:, 1], pts[:, 2], pts[:, 0] * pts[:, 2],
                  (pts[:, 2] > 0.5).astype(float)], 1)
    assert a == pytest.approx(q @ f, abs=6e-3)
    assert q.sum() <= 1.0 + 5e-3


def test_ggx_concentrates_at_the_mirror_direction():
    O = _sig(oracle())
    V = _view(0.5)
    mirror = np.float32([-V[0], -V[1], V[2]])
    out = np.zeros(3, np.float32)
    spread = []
    for rough in (0.8, 0.5, 0.3, 0.1):
        d = []
        for i in range(64):
            for j in range(16):
                O.spo_ggx_direction(N, V, rough, (i + 0.5) / 64, (j + 0.5) / 16, 1, out)
                d.append(float(out @ mirror))
        spread.append(1 - np.median(d))
        assert np.allclose(np.linalg.norm(out), 1.0, atol=1e-5)
    assert spread == sorted(spread, reverse=True)
    assert spread[-1] < 1e-3
