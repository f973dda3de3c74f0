"""The double-evaluated sin/cos of the GPU's bit-reproducible mode
(SPG_TX_DOUBLE) against the C library's double sin and cos rounded once to
float (the oracle's SPO_TX_CR): exhaustively on the argument sets the
samplers produce, and on random arguments (CPU; the product's host build of
the same code)."""
import numpy as np
import pytest

import vk_cinematic_amd as V
from _libs import oracle


def _both(x):
    x = np.ascontiguousarray(x, np.float32)
    n = x.size
    s1, c1 = np.zeros(n, np.float32), np.zeros(n, np.float32)
    s2, c2 = np.zeros(n, np.float32), np.zeros(n, np.float32)
    V.check(V.lib().spg_sincos_host(x, n, s1, c1))
    oracle().spo_sincos_cr_array(x, n, s2, c2)
    return s1, c1, s2, c2


@pytest.mark.parametrize("scale", [6.28318548, 3.14159265359, -3.14159265359])
def test_sincos_exhaustive_on_sampler_arguments(scale):
    """Every phi = 2 pi u (cosine and GGX lobes) and every +-pi u for the
    2^24 values of a 24-bit uniform u: identical bits."""
    u = np.arange(1 << 24, dtype=np.float32) * np.float32(1.0 / 16777216.0)
    s1, c1, s2, c2 = _both(np.float32(scale) * u)
    assert np.array_equal(s1.view(np.uint32), s2.view(np.uint32))
    assert np.array_equal(c1.view(np.uint32), c2.view(np.uint32))


def test_sincos_random_and_special_arguments():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-2000, 2000, 4_000_000),
                        rng.uniform(-1e7, 1e7, 100_000),     # library path
                        [0.0, -0.0, 1e-30, -1e-30, 1647098.0, 1647100.0,
                         np.inf, -np.inf, np.nan]]).astype(np.float32)
    s1, c1, s2, c2 = _both(x)
    for a, b in ((s1, s2), (c1, c2)):
        same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
        assert same.all(), x[~same][:5]
