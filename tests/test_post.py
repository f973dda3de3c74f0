"""The step after the path: the oracle's tone mapping (PerformToneMapping,
src/shaders/post_processing.frag.glsl:19-26) and progressive accumulation
against closed forms, and the PFM writer (a host function of the library)."""
import numpy as np
import pytest

import vk_cinematic_amd as V
from _libs import oracle

SPO_TX_LIBM, SPO_TX_CR = 0, 1


def _tonemap(rgba, exposure=1.0, tx=SPO_TX_CR):
    rgba = np.ascontiguousarray(rgba, np.float32).ravel()
    out = np.zeros(rgba.size, np.uint8)
    oracle().spo_tonemap(rgba, rgba.size // 4, exposure, tx, out)
    return out.reshape(-1, 4)


@pytest.mark.parametrize("tx", [SPO_TX_LIBM, SPO_TX_CR])
def test_tonemap_closed_forms(tx):
    px = np.array([[0, 1, 1e30, 1], [0.5, np.nan, -0.5, 1], [3, 0.25, 9, 1]], np.float32)
    q = _tonemap(px, 1.0, tx)
    # 255 * (c / (1 + c))^(1/2.2), rounded
    want = np.round(255 * (np.array([0, .5, 1, 1 / 3, 0, 0, .75, .2, .9]) **
                           (1 / 2.2))).astype(np.uint8)
    want[4] = want[5] = 0  # NaN, and -0.5 (c / (1 + c) = -1, pow NaN)
    assert q[:, :3].ravel().tolist() == want.tolist()
    assert np.all(q[:, 3] == 255)
    # exposure 2 on c is c' = 2c
    assert np.array_equal(_tonemap(px * 2, 1.0, tx), _tonemap(px, 2.0, tx))


def test_accumulate_is_the_weighted_mean():
    rng = np.random.default_rng(3)
    a = rng.uniform(0, 4, (50, 4)).astype(np.float32)
    b = rng.uniform(0, 4, (50, 4)).astype(np.float32)
    d = a.copy()
    oracle().spo_accumulate(d.ravel(), np.ascontiguousarray(b.ravel()), 50, 3, 5)
    want = (a.astype(np.float64) * 3 + b * 5) / 8
    assert np.allclose(d[:, :3], want[:, :3], rtol=1e-6)
    assert np.all(d[:, 3] == 1.0)


def test_pfm_round_trip(tmp_path):
    img = np.arange(3 * 2 * 4, dtype=np.float32).reshape(2, 3, 4) / 7.0
    path = tmp_path / "x.pfm"
    V.write_pfm(path, 3, 2, img)
    data = path.read_bytes()
    head = b"PF\n3 2\n-1.0\n"
    assert data[:len(head)] == head
    body = np.frombuffer(data[len(head):], "<f4").reshape(2, 3, 3)
    assert np.array_equal(body[::-1], img[:, :, :3])
