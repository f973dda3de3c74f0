"""The step after the path on the GPU: tone mapping equals the oracle's
bytes (double-evaluated pow), progressive accumulation of two trace calls
equals the oracle's accumulation of the same two images and the single
render of all their samples to FP32 rounding."""
import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes
from test_post import _tonemap, SPO_TX_CR
from _libs import oracle

pytestmark = pytest.mark.gpu


def _render(device, s, cam, w, h, spp, offset, img):
    ctx = V.make_context(cam, s, spp=spp, bounces=3, clamp=0.0,
                         estimator=V.EST_UNBIASED, rng=V.RNG_PCG, seed=9,
                         sample_offset=offset)
    V.path_trace_tiles_device(ctx, V.ComputeTiles(w, h, 64, 64), img)
    return img.download(np.float32, w * h * 4)


def test_tonemap_matches_oracle(device):
    desc = scenes.config1(64, 48)
    s, cam = scenes.load_gpu(device, desc)
    img = device.alloc(64 * 48 * 16)
    px = _render(device, s, cam, 64, 48, 4, 0, img)
    for exposure in (1.0, 0.3):
        got = device.tonemap(img, 64 * 48, exposure, V.TX_DOUBLE)
        assert np.array_equal(got, _tonemap(px, exposure, SPO_TX_CR))
        fast = device.tonemap(img, 64 * 48, exposure, V.TX_FAST)
        assert np.max(np.abs(fast.astype(int) - got.astype(int))) <= 1
    img.free()
    s.close()


def test_progressive_accumulation(device):
    desc = scenes.config1(64, 48)
    s, cam = scenes.load_gpu(device, desc)
    n = 64 * 48
    acc, part, full = device.alloc(n * 16), device.alloc(n * 16), device.alloc(n * 16)
    a = _render(device, s, cam, 64, 48, 2, 0, acc)
    b = _render(device, s, cam, 64, 48, 2, 2, part)
    device.accumulate(acc, part, n, 2, 2)
    got = acc.download(np.float32, n * 4)
    want = a.copy()
    oracle().spo_accumulate(want, np.ascontiguousarray(b), n, 2, 2)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    whole = _render(device, s, cam, 64, 48, 4, 0, full)
    assert np.allclose(got, whole, rtol=1e-5, atol=1e-7)
    for buf in (acc, part, full):
        buf.free()
    s.close()
