"""Irradiance cube map on the GPU (spg_create_irradiance_cubemap,
CreateIrradianceCubeMap src/cubemap.cpp:108-233) against the oracle: bit
for bit with double-evaluated transcendentals (one wave per texel, the
reference's sequential sum order kept), within 1e-3 with fast ones."""
import numpy as np
import pytest

import vk_cinematic_amd as V
from test_oracle_irradiance import _irradiance, SPO_TX_CR

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("fw,fh,clamp", [(32, 32, 10.0), (5, 3, 10.0), (8, 8, 0.0)])
def test_sky_matches_oracle_bit_exact(device, fw, fh, clamp):
    sky = V.gen_sky_equirect(128, 64).reshape(64, 128, 4)
    got, ms = device.create_irradiance_cubemap(sky, fw, fh, clamp, V.TX_DOUBLE)
    want = _irradiance(sky, fw, fh, clamp, SPO_TX_CR)
    assert np.array_equal(_bits(got), _bits(want))
    assert ms > 0


def test_random_hdr_matches_oracle_bit_exact(device):
    rng = np.random.default_rng(23)
    img = rng.exponential(3.0, (33, 70, 4)).astype(np.float32)
    img[rng.integers(0, 33, 20), rng.integers(0, 70, 20), :3] = 500.0
    got, _ = device.create_irradiance_cubemap(img, 6, 6, 10.0, V.TX_DOUBLE)
    want = _irradiance(img, 6, 6, 10.0, SPO_TX_CR)
    assert np.array_equal(_bits(got), _bits(want))


def test_fast_transcendentals_close(device):
    sky = V.gen_sky_equirect(128, 64).reshape(64, 128, 4)
    got, _ = device.create_irradiance_cubemap(sky, 16, 16, 10.0, V.TX_FAST)
    want = _irradiance(sky, 16, 16, 10.0, SPO_TX_CR)
    assert np.allclose(got, want, rtol=1e-3, atol=1e-4)


def test_rejects_bad_arguments(device):
    img = np.ones((4, 8, 4), np.float32)
    with pytest.raises(V.SpgError):
        device.create_irradiance_cubemap(img, 0, 4)
    with pytest.raises(V.SpgError):
        device.create_irradiance_cubemap(img, 4, 4, 10.0, 7)
