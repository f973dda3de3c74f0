"""The oracle's irradiance cube map (CreateIrradianceCubeMap,
src/cubemap.cpp:108-233, production switches: uniform sampling, radiance
clamp 10) pinned against the reference build and closed forms."""
import numpy as np
import pytest

from _libs import oracle, ref

SPO_TX_LIBM, SPO_TX_CR = 0, 1


def _grid():
    """The reference's loop counters: phi += 0.1f below 2 pi, theta += 0.1f
    below pi/2 (float arithmetic)."""
    pi2 = np.float32(2.0) * np.float32(3.14159265359)
    half = np.float32(0.5) * np.float32(3.14159265359)
    phis, thetas = [], []
    p = np.float32(0)
    while p < pi2:
        phis.append(p)
        p = np.float32(p + np.float32(0.1))
    t = np.float32(0)
    while t < half:
        thetas.append(t)
        t = np.float32(t + np.float32(0.1))
    return np.array(phis, np.float32), np.array(thetas, np.float32)


def _irradiance(img, fw, fh, clamp=10.0, tx=SPO_TX_LIBM):
    h, w = img.shape[:2]
    out = np.zeros(6 * fw * fh * 4, np.float32)
    oracle().spo_create_irradiance_cubemap(np.ascontiguousarray(img, np.float32).ravel(),
                                           w, h, fw, fh, clamp, tx, out)
    return out.reshape(6, fh, fw, 4)


def test_grid_size():
    phis, thetas = _grid()
    assert len(phis) == 63 and len(thetas) == 16


@pytest.mark.parametrize("L,clamp,expect", [(2.0, 10.0, 2.0), (50.0, 10.0, 10.0),
                                            (50.0, 0.0, 50.0)])
def test_constant_environment_closed_form(L, clamp, expect):
    """Constant radiance L (clamped to `clamp`): every texel holds
    pi * L' * mean(cos(theta) sin(theta)) over the sample grid, whatever its
    direction, and alpha 1."""
    phis, thetas = _grid()
    s = float(np.mean(np.cos(thetas.astype(np.float64)) * np.sin(thetas.astype(np.float64))))
    img = np.full((8, 16, 4), L, np.float32)
    for tx in (SPO_TX_LIBM, SPO_TX_CR):
        cube = _irradiance(img, 3, 3, clamp, tx)
        want = np.pi * expect * s
        assert np.allclose(cube[..., :3], want, rtol=2e-5)
        assert np.all(cube[..., 3] == 1.0)


def test_upper_hemisphere_light():
    """Radiance 1 only in the image's upper half (directions with y > 0):
    the +Y face's centre texel (normal straight up) gets the full
    constant-environment value, the -Y face's centre none, and a texel
    whose normal is horizontal (side faces, middle row) half of it."""
    phis, thetas = _grid()
    full = np.pi * float(np.mean(np.cos(thetas.astype(np.float64)) *
                                 np.sin(thetas.astype(np.float64))))
    img = np.zeros((16, 32, 4), np.float32)
    img[:8] = 1.0
    cube = _irradiance(img, 4, 4)
    assert cube[2, 2, 2, 0] == pytest.approx(full, rel=2e-2)
    assert np.all(cube[3, 2, 2, :3] == 0.0)
    for face in (0, 1, 4, 5):
        assert np.allclose(cube[face, 2, :, :3], full / 2, rtol=3e-2)
    assert np.all(cube[2, ..., 0] > cube[3, ..., 0])


def test_oracle_equals_reference_build():
    """libm transcendentals: bit for bit the reference's own function on an
    HDR image with values above the clamp."""
    R = ref()
    if R is None or not hasattr(R, "ref_create_irradiance_cubemap"):
        pytest.skip("reference build not available")
    rng = np.random.default_rng(17)
    img = rng.exponential(2.0, (12, 24, 4)).astype(np.float32)
    img[3, 5, :3] = 80.0
    fw, fh = 3, 2
    want = np.zeros(6 * fw * fh * 4, np.float32)
    R.ref_create_irradiance_cubemap(np.ascontiguousarray(img.ravel()), 24, 12, fw, fh, want)
    got = _irradiance(img, fw, fh, 10.0, SPO_TX_LIBM)
    assert np.array_equal(got.ravel().view(np.uint32), want.view(np.uint32))
