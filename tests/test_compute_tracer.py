"""The oracle's restatement of the reference's parked Vulkan compute
tracer (src/shaders/test_compute.comp.glsl; oracle/spo_compute.c) pinned by
closed forms (CPU only): the checkerboard texels and bilinear midpoints of
CreateCheckerBoardImage (src/image.cpp:134-160), a light seen directly is
exactly 1, nothing in view is 0, a red surface lit by white light keeps
green equal to blue and both below the albedo bound, tiles write only their
own pixels and differ by their seeds."""
import numpy as np
import pytest

from _compute_scene import camera_transform, compute_scene
from _libs import oracle
from _scenes import load_oracle


def test_checkerboard_texels_and_bilinear():
    O = oracle()
    for tx, ty in [(0, 0), (15, 0), (16, 0), (16, 16), (37, 200), (255, 255)]:
        u, v = (tx + 0.5) / 256, (ty + 0.5) / 256
        want = 0.04 if ((tx // 16) + (ty // 16)) % 2 else 0.18
        assert O.spo_checker_bilinear(u, v) == pytest.approx(want, abs=1e-7)
        assert O.spo_checker_bilinear(u + 1.0, v - 1.0) == pytest.approx(want, abs=1e-7)
    # halfway between texels 15 and 16 (a square edge): the mean of the two
    assert O.spo_checker_bilinear(16 / 256, 8.5 / 256) == pytest.approx(0.11, abs=1e-7)
    # the edge at u = 0 wraps to texel 255 (another square colour)
    assert O.spo_checker_bilinear(0.0, 0.5 / 256) == pytest.approx(
        0.5 * (0.18 + 0.04), abs=1e-7)


def _render(desc, tiles):
    O = oracle()
    s, _ = load_oracle(desc, build_midphase=False)
    img = np.full(256 * 192 * 4, -1.0, np.float32)
    cam = camera_transform(desc)
    for t in tiles:
        O.spo_compute_tracer_tile(s, cam, t, img)
    O.spo_scene_destroy(s)
    return img.reshape(192, 256, 4)


def test_light_seen_directly_is_one_and_empty_is_zero():
    desc = compute_scene(sphere=False, ground=False)
    # look straight up at the light from below it
    desc["camera"] = (256, 192, (0.0, 1.0, 0.0), (0.70710677, 0.0, 0.0, 0.70710677), 0.8)
    img = _render(desc, range(192))
    assert np.all(img[..., :3] == 1.0) and np.all(img[..., 3] == 1.0)
    empty = compute_scene(light=False, sphere=False, ground=False)
    assert np.all(_render(empty, range(192))[..., :3] == 0.0)


def test_red_surface_under_white_light():
    desc = compute_scene(ground=False)
    desc["camera"] = (256, 192, (0.0, 1.0, 3.0), (0, 0, 0, 1), 0.8)
    img = _render(desc, range(192))
    rgb = img[..., :3]
    sphere = rgb[..., 0] > 0
    assert sphere.mean() > 0.01
    assert np.array_equal(rgb[..., 1], rgb[..., 2])          # G = B exactly
    assert np.all(rgb[..., 0] <= 0.18 * (1 + 1e-6))          # albedo x 1 x cos
    assert np.all(rgb[..., 1] <= 0.1 * (1 + 1e-6))
    assert np.all(rgb[sphere, 0] > rgb[sphere, 1])


def test_tiles_write_their_pixels_and_differ_by_seed():
    desc = compute_scene()
    img = _render(desc, [0, 17, 191])
    written = img[..., 3] == 1.0
    want = np.zeros((192, 256), bool)
    for t in (0, 17, 191):
        x, y = (t % 16) * 16, (t // 16) * 16
        want[y:y + 16, x:x + 16] = True
    assert np.array_equal(written, want)
    a = _render(desc, [40])
    b = _render(desc, [40])
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
