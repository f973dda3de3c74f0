"""The parked Vulkan compute tracer's semantics on the GPU
(spg_compute_tracer_tiles, src/shaders/test_compute.comp.glsl) against the
oracle's restatement (oracle/spo_compute.c): every pixel of every tile bit
for bit, a subset of tiles writing only its own pixels."""
import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes
from _compute_scene import camera_transform, compute_scene
from _libs import oracle
from _scenes import load_oracle

pytestmark = pytest.mark.gpu


def _oracle(desc, tiles, fill):
    O = oracle()
    s, _ = load_oracle(desc, build_midphase=False)
    img = np.full(256 * 192 * 4, fill, np.float32)
    cam = camera_transform(desc)
    for t in tiles:
        O.spo_compute_tracer_tile(s, cam, int(t), img)
    O.spo_scene_destroy(s)
    return img


@pytest.mark.parametrize("tiles", [list(range(192)), [0, 5, 77, 150, 191]])
def test_compute_tracer_matches_oracle(device, tiles):
    desc = compute_scene()
    s, _ = scenes.load_gpu(device, desc)
    img = np.full(256 * 192 * 4, -7.0, np.float32)
    V.check(V.lib().spg_compute_tracer_tiles(s.handle, camera_transform(desc),
                                             np.asarray(tiles, np.uint32), len(tiles), img))
    want = _oracle(desc, tiles, -7.0)
    assert np.array_equal(img.view(np.uint32), want.view(np.uint32))
    lit = img.reshape(-1, 4)[:, 0]
    assert (lit > 0).mean() > 0.1 * len(tiles) / 192
    s.close()


def test_compute_tracer_rejects_bad_tiles(device):
    desc = compute_scene()
    s, _ = scenes.load_gpu(device, desc)
    img = np.zeros(256 * 192 * 4, np.float32)
    assert V.lib().spg_compute_tracer_tiles(s.handle, camera_transform(desc),
                                            np.asarray([192], np.uint32), 1, img) == -1
    s.close()
