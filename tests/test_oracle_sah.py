"""The oracle walking the product's SAH hierarchy (the mode used for the
million-triangle scenes, where the reference's greedy builder does not
scale) returns exactly what the reference-restating greedy walk and the
brute-force closest hit return."""
import ctypes as C

import numpy as np
import pytest

from vk_cinematic_amd import scenes
from _libs import (SPO_EST_PARITY, SPO_EST_UNBIASED, SPO_HIT_BRUTE,
                   SPO_HIT_REFERENCE, SPO_RNG_PCG, SPO_TX_CR, SpoConfig, oracle)
from _scenes import load_oracle


def _render(s, cam, w, h, est, hit, spp=2, bounces=3, seed=5):
    cfg = SpoConfig(rng=SPO_RNG_PCG, tx=SPO_TX_CR, hit=hit, estimator=est,
                    samplesPerPixel=spp, bounces=bounces, radianceClamp=10.0,
                    jitterYFirst=1, seed=seed)
    px = np.zeros(w * h * 4, np.float32)
    met = np.zeros(12, np.uint64)
    assert oracle().spo_path_trace_tile(s, C.byref(cam), C.byref(cfg), 0, 0, w, h,
                                        None, px, met) == 0
    return px, met


@pytest.mark.parametrize("est", [SPO_EST_PARITY, SPO_EST_UNBIASED])
def test_sah_walk_equals_greedy_walk_config1(est):
    desc = scenes.config1(64, 64, env="cube", env_size=32)
    a, cam = load_oracle(desc)
    b, _ = load_oracle(desc, sah=True)
    pa, ma = _render(a, cam, 64, 64, est, SPO_HIT_REFERENCE)
    pb, mb = _render(b, cam, 64, 64, est, SPO_HIT_REFERENCE)
    oracle().spo_scene_destroy(a)
    oracle().spo_scene_destroy(b)
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
    assert [int(ma[k]) for k in (1, 2, 3, 4)] == [int(mb[k]) for k in (1, 2, 3, 4)]


def test_sah_walk_equals_brute_force_box_scene():
    """Rotated, scaled instances of displaced icospheres inside the box with
    the opening (the large workloads' layout at a small size): the SAH walk
    and the brute-force closest hit agree on every pixel."""
    desc = scenes.box_scene(3, 2, 48, 32, seed=11)
    s, cam = load_oracle(desc, sah=True)
    pa, ma = _render(s, cam, 48, 32, SPO_EST_UNBIASED, SPO_HIT_REFERENCE,
                     bounces=6)
    pb, mb = _render(s, cam, 48, 32, SPO_EST_UNBIASED, SPO_HIT_BRUTE, bounces=6)
    oracle().spo_scene_destroy(s)
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
    assert int(ma[2]) == int(mb[2])
    # the scene is lit through the opening and the rays bounce inside
    assert pa.reshape(-1, 4)[:, :3].mean() > 0
    assert int(ma[2]) > 48 * 32 * 2


def test_box_scene_sizes():
    d4 = scenes.config4()
    d5 = scenes.config5()
    assert d4["camera"][:2] == (512, 512) and d5["camera"][:2] == (1920, 1080)
    assert 1_000_000 <= d4["triangles"] < 1_100_000
    assert 4_000_000 <= d5["triangles"] < 4_100_000
    # every object's material is registered
    ids = {m[0] for m in d5["materials"]}
    assert all(o[1] in ids for o in d5["objects"])


def test_material_registry_capacity():
    """sp_RegisterMaterial (src/sp_material_system.cpp:1-13) refuses the
    33rd material (SP_MAX_MATERIALS 32); the large scenes lift the cap."""
    from _libs import f32
    O = oracle()
    s = O.spo_scene_create()
    ok = [O.spo_material_register(s, f32(1, 1, 1), 0, f32(0, 0, 0), 0, 0.0, k)
          for k in range(33)]
    assert ok == [1] * 32 + [0]
    O.spo_scene_set_limits(s, 32, 128, 64)
    assert O.spo_material_register(s, f32(1, 1, 1), 0, f32(0, 0, 0), 0, 0.0, 99) == 1
    O.spo_scene_destroy(s)
