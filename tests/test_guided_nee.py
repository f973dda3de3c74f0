"""The guided estimator with next-event estimation (SPG_EST_GUIDED_NEE; no
reference counterpart) in the oracle, with the verification mixture: three
sampling techniques -- BSDF and guide mixture by one-sample MIS, the light
sampler of SPG_EST_NEE -- combined by the power heuristic.  Checked against
a closed form (the inside of an emitting sphere, where every vertex sees the
light and a dropped or doubled MIS weight moves the mean), against the plain
unbiased estimator's mean, and for the error reduction NEE brings over the
guided estimator alone; any sky/emitter split of the light samples (the
product learns one) leaves the mean unchanged (CPU only)."""
import ctypes as C
import math

import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes
from _libs import (SPO_EST_GUIDED, SPO_EST_GUIDED_NEE, SPO_EST_NEE, SPO_EST_UNBIASED,
                   SPO_HIT_BRUTE, SPO_RNG_PCG, SPO_TX_CR, SpoConfig, oracle)
from _scenes import load_oracle
from test_nee import _lamps_scene


def _cfg(spp, bounces, seed, est, fb=0.5):
    return SpoConfig(rng=SPO_RNG_PCG, tx=SPO_TX_CR, hit=SPO_HIT_BRUTE, estimator=est,
                     samplesPerPixel=spp, bounces=bounces, radianceClamp=0.0,
                     jitterYFirst=1, seed=seed, bsdfFraction=fb)


def _render(O, s, cam, w, h, cfg):
    px = np.zeros(w * h * 4, np.float32)
    met = np.zeros(12, np.uint64)
    assert O.spo_path_trace_tile(s, C.byref(cam), C.byref(cfg), 0, 0, w, h, None, px,
                                 met) == 0
    return px.reshape(-1, 4)[:, :3].astype(np.float64), met


def test_closed_emitting_sphere_reaches_the_geometric_series():
    """Inside a closed sphere of albedo a whose every triangle emits E, N
    path vertices: the expectation is E (1 - q^N) / (1 - q) with q = a (1 -
    F) and F between 0.04 and 1 (the reference's Schlick term), so it lies in
    (E, E (1 - a^N) / (1 - a)) and near the F = 0.04 value.  Every scattering
    vertex takes a light sample of the sphere itself; emission reached by a
    scattered ray counts with its power-heuristic share only, so counting
    it whole (or dropping the light sample) moves the mean out of the
    bracket.  Shadow rays count as rays."""
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
    px, met = _render(O, s, cam, 16, 16, _cfg(64, N, 3, SPO_EST_GUIDED_NEE))
    q = a * (1 - 0.04)
    lo, mid, hi = E, E * (1 - q ** N) / (1 - q), E * (1 - a ** N) / (1 - a)
    mean = px[:, 0].mean()
    assert lo < mean < hi and mean == pytest.approx(mid, rel=0.05)
    # at most a camera ray, N - 1 scattered and N - 1 shadow rays per path
    # (a mixture direction below the surface ends the path, a light sample
    # facing away casts no ray); most paths have them all
    paths = 16 * 16 * 64
    assert 0.95 * paths * (2 * N - 1) < int(met[2]) <= paths * (2 * N - 1)
    O.spo_scene_destroy(s)


def test_agrees_with_unbiased_and_cuts_the_guided_error():
    """Five emissive spheres over a floor, black sky: the guided + NEE
    image mean equals the plain unbiased estimator's within 4 standard
    errors, and at equal samples its squared error against a long plain
    render is well below the guided estimator's alone (the verification
    mixture does not know where the lamps are; the light sampler does).
    Every training-sink weight is positive and finite."""
    O = oracle()
    desc = _lamps_scene()
    s, cam = load_oracle(desc)
    ref, _ = _render(O, s, cam, 16, 12, _cfg(4096, 3, 99, SPO_EST_UNBIASED))
    sink = np.zeros((100000, 16), np.float32)
    O.spo_guided_sink(sink.ctypes.data, len(sink))
    gn = [_render(O, s, cam, 16, 12, _cfg(16, 3, k, SPO_EST_GUIDED_NEE))[0] for k in range(8)]
    n = O.spo_guided_sink_count()
    O.spo_guided_sink(None, 0)
    g = [_render(O, s, cam, 16, 12, _cfg(16, 3, k, SPO_EST_GUIDED))[0] for k in range(8)]
    u = [_render(O, s, cam, 16, 12, _cfg(16, 3, 50 + k, SPO_EST_UNBIASED))[0]
         for k in range(8)]
    mu, mg = np.mean(u), np.mean(gn)
    se = math.sqrt(np.var([x.mean() for x in u]) / 8 + np.var([x.mean() for x in gn]) / 8)
    assert abs(mu - mg) < 4 * se + 1e-4
    eg = np.mean([((x - ref) ** 2).mean() for x in g])
    egn = np.mean([((x - ref) ** 2).mean() for x in gn])
    assert egn < eg / 3, (egn, eg)
    assert n > 100
    w = sink[:n, 12]
    assert np.all(w > 0) and np.all(np.isfinite(w))
    O.spo_scene_destroy(s)


def _sky_and_lamp_scene(w=20, h=16):
    """Config 1 (sphere on a ground quad under the sky cube with its sun)
    plus one emissive sphere: light samples split between sky and
    emitter."""
    desc = scenes.config1(w, h, env="cube", env_size=16)
    desc["objects"] = list(desc["objects"]) + [
        (0, 777, (1.6, 0.6, 0.8), (0.0, 0.0, 0.0, 1.0), (0.3, 0.3, 0.3))]
    desc["materials"] = list(desc["materials"]) + [
        (777, (0.0, 0.0, 0.0), (12.0, 9.0, 6.0), 0.0, V.SPG_NO_TEXTURE, V.SPG_NO_TEXTURE)]
    return desc


@pytest.mark.parametrize("est", [SPO_EST_GUIDED_NEE, SPO_EST_NEE])
def test_any_light_split_is_unbiased(est):
    """The sky share of the light samples (the product's learned light
    split) changes only the variance: with the sky chosen for 0.1 or 0.8 of
    the light samples the image mean equals the plain estimator's within 4
    standard errors."""
    O = oracle()
    desc = _sky_and_lamp_scene()
    s, cam = load_oracle(desc)
    u = [_render(O, s, cam, 20, 16, _cfg(32, 3, 300 + k, SPO_EST_UNBIASED))[0].mean()
         for k in range(6)]
    for split in (0.1, 0.8):
        O.spo_scene_set_light_split(s, split)
        g = [_render(O, s, cam, 20, 16, _cfg(32, 3, 400 + k, est))[0].mean() for k in range(6)]
        se = math.sqrt(np.var(u) / 6 + np.var(g) / 6)
        assert abs(np.mean(u) - np.mean(g)) < 4 * se + 1e-4, (split, np.mean(u), np.mean(g))
    O.spo_scene_destroy(s)
