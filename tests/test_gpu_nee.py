"""Next-event estimation on the GPU (SPG_EST_NEE): bit-exact against the
oracle's restatement (same per-pixel PCG or xorshift streams, double-
evaluated transcendentals), with and without a cube-map environment, on
config 1 and on tiles of the million-triangle box scene; and the error it
removes at equal samples."""
import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes
from _libs import SPO_EST_NEE, SPO_RNG_PCG, SPO_RNG_XORSHIFT_PIXEL, oracle
from _scenes import load_oracle
from test_gpu_path_trace import _oracle_tile_sah, gpu_render, oracle_render

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rng,orng", [(V.RNG_PCG, SPO_RNG_PCG),
                                      (V.RNG_XORSHIFT_PIXEL, SPO_RNG_XORSHIFT_PIXEL)])
@pytest.mark.parametrize("env", ["cube", "equirect"])
def test_nee_config1_matches_oracle_bit_exact(device, rng, orng, env):
    """Every pixel's bits and the ray (shadow rays included), hit and miss
    counts equal the oracle's; an equirect background has no table, so the
    estimator draws its numbers but casts no shadow rays."""
    desc = scenes.config1(64, 48, env=env, env_size=32)
    img, m = gpu_render(device, desc, 4, 3, V.EST_NEE, rng, clamp=0.0)
    px, met = oracle_render(desc, 4, 3, SPO_EST_NEE, orng, clamp=0.0)
    assert np.array_equal(img.view(np.uint32), px.view(np.uint32))
    assert [m.values[k] for k in (1, 2, 3, 4)] == [int(met[k]) for k in (1, 2, 3, 4)]


@pytest.mark.parametrize("name,tiles", [
    ("config4", [(224, 224, 288, 288), (0, 448, 64, 512)]),
    ("config5", [(928, 508, 992, 572), (0, 1040, 64, 1080)]),
])
def test_nee_box_scene_tiles_match_oracle_bit_exact(device, name, tiles):
    """The million-triangle box scenes, 2 spp and 6 bounces: config 4 (the
    sky only, through the opening) and config 5 (the sky and 54 emissive
    spheres: light samples split between the sky table and the ~1.1M
    emissive triangles): tiles equal the oracle's over the same SAH
    hierarchy bit for bit, as do the ray (shadow rays included), hit and
    miss counts."""
    desc = getattr(scenes, name)()
    w, h = desc["camera"][0], desc["camera"][1]
    s, cam = scenes.load_gpu(device, desc)
    ctx = V.make_context(cam, s, spp=2, bounces=6, clamp=0.0, estimator=V.EST_NEE,
                         rng=V.RNG_PCG, seed=0x5EED, collect=True)
    os_, ocam = load_oracle(desc, sah=True)
    for t in tiles:
        tile = V.Tile(*t)
        img = np.zeros(w * h * 4, np.float32)
        m = V.sp_PathTraceTile(ctx, tile, img)
        px, met = _oracle_tile_sah(os_, ocam, desc, tile, 2, 6, SPO_EST_NEE)
        a = img.reshape(h, w, 4)[t[1]:t[3], t[0]:t[2]]
        b = px.reshape(h, w, 4)[t[1]:t[3], t[0]:t[2]]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (name, t)
        assert [m.values[k] for k in (2, 3, 4)] == [int(met[k]) for k in (2, 3, 4)]
    oracle().spo_scene_destroy(os_)
    s.close()


def test_nee_cuts_error_at_equal_samples(device):
    """Config 1 at 128x128 (sun of radiance 50 in the sky cube): against a
    4096-spp plain render, 16-spp NEE images have less than a third of the
    plain tracer's relative squared error, and the two estimators' image
    means agree within 1%."""
    desc = scenes.config1(128, 128)
    ref, _ = gpu_render(device, desc, 4096, 3, V.EST_UNBIASED, V.RNG_PCG, clamp=0.0,
                        seed=1, collect=False)
    ref = ref.reshape(-1, 4)[:, :3].astype(np.float64)

    def rel(img):
        x = img.reshape(-1, 4)[:, :3].astype(np.float64)
        return float((((x - ref) ** 2) / (ref * ref + 1e-2)).mean()), x.mean()

    eu, mu = zip(*[rel(gpu_render(device, desc, 16, 3, V.EST_UNBIASED, V.RNG_PCG,
                                  clamp=0.0, seed=100 + k, collect=False)[0])
                   for k in range(4)])
    en, mn = zip(*[rel(gpu_render(device, desc, 16, 3, V.EST_NEE, V.RNG_PCG,
                                  clamp=0.0, seed=200 + k, collect=False)[0])
                   for k in range(4)])
    assert np.mean(en) < np.mean(eu) / 3, (en, eu)
    assert np.mean(mn) == pytest.approx(np.mean(mu), rel=0.01)


def test_nee_emitter_scene_matches_oracle_bit_exact(device):
    """A floor lit only by a small emissive sphere (no sky table: every light
    sample goes to the emitter table): bit-exact against the oracle."""
    from test_nee import _lamp_scene
    desc = _lamp_scene()
    for rng, orng in [(V.RNG_PCG, SPO_RNG_PCG), (V.RNG_XORSHIFT_PIXEL, SPO_RNG_XORSHIFT_PIXEL)]:
        img, m = gpu_render(device, desc, 8, 3, V.EST_NEE, rng, clamp=0.0)
        px, met = oracle_render(desc, 8, 3, SPO_EST_NEE, orng, clamp=0.0)
        assert np.array_equal(img.view(np.uint32), px.view(np.uint32))
        assert [m.values[k] for k in (1, 2, 3, 4)] == [int(met[k]) for k in (1, 2, 3, 4)]


def test_nee_error_paths(device):
    """An unknown estimator is an error, not an abort, in a trace call and
    in a batch; NEE with float transcendentals renders a finite image."""
    desc = scenes.config1(32, 24)
    s, cam = scenes.load_gpu(device, desc)
    img = np.zeros(32 * 24 * 4, np.float32)
    ctx = V.make_context(cam, s, spp=1, bounces=3, estimator=7)
    with pytest.raises(V.SpgError):
        V.path_trace_tiles(ctx, V.ComputeTiles(32, 24, 64, 64), img)
    buf = device.alloc(32 * 24 * 16)
    with pytest.raises(V.SpgError):
        V.Batch(device, [(cam, s, buf)], spp=1, bounces=3, estimator=7)
    ctx = V.make_context(cam, s, spp=4, bounces=3, clamp=0.0, estimator=V.EST_NEE,
                         transcendentals=V.TX_FAST)
    V.path_trace_tiles(ctx, V.ComputeTiles(32, 24, 64, 64), img)
    assert np.isfinite(img).all() and img.reshape(-1, 4)[:, :3].mean() > 0
    buf.free()
    s.close()


def test_nee_spatial_emitter_choice_matches_oracle_bit_exact(device):
    """Five emissive spheres of different size, height and power over a
    floor (the spatial emitter choice: an emitter instance by power over
    squared distance from the shading point, then a triangle by area):
    bit-exact against the oracle in both RNG modes, with equal ray, hit and
    miss counts."""
    from test_nee import _lamps_scene
    desc = _lamps_scene()
    for rng, orng in [(V.RNG_PCG, SPO_RNG_PCG), (V.RNG_XORSHIFT_PIXEL, SPO_RNG_XORSHIFT_PIXEL)]:
        img, m = gpu_render(device, desc, 8, 4, V.EST_NEE, rng, clamp=0.0)
        px, met = oracle_render(desc, 8, 4, SPO_EST_NEE, orng, clamp=0.0)
        assert np.array_equal(img.view(np.uint32), px.view(np.uint32))
        assert [m.values[k] for k in (1, 2, 3, 4)] == [int(met[k]) for k in (1, 2, 3, 4)]


@pytest.mark.parametrize("which", ["buried", "many"])
def test_nee_emitter_choice_edges_match_oracle_bit_exact(device, which):
    """Edges of the emitter choice, bit-exact against the oracle: lamps
    below the floor with one above (at points on the upper lamp every
    emitter's bounding sphere lies behind the tangent plane, total weight 0,
    no emitter sample); 300 tiny lamps (more than the spatial choice scans:
    the global table by power)."""
    from test_nee import _buried_lamps_scene, _many_lamps_scene
    desc = _buried_lamps_scene() if which == "buried" else _many_lamps_scene()
    img, m = gpu_render(device, desc, 8, 4, V.EST_NEE, V.RNG_PCG, clamp=0.0)
    px, met = oracle_render(desc, 8, 4, SPO_EST_NEE, SPO_RNG_PCG, clamp=0.0)
    assert np.array_equal(img.view(np.uint32), px.view(np.uint32))
    assert [m.values[k] for k in (1, 2, 3, 4)] == [int(met[k]) for k in (1, 2, 3, 4)]
    assert img.reshape(-1, 4)[:, :3].max() > 0


def test_nee_learned_light_split(device):
    """A scene told to learn its light split: NEE calls on config 5
    (480x270) add the luminance their reached light samples delivered and
    later calls draw with the learned sky share -- none before the first
    call, within [0.05, 0.95] and below the default 0.5 after (the sky is
    mostly hidden by the box), the same bits on a second scene given the
    same calls, forgotten when learning stops.  Images stay finite and
    their mean matches the default split's within the Monte Carlo error."""
    desc = scenes.config5(480, 270)
    tiles = V.ComputeTiles(480, 270, 64, 64)
    splits, means = [], []
    for rep in range(2):
        s, cam = scenes.load_gpu(device, desc)
        s.learn_light_split(True)
        assert s.light_split() == -1.0
        img = np.zeros(480 * 270 * 4, np.float32)
        m = []
        for k in range(4):
            V.path_trace_tiles(V.make_context(cam, s, spp=4, bounces=6, clamp=0.0,
                                              estimator=V.EST_NEE, rng=V.RNG_PCG,
                                              seed=11, sample_offset=4 * k,
                                              transcendentals=V.TX_FAST), tiles, img)
            assert np.isfinite(img).all()
            m.append(img.reshape(-1, 4)[:, :3].mean())
        splits.append(s.light_split())
        means.append(m)
        s.learn_light_split(False)
        assert s.light_split() == -1.0
        s.close()
    assert splits[0] == splits[1] and means[0] == means[1]
    assert 0.05 <= splits[0] < 0.5
    # the default split's mean over the same samples, for comparison
    s, cam = scenes.load_gpu(device, desc)
    img = np.zeros(480 * 270 * 4, np.float32)
    d = []
    for k in range(4):
        V.path_trace_tiles(V.make_context(cam, s, spp=4, bounces=6, clamp=0.0,
                                          estimator=V.EST_NEE, rng=V.RNG_PCG, seed=11,
                                          sample_offset=4 * k, transcendentals=V.TX_FAST),
                           tiles, img)
        d.append(img.reshape(-1, 4)[:, :3].mean())
    s.close()
    learned, default = np.array(means[0][1:]), np.array(d[1:])
    se = np.sqrt(learned.var(ddof=1) / 3 + default.var(ddof=1) / 3)
    assert abs(learned.mean() - default.mean()) < 5 * se + 0.02 * default.mean()
