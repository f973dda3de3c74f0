"""The guided estimator with next-event estimation (SPG_EST_GUIDED_NEE) on
the GPU against the oracle's restatement (spo_trace.c trace_guided with
SPO_EST_GUIDED_NEE), with the verification mixture (spg_guide_debug
SPG_GUIDE_ANALYTIC_MIXTURE) and double-evaluated transcendentals: bit for
bit on config 1 (light samples of the sky), on emissive spheres over a
floor (the spatial emitter choice) and on tiles of config 5 (sky and 54
emissive spheres over the SAH hierarchy); the training samples of a pass;
the graph-captured batch against eager calls; and config 5 at full size
with an online-trained network against the plain tracer's mean."""
import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes
from _libs import (SPO_EST_GUIDED_NEE, SPO_HIT_REFERENCE, SPO_RNG_PCG,
                   SPO_RNG_XORSHIFT_PIXEL)
from test_gpu_guided_parity import _oracle_guided

pytestmark = pytest.mark.gpu


def _gpu(device, desc, spp, bounces, rng, seed=0x5EED, train=False, tiles=None):
    w, h = desc["camera"][0], desc["camera"][1]
    s, cam = scenes.load_gpu(device, desc)
    g = V.Guide(device, seed=1)
    g.debug(analytic=True)
    ctx = V.make_context(cam, s, spp=spp, bounces=bounces, clamp=0.0,
                         estimator=V.EST_GUIDED_NEE, rng=rng, seed=seed, guide=g,
                         collect=True, train=train)
    img = np.zeros(w * h * 4, np.float32)
    m = V.path_trace_tiles(ctx, tiles or V.ComputeTiles(w, h, 64, 64), img)
    samples = g.samples() if train else None
    g.close()
    s.close()
    return img, m, samples


def _lamps():
    from test_nee import _lamps_scene
    return _lamps_scene()


@pytest.mark.parametrize("rng,orng", [(V.RNG_PCG, SPO_RNG_PCG),
                                      (V.RNG_XORSHIFT_PIXEL, SPO_RNG_XORSHIFT_PIXEL)])
@pytest.mark.parametrize("scene", ["config1", "lamps"])
def test_small_scenes_bit_exact(device, rng, orng, scene):
    """Config 1 with its sky cube (every light sample goes to the sky
    table) and five emissive spheres over a floor under a black sky (every
    light sample goes to the spatial emitter choice): pixels, ray (shadow
    rays included), hit and miss counts equal the oracle's."""
    desc = (scenes.config1(64, 48, env="cube", env_size=32) if scene == "config1"
            else _lamps())
    img, m, _ = _gpu(device, desc, 4, 4, rng)
    px, met, _ = _oracle_guided(desc, 4, 4, orng, 0x5EED, est=SPO_EST_GUIDED_NEE)
    assert np.array_equal(img.view(np.uint32), px.view(np.uint32))
    assert [m.values[k] for k in (2, 3, 4)] == [int(met[k]) for k in (2, 3, 4)]
    # shadow rays count as rays, not as hits or misses of the paths
    assert m.values[3] + m.values[4] < m.values[2] < 2 * (m.values[3] + m.values[4])
    assert np.isfinite(img).all() and img.reshape(-1, 4)[:, :3].max() > 0


@pytest.mark.parametrize("split", [None, 0.2])
def test_config5_tiles_bit_exact(device, split):
    """Tiles of the full-size config 5 (4M triangles; the sky through the
    opening and 54 emissive spheres), 2 spp, 6 bounces, with the default
    sky/emitter split of the light samples and with a fixed one (what the
    learned split sets): bit for bit against the oracle over the same SAH
    hierarchy, equal ray, hit and miss counts."""
    desc = scenes.config5()
    w, h = desc["camera"][0], desc["camera"][1]
    s, cam = scenes.load_gpu(device, desc)
    g = V.Guide(device, seed=1)
    g.debug(analytic=True)
    if split is not None:
        g.set_light_split(split)
        assert g.light_split() == np.float32(split)
    ctx = V.make_context(cam, s, spp=2, bounces=6, clamp=0.0,
                         estimator=V.EST_GUIDED_NEE, rng=V.RNG_PCG, seed=0x5EED, guide=g,
                         collect=True)
    for t in [(928, 508, 992, 572), (0, 1040, 64, 1080)]:
        img = np.zeros(w * h * 4, np.float32)
        m = V.sp_PathTraceTile(ctx, V.Tile(*t), img)
        px, met, _ = _oracle_guided(desc, 2, 6, SPO_RNG_PCG, 0x5EED, hit=SPO_HIT_REFERENCE,
                                    sah=True, tile=t, est=SPO_EST_GUIDED_NEE,
                                    light_split=split)
        a = img.reshape(h, w, 4)[t[1]:t[3], t[0]:t[2]]
        b = px.reshape(h, w, 4)[t[1]:t[3], t[0]:t[2]]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), t
        assert [m.values[k] for k in (2, 3, 4)] == [int(met[k]) for k in (2, 3, 4)]
    g.close()
    s.close()


def test_training_samples_bit_exact(device):
    """A pass small enough that every path is selected, on the lamps scene:
    the collected training samples are exactly the oracle's candidates
    (matched by the bits of n, v, wi), each weight -- which now holds the
    MIS-weighted emission and the light samples of the vertices after it --
    bit for bit."""
    desc = _lamps()
    img, _, got = _gpu(device, desc, 32, 4, V.RNG_PCG, seed=0x77, train=True)
    px, _, cand = _oracle_guided(desc, 32, 4, SPO_RNG_PCG, 0x77, sink=100000,
                                 est=SPO_EST_GUIDED_NEE)
    assert np.array_equal(img.view(np.uint32), px.view(np.uint32))
    assert len(cand) > 50 and len(got) == len(cand)
    key = lambda r: r.view(np.uint32).tobytes()  # noqa: E731
    want = {key(r[3:12]): r for r in cand}
    assert len(want) == len(cand)
    for r in got:
        assert r[12].view(np.uint32) == want[key(r[3:12])][12].view(np.uint32)


def test_batch_frozen_equals_eager(device):
    """The graph-captured batch (guide inference, light samples and shadow
    rays in the graph) with a frozen network: two cameras on the lamps
    scene, each item equal to the eager render bit for bit."""
    desc = _lamps()
    s, cam = scenes.load_gpu(device, desc)
    cam2 = scenes.sp_ConfigureCamera(24, 20, (1.0, 4.0, 6.0), (0, 0, 0, 1), 0.8)
    g = V.Guide(device, seed=3, batch_size=4096, steps_per_pass=2)
    for k in range(4):
        V.path_trace_tiles(V.make_context(cam, s, spp=4, bounces=4, clamp=0.0,
                                          estimator=V.EST_GUIDED_NEE, rng=V.RNG_PCG,
                                          seed=9, sample_offset=4 * k, guide=g, train=True),
                           V.ComputeTiles(16, 12, 64, 64), np.zeros(16 * 12 * 4, np.float32))
    kw = dict(spp=3, bounces=4, clamp=0.0, estimator=V.EST_GUIDED_NEE, rng=V.RNG_PCG,
              seed=0x77, transcendentals=V.TX_DOUBLE, guide=g)
    imgs = [device.alloc(16 * 12 * 16), device.alloc(24 * 20 * 16)]
    b = V.Batch(device, [(cam, s, imgs[0]), (cam2, s, imgs[1])], autotune=False, **kw)
    b.run()
    for c, img, w, h in ((cam, imgs[0], 16, 12), (cam2, imgs[1], 24, 20)):
        got = img.download(np.float32, w * h * 4)
        want = np.zeros(w * h * 4, np.float32)
        V.path_trace_tiles(V.make_context(c, s, **kw), V.ComputeTiles(w, h, 64, 64), want)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert np.isfinite(got).all()
    b.close()
    for img in imgs:
        img.free()
    g.close()
    s.close()


def test_config5_trained_mean_equals_plain(device):
    """Config 5 at 1920x1080 with an online-trained network (12 passes with
    light samples): every pixel finite, the image mean equal to the plain
    tracer's within the Monte Carlo error."""
    desc = scenes.config5()
    w, h = desc["camera"][0], desc["camera"][1]
    s, cam = scenes.load_gpu(device, desc)
    tiles = V.ComputeTiles(w, h, 64, 64)
    dimg = device.alloc(w * h * 16)
    g = V.Guide(device, seed=1, max_weight=10.0)
    for k in range(12):
        V.path_trace_tiles_device(V.make_context(cam, s, spp=1, bounces=6, clamp=0.0,
                                                 estimator=V.EST_GUIDED_NEE, rng=V.RNG_PCG,
                                                 seed=5, sample_offset=k, guide=g,
                                                 train=True, transcendentals=V.TX_FAST),
                                  tiles, dimg)
    assert g.stats()["steps"] == 48

    def means(est, seed, reps, spp):
        out = []
        for k in range(reps):
            ctx = V.make_context(cam, s, spp=spp, bounces=6, clamp=0.0, estimator=est,
                                 rng=V.RNG_PCG, seed=seed + k, guide=g,
                                 transcendentals=V.TX_FAST)
            V.path_trace_tiles_device(ctx, tiles, dimg)
            img = dimg.download(np.float32, w * h * 4).reshape(h, w, 4)[:, :, :3]
            assert np.isfinite(img).all()
            out.append(img.mean())
        return np.array(out)

    gm = means(V.EST_GUIDED_NEE, 100, 4, 2)
    pm = means(V.EST_UNBIASED, 200, 4, 2)
    se = np.sqrt(gm.var(ddof=1) / 4 + pm.var(ddof=1) / 4)
    print(f"\nconfig5: guided+NEE {gm.mean():.5f} +- {gm.std():.5f}, plain "
          f"{pm.mean():.5f} +- {pm.std():.5f}")
    assert abs(gm.mean() - pm.mean()) < 5 * se + 2e-3 * pm.mean()
    g.close()
    dimg.free()
    s.close()


def test_light_split_is_learned_and_reproducible(device, tmp_path):
    """Training passes on config 5 (480x270) learn the sky share of the
    light samples from the luminance they delivered: none before training,
    within [0.05, 0.95] after, below the default 0.5 (most sky samples end
    on the box's walls; the 54 emissive spheres deliver most of the direct
    light), the same bits from two trainings with the same seed, and kept
    through a checkpoint."""
    desc = scenes.config5(480, 270)
    s, cam = scenes.load_gpu(device, desc)
    tiles = V.ComputeTiles(480, 270, 64, 64)
    img = np.zeros(480 * 270 * 4, np.float32)
    splits = []
    for rep in range(2):
        g = V.Guide(device, seed=2, batch_size=4096, steps_per_pass=2)
        assert g.light_split() == -1.0
        for k in range(6):
            V.path_trace_tiles(V.make_context(cam, s, spp=1, bounces=6, clamp=0.0,
                                              estimator=V.EST_GUIDED_NEE, rng=V.RNG_PCG,
                                              seed=3, sample_offset=k, guide=g, train=True,
                                              transcendentals=V.TX_FAST), tiles, img)
        splits.append(g.light_split())
        if rep == 1:
            g.save(tmp_path / "g.ckpt")
            g2 = V.Guide(device, seed=9)
            g2.load(tmp_path / "g.ckpt")
            assert g2.light_split() == splits[-1]
            g2.set_light_split(-1)
            assert g2.light_split() == -1.0
            g2.close()
        g.close()
    print(f"\nlearned sky share {splits[0]:.4f}")
    assert splits[0] == splits[1]
    assert 0.05 <= splits[0] < 0.5
    s.close()


def test_batch_training_replays_equal_eager_training(device):
    """Online training in the captured graph with light samples: config 1
    with an emissive sphere (sky and emitter light), each replay one pass
    with its sample collection, Adam steps and light sums.  After five
    replays the guide's parameters and learned light split equal those of
    five eager training traces bit for bit (the split learned after the
    first pass is the one the later passes draw with), and so do the last
    images."""
    from test_guided_nee import _sky_and_lamp_scene
    desc = _sky_and_lamp_scene(96, 64)
    s, cam = scenes.load_gpu(device, desc)
    w, h = 96, 64
    kw = dict(spp=1, bounces=4, clamp=0.0, estimator=V.EST_GUIDED_NEE, rng=V.RNG_PCG,
              seed=0x99, transcendentals=V.TX_FAST)
    ga = V.Guide(device, seed=5, batch_size=2048, steps_per_pass=2)
    gb = V.Guide(device, seed=5, batch_size=2048, steps_per_pass=2)
    img = device.alloc(w * h * 16)
    b = V.Batch(device, [(cam, s, img)], autotune=False, guide=ga, train=True, **kw)
    eager = device.alloc(w * h * 16)
    for _ in range(5):
        b.run()
        V.path_trace_tiles_device(V.make_context(cam, s, guide=gb, train=True, **kw),
                                  [V.Tile(0, 0, w, h)], eager)
    assert ga.stats()["steps"] == gb.stats()["steps"] == 10
    assert np.array_equal(ga.params().view(np.uint32), gb.params().view(np.uint32))
    assert ga.light_split() == gb.light_split() and 0.05 <= ga.light_split() <= 0.95
    a = img.download(np.float32, w * h * 4)
    e = eager.download(np.float32, w * h * 4)
    assert np.array_equal(a.view(np.uint32), e.view(np.uint32))
    b.close()
    for x in (img, eager):
        x.free()
    ga.close()
    gb.close()
    s.close()
