"""Batched rendering (spg_batch_*): many scenes in one wavefront replayed
from a HIP graph.  Every item's image equals the single-scene render of the
same camera and scene bit for bit (same per-pixel streams), across replays
and with other work on the device in between; the autotuner's choice is one
of its candidates."""
import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes

pytestmark = pytest.mark.gpu


def _load(device, descs):
    out = []
    for d in descs:
        s, cam = scenes.load_gpu(device, d)
        w, h = d["camera"][0], d["camera"][1]
        out.append((s, cam, device.alloc(w * h * 16), w, h))
    return out


def _single(device, s, cam, w, h, **kw):
    img = device.alloc(w * h * 16)
    ctx = V.make_context(cam, s, **kw)
    V.path_trace_tiles_device(ctx, V.ComputeTiles(w, h, 64, 64), img)
    out = img.download(np.float32, w * h * 4)
    img.free()
    return out


def _images(loaded):
    return [img.download(np.float32, w * h * 4) for _, _, img, w, h in loaded]


def _close(loaded):
    for s, _, img, _, _ in loaded:
        img.free()
        s.close()


@pytest.mark.parametrize("est,tx", [(V.EST_UNBIASED, V.TX_DOUBLE),
                                    (V.EST_PARITY, V.TX_DOUBLE),
                                    (V.EST_UNBIASED, V.TX_FAST)])
def test_batch_items_equal_single_scene_renders(device, est, tx):
    descs = scenes.config3(count=5, seed=11, min_side=40, max_side=150,
                           min_tris=2e4, max_tris=2e5)
    loaded = _load(device, descs)
    kw = dict(spp=3, bounces=5, clamp=0.0 if est == V.EST_UNBIASED else 10.0,
              estimator=est, rng=V.RNG_PCG, seed=0x5EED, transcendentals=tx)
    b = V.Batch(device, [(cam, s, img) for s, cam, img, _, _ in loaded],
                autotune=False, **kw)
    m = b.run()
    got = _images(loaded)
    for (s, cam, _, w, h), g in zip(loaded, got):
        want = _single(device, s, cam, w, h, **kw)
        assert np.array_equal(g.view(np.uint32), want.view(np.uint32))
    assert m.values[V.METRIC_PATHS] == sum(w * h for *_, w, h in loaded) * 3
    assert m.values[V.METRIC_RAYS] == m.values[V.METRIC_HITS] + m.values[V.METRIC_MISSES]
    b.close()
    _close(loaded)


def test_graph_replays_are_identical_with_work_between(device):
    descs = scenes.config3(count=3, seed=5, min_side=64, max_side=128,
                           min_tris=1e4, max_tris=5e4)
    loaded = _load(device, descs)
    b = V.Batch(device, [(cam, s, img) for s, cam, img, _, _ in loaded], spp=2,
                bounces=4, autotune=False)
    b.run()
    first = _images(loaded)
    for rep in range(3):
        # eager single-scene work on the same device and stream between replays
        s, cam, _, w, h = loaded[rep % 3]
        _single(device, s, cam, w, h, spp=1, bounces=2, estimator=V.EST_UNBIASED,
                clamp=0.0)
        for _, _, img, w, h in loaded:  # clear the outputs
            img.upload(np.zeros(w * h * 4, np.float32))
        b.run()
        again = _images(loaded)
        for a, g in zip(first, again):
            assert np.array_equal(a.view(np.uint32), g.view(np.uint32))
    b.close()
    _close(loaded)


def test_batch_larger_than_the_wavefront(device):
    """More pixels than paths in flight: slot chunks, each with its own
    sample passes."""
    descs = [scenes.grid_scene(100 + k, 3e4, 1024) for k in range(3)]
    loaded = _load(device, descs)
    kw = dict(spp=2, bounces=3, clamp=0.0, estimator=V.EST_UNBIASED, rng=V.RNG_PCG,
              transcendentals=V.TX_FAST)
    b = V.Batch(device, [(cam, s, img) for s, cam, img, _, _ in loaded],
                autotune=False, **kw)
    t = b.tuning()
    assert t["pixels"] == 3 * 1024 * 1024 > t["pathsInFlight"]
    b.run()
    got = _images(loaded)
    s, cam, _, w, h = loaded[2]
    want = _single(device, s, cam, w, h, **kw)
    assert np.array_equal(got[2].view(np.uint32), want.view(np.uint32))
    b.close()
    _close(loaded)


def test_autotuner_choice(device):
    """The tuner times capacity x traversal grid (3 shapes here: one
    capacity fits the 4 small images) and then two more refill thresholds;
    a second batch of the same configuration takes the remembered shape
    without timing; neither the shape nor the refill threshold changes the
    image (an untuned batch renders the same bits)."""
    descs = scenes.config3(count=4, seed=9, min_side=128, max_side=256,
                           min_tris=1e4, max_tris=1e5)
    loaded = _load(device, descs)
    items = [(cam, s, img) for s, cam, img, _, _ in loaded]
    b = V.Batch(device, items, spp=4, bounces=5, autotune=True)
    t = b.tuning()
    assert t["candidates"] == 5 and t["cached"] == 0
    assert t["pathsInFlight"] == 1 << 20
    assert t["refill"] in (0, 32, 48)
    assert t["tunedMsPerSample"] > 0
    assert t["graphNodes"] > 10
    b.run()
    assert b.last_ms() > 0
    tuned = _images(loaded)
    assert all(np.isfinite(g).all() and g.reshape(-1, 4)[:, :3].mean() > 0 for g in tuned)
    b.close()
    b2 = V.Batch(device, items, spp=4, bounces=5, autotune=True)
    t2 = b2.tuning()
    assert t2["cached"] == 1 and t2["candidates"] == 0
    for k in ("pathsInFlight", "extendBlocks", "refill", "tunedMsPerSample"):
        assert t2[k] == t[k], k
    b2.close()
    b3 = V.Batch(device, items, spp=4, bounces=5, autotune=False)
    assert b3.tuning()["refill"] == 0
    b3.run()
    for g, want in zip(_images(loaded), tuned):
        assert np.array_equal(g.view(np.uint32), want.view(np.uint32))
    b3.close()
    _close(loaded)


def test_batch_rejects_bad_items(device):
    d = scenes.config3(count=1, seed=1, min_side=16, max_side=16, min_tris=1e4,
                       max_tris=1e4)[0]
    s, cam = scenes.load_gpu(device, d)
    img = device.alloc(16 * 16 * 16)
    with pytest.raises(V.SpgError):
        V.Batch(device, [(cam, s, img)], estimator=V.EST_GUIDED)
    unbuilt = V.Scene(device)
    with pytest.raises(V.SpgError):
        V.Batch(device, [(cam, unbuilt, img)])
    unbuilt.close()
    img.free()
    s.close()


def test_batch_item_tiles_match_oracle(device):
    """A batch item against the oracle over the same SAH hierarchy (plain
    unbiased tracer, double-evaluated transcendentals, per-pixel PCG): two
    tiles of two items of a config-3 style batch, bit for bit."""
    import ctypes as C

    from _libs import SPO_HIT_REFERENCE, SPO_RNG_PCG, SPO_TX_CR, SpoConfig, oracle
    from _scenes import load_oracle

    descs = scenes.config3(count=4, seed=21, min_side=96, max_side=160,
                           min_tris=2e4, max_tris=3e5)
    loaded = _load(device, descs)
    kw = dict(spp=2, bounces=5, clamp=0.0, estimator=V.EST_UNBIASED, rng=V.RNG_PCG,
              seed=0x5EED, transcendentals=V.TX_DOUBLE)
    b = V.Batch(device, [(cam, s, img) for s, cam, img, _, _ in loaded],
                autotune=False, **kw)
    b.run()
    got = _images(loaded)
    O = oracle()
    for k in (1, 3):
        w, h = loaded[k][3], loaded[k][4]
        os_, ocam = load_oracle(descs[k], sah=True)
        cfg = SpoConfig(rng=SPO_RNG_PCG, tx=SPO_TX_CR, hit=SPO_HIT_REFERENCE,
                        estimator=V.EST_UNBIASED, samplesPerPixel=2, bounces=5,
                        radianceClamp=0.0, jitterYFirst=1, seed=0x5EED)
        for t in ((0, 0, 32, 32), (w // 2 - 16, h // 2 - 16, w // 2 + 16, h // 2 + 16)):
            px = np.zeros(w * h * 4, np.float32)
            assert O.spo_path_trace_tile(os_, C.byref(ocam), C.byref(cfg), *t, None, px,
                                         np.zeros(12, np.uint64)) == 0
            a = got[k].reshape(h, w, 4)[t[1]:t[3], t[0]:t[2]]
            r = px.reshape(h, w, 4)[t[1]:t[3], t[0]:t[2]]
            assert np.array_equal(a.view(np.uint32), r.view(np.uint32)), (k, t)
        O.spo_scene_destroy(os_)
    b.close()
    _close(loaded)


def test_full_config3_batch(device):
    """The full batched workload: 64 seeded scenes of 10k-2M triangles,
    images 128^2..1024^2, 16 spp, 5 bounces, one autotuned graph.  Every
    pixel finite, every path traced once per sample, every ray a hit or a
    miss, and the largest item equal to its single-scene render."""
    descs = scenes.config3()
    loaded = _load(device, descs)
    kw = dict(spp=16, bounces=5, clamp=0.0, estimator=V.EST_UNBIASED, rng=V.RNG_PCG,
              seed=0x5EED, transcendentals=V.TX_FAST)
    b = V.Batch(device, [(cam, s, img) for s, cam, img, _, _ in loaded],
                autotune=True, **kw)
    t = b.tuning()
    pixels = sum(w * h for *_, w, h in loaded)
    assert t["pixels"] == pixels
    m = b.run()
    assert m.values[V.METRIC_PATHS] == pixels * 16
    rays = m.values[V.METRIC_RAYS]
    assert rays == m.values[V.METRIC_HITS] + m.values[V.METRIC_MISSES]
    assert pixels * 16 <= rays <= pixels * 16 * 5
    got = _images(loaded)
    assert all(np.isfinite(g).all() for g in got)
    assert all(g.reshape(-1, 4)[:, :3].mean() > 0 for g in got)
    k = int(np.argmax([w * h for *_, w, h in loaded]))
    s, cam, _, w, h = loaded[k]
    want = _single(device, s, cam, w, h, **kw)
    assert np.array_equal(got[k].view(np.uint32), want.view(np.uint32))
    b.close()
    _close(loaded)


def test_nee_batch_items_equal_single_scene_renders(device):
    """Next-event estimation in the batch (shadow rays of every item through
    the batch variant of the extension kernel): each item equals its
    single-scene NEE render bit for bit."""
    descs = scenes.config3(count=3, seed=17, min_side=48, max_side=120,
                           min_tris=2e4, max_tris=1e5)
    loaded = _load(device, descs)
    kw = dict(spp=2, bounces=5, clamp=0.0, estimator=V.EST_NEE, rng=V.RNG_PCG,
              seed=0x5EED, transcendentals=V.TX_DOUBLE)
    b = V.Batch(device, [(cam, s, img) for s, cam, img, _, _ in loaded],
                autotune=False, **kw)
    m = b.run()
    got = _images(loaded)
    for (s, cam, _, w, h), g in zip(loaded, got):
        want = _single(device, s, cam, w, h, **kw)
        assert np.array_equal(g.view(np.uint32), want.view(np.uint32))
    # the rays counter includes the shadow rays; hits and misses count path rays
    paths = sum(w * h for *_, w, h in loaded) * 2
    path_rays = m.values[V.METRIC_HITS] + m.values[V.METRIC_MISSES]
    assert paths <= path_rays < m.values[V.METRIC_RAYS] <= 2 * path_rays
    b.close()
    _close(loaded)


def _guided_setup(device, w, h):
    desc = scenes.config1(w, h, env="cube", env_size=32)
    s, cam = scenes.load_gpu(device, desc)
    return desc, s, cam


def test_guided_batch_frozen_equals_eager(device):
    """The guided estimator in the graph-captured batch with a frozen guide
    (inference every bounce, no training): two cameras on one scene, each
    item equal to the eager guided render bit for bit."""
    desc, s, cam = _guided_setup(device, 96, 64)
    cam2 = scenes.sp_ConfigureCamera(64, 80, (1.0, 1.5, 4.0), (0, 0, 0, 1), 0.8)
    g = V.Guide(device, seed=3, batch_size=4096, steps_per_pass=2)
    for k in range(4):  # a few training passes so the mixture is not the initial one
        V.path_trace_tiles(V.make_context(cam, s, spp=1, bounces=4, clamp=0.0,
                                          estimator=V.EST_GUIDED, rng=V.RNG_PCG,
                                          seed=9, sample_offset=k, guide=g, train=True),
                           V.ComputeTiles(96, 64, 64, 64), np.zeros(96 * 64 * 4, np.float32))
    kw = dict(spp=3, bounces=4, clamp=0.0, estimator=V.EST_GUIDED, rng=V.RNG_PCG,
              seed=0x77, transcendentals=V.TX_DOUBLE, guide=g)
    imgs = [device.alloc(96 * 64 * 16), device.alloc(64 * 80 * 16)]
    b = V.Batch(device, [(cam, s, imgs[0]), (cam2, s, imgs[1])], autotune=False, **kw)
    b.run()
    for c, img, w, h in ((cam, imgs[0], 96, 64), (cam2, imgs[1], 64, 80)):
        got = img.download(np.float32, w * h * 4)
        want = _single(device, s, c, w, h, **kw)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert np.isfinite(got).all()
    # the guide is pinned while the batch lives
    with pytest.raises(V.SpgError):
        V.check(V.lib().spg_guide_destroy(g.handle))
    b.close()
    for img in imgs:
        img.free()
    g.close()
    s.close()


def test_guided_batch_training_replays_equal_eager_training(device):
    """Online training inside the captured graph: each replay is one guided
    pass with its sample collection and Adam steps.  Replaying it R times
    leaves the guide bit-identical to R eager training traces of the same
    slots (one tile over the image, sample offset 0), and the images of the
    last pass equal."""
    desc, s, cam = _guided_setup(device, 96, 64)
    w, h = 96, 64
    kw = dict(spp=1, bounces=4, clamp=0.0, estimator=V.EST_GUIDED, rng=V.RNG_PCG,
              seed=0x99, transcendentals=V.TX_FAST)
    ga = V.Guide(device, seed=5, batch_size=2048, steps_per_pass=2)
    gb = V.Guide(device, seed=5, batch_size=2048, steps_per_pass=2)
    img = device.alloc(w * h * 16)
    b = V.Batch(device, [(cam, s, img)], autotune=False, guide=ga, train=True, **kw)
    assert ga.stats()["steps"] == 0       # creating the batch did not train
    eager = device.alloc(w * h * 16)
    for _ in range(5):
        b.run()
        V.path_trace_tiles_device(V.make_context(cam, s, guide=gb, train=True, **kw),
                                  [V.Tile(0, 0, w, h)], eager)
    assert ga.stats()["steps"] == gb.stats()["steps"] == 10
    assert np.array_equal(ga.params().view(np.uint32), gb.params().view(np.uint32))
    a = img.download(np.float32, w * h * 4)
    e = eager.download(np.float32, w * h * 4)
    assert np.array_equal(a.view(np.uint32), e.view(np.uint32))
    b.close()
    for x in (img, eager):
        x.free()
    ga.close()
    gb.close()
    s.close()


def test_guided_batch_needs_one_scene(device):
    d1, s1, c1 = _guided_setup(device, 32, 32)
    d2, s2, c2 = _guided_setup(device, 32, 32)
    g = V.Guide(device, seed=1)
    img = [device.alloc(32 * 32 * 16) for _ in range(2)]
    with pytest.raises(V.SpgError):
        V.Batch(device, [(c1, s1, img[0]), (c2, s2, img[1])], estimator=V.EST_GUIDED,
                guide=g, autotune=False)
    for x in img:
        x.free()
    g.close()
    s1.close()
    s2.close()
