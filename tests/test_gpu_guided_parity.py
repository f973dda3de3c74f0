"""The guided estimator on the GPU against the oracle's restatement of it.

* Verification mixture (spg_guide_debug SPG_GUIDE_ANALYTIC_MIXTURE, no
  network): the guided scatter -- sampler choice, vMF draw, one-sample MIS
  density 0.5 p_bsdf + 0.5 q, the (E, X, cos/pdf) vertex records and the
  reverse recurrence -- equals the oracle's (spo_trace.c trace_guided) bit
  for bit with double-evaluated transcendentals, and so do the training
  samples a pass collects (weights Y(f cos Li / pdf)).
* The network's mixture (recorded scatter events of trained-guide traces):
  the lobes k_guide_infer produced equal k_guide_eval's on the same inputs
  bit for bit and the FP64 oracle network's within FP16 rounding, and the
  sampled direction and density equal spo_guided_scatter's for the
  recorded lobes and uniforms bit for bit.
* The full-size guided workloads (config 4, config 5) with a trained
  network: finite images, ray conservation, and means equal to the plain
  tracer's within the Monte Carlo error.
"""
import ctypes as C

import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes
from _libs import (SPO_EST_GUIDED, SPO_HIT_BRUTE, SPO_HIT_REFERENCE, SPO_RNG_PCG,
                   SPO_RNG_XORSHIFT_PIXEL, SPO_TX_CR, SpoConfig, SpoLobes, oracle)
from _guide import oracle_forward
from _scenes import load_oracle

pytestmark = pytest.mark.gpu


def _oracle_guided(desc, spp, bounces, orng, seed, hit=SPO_HIT_BRUTE, sah=False,
                   tile=None, sink=0, est=SPO_EST_GUIDED, light_split=None):
    O = oracle()
    s, cam = load_oracle(desc, sah=sah)
    if light_split is not None:
        O.spo_scene_set_light_split(s, light_split)
    w, h = desc["camera"][0], desc["camera"][1]
    cfg = SpoConfig(rng=orng, tx=SPO_TX_CR, hit=hit, estimator=est,
                    samplesPerPixel=spp, bounces=bounces, radianceClamp=0.0,
                    jitterYFirst=1, seed=seed, bsdfFraction=0.5)
    px = np.zeros(w * h * 4, np.float32)
    met = np.zeros(12, np.uint64)
    rows = np.zeros((max(sink, 1), 16), np.float32)
    O.spo_guided_sink(rows.ctypes.data if sink else None, sink)
    t = tile or (0, 0, w, h)
    assert O.spo_path_trace_tile(s, C.byref(cam), C.byref(cfg), *t, None, px, met) == 0
    n = O.spo_guided_sink_count()
    O.spo_guided_sink(None, 0)
    O.spo_scene_destroy(s)
    return px, met, rows[:n]


@pytest.mark.parametrize("rng,orng", [(V.RNG_PCG, SPO_RNG_PCG),
                                      (V.RNG_XORSHIFT_PIXEL, SPO_RNG_XORSHIFT_PIXEL)])
@pytest.mark.parametrize("bounces", [3, 6])
def test_guided_scatter_bit_exact_with_verification_mixture(device, rng, orng, bounces):
    desc = scenes.config1(64, 48, env="cube", env_size=32)
    s, cam = scenes.load_gpu(device, desc)
    g = V.Guide(device, seed=1)
    g.debug(analytic=True)
    ctx = V.make_context(cam, s, spp=4, bounces=bounces, clamp=0.0,
                         estimator=V.EST_GUIDED, rng=rng, seed=0x5EED, guide=g,
                         collect=True)
    img = np.zeros(64 * 48 * 4, np.float32)
    m = V.path_trace_tiles(ctx, V.ComputeTiles(64, 48, 64, 64), img)
    px, met, _ = _oracle_guided(desc, 4, bounces, orng, 0x5EED)
    assert np.array_equal(img.view(np.uint32), px.view(np.uint32))
    assert [m.values[k] for k in (2, 3, 4)] == [int(met[k]) for k in (2, 3, 4)]
    assert np.isfinite(img).all() and img[0::4].mean() > 0
    g.close()
    s.close()


def test_guided_training_samples_bit_exact_with_verification_mixture(device):
    """One training pass small enough that every path is selected: the
    collected samples are exactly the oracle's candidate samples (pn from
    the hit point and the guide's scene bounds, n, v, wi, weight), each
    weight bit for bit."""
    W, H = 48, 32
    desc = scenes.config1(W, H, env="cube", env_size=32)
    s, cam = scenes.load_gpu(device, desc)
    g = V.Guide(device, seed=1, batch_size=16384, steps_per_pass=4)
    g.debug(analytic=True)
    ctx = V.make_context(cam, s, spp=2, bounces=4, clamp=0.0, estimator=V.EST_GUIDED,
                         rng=V.RNG_PCG, seed=0x77, guide=g, train=True)
    img = np.zeros(W * H * 4, np.float32)
    V.path_trace_tiles(ctx, V.ComputeTiles(W, H, 64, 64), img)
    got = g.samples()
    px, _, cand = _oracle_guided(desc, 2, 4, SPO_RNG_PCG, 0x77, sink=100000)
    assert np.array_equal(img.view(np.uint32), px.view(np.uint32))
    assert 500 < len(cand) < 4 * 16384 * 4
    assert len(got) == len(cand)
    # match by (n, v, wi) bits
    key = lambda r: r.view(np.uint32).tobytes()  # noqa: E731
    want = {key(r[3:12]): r for r in cand}
    assert len(want) == len(cand)
    P = np.zeros((len(got), 3))
    for j, r in enumerate(got):
        c = want[key(r[3:12])]
        assert r[12].view(np.uint32) == c[12].view(np.uint32)
        P[j] = c[0:3]
    # pn is the hit point mapped affinely into the scene bounds, per axis
    for a in range(3):
        A = np.stack([P[:, a], np.ones(len(P))], 1)
        coef, *_ = np.linalg.lstsq(A, got[:, a].astype(np.float64), rcond=None)
        assert coef[0] > 0
        assert np.max(np.abs(A @ coef - got[:, a])) < 2e-6
    g.close()
    s.close()


@pytest.mark.parametrize("name,tiles", [
    ("config4", [(224, 224, 288, 288), (448, 0, 512, 64)]),
    ("config5", [(928, 508, 992, 572), (0, 1040, 64, 1080)]),
])
def test_box_scene_guided_tiles_bit_exact(device, name, tiles):
    """Tiles of the full-size million-triangle workloads, 6 bounces, the
    guided estimator with the verification mixture, against the oracle over
    the same SAH hierarchy: bit for bit, equal ray, hit and miss counts."""
    desc = getattr(scenes, name)()
    w, h = desc["camera"][0], desc["camera"][1]
    s, cam = scenes.load_gpu(device, desc)
    g = V.Guide(device, seed=1)
    g.debug(analytic=True)
    ctx = V.make_context(cam, s, spp=2, bounces=6, clamp=0.0, estimator=V.EST_GUIDED,
                         rng=V.RNG_PCG, seed=0x5EED, guide=g, collect=True)
    for t in tiles:
        img = np.zeros(w * h * 4, np.float32)
        m = V.sp_PathTraceTile(ctx, V.Tile(*t), img)
        px, met, _ = _oracle_guided(desc, 2, 6, SPO_RNG_PCG, 0x5EED, hit=SPO_HIT_REFERENCE,
                                    sah=True, tile=t)
        a = img.reshape(h, w, 4)[t[1]:t[3], t[0]:t[2]]
        b = px.reshape(h, w, 4)[t[1]:t[3], t[0]:t[2]]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (name, t)
        assert [m.values[k] for k in (2, 3, 4)] == [int(met[k]) for k in (2, 3, 4)]
    g.close()
    s.close()


def test_network_scatter_records_match_eval_and_oracle(device):
    """Scatter events recorded from traces with a trained network (double-
    evaluated transcendentals): the lobes k_guide_infer wrote equal
    k_guide_eval's for the recorded (pn, n, v) bit for bit and the FP64
    oracle network's within FP16 rounding; the direction and density equal
    spo_guided_scatter's for the recorded lobes and uniforms bit for bit."""
    desc = scenes.config1(64, 64, env="cube", env_size=32)
    s, cam = scenes.load_gpu(device, desc)
    g = V.Guide(device, seed=4, batch_size=4096, steps_per_pass=2)
    tiles = V.ComputeTiles(64, 64, 64, 64)
    img = np.zeros(64 * 64 * 4, np.float32)
    for k in range(8):
        V.path_trace_tiles(V.make_context(cam, s, spp=1, bounces=4, clamp=0.0,
                                          estimator=V.EST_GUIDED, rng=V.RNG_PCG,
                                          seed=3, sample_offset=k, guide=g, train=True),
                           tiles, img)
    g.debug(record=4000)
    V.path_trace_tiles(V.make_context(cam, s, spp=2, bounces=4, clamp=0.0,
                                      estimator=V.EST_GUIDED, rng=V.RNG_PCG, seed=9,
                                      guide=g), tiles, img)
    rec = g.records()
    g.debug(record=0)
    assert len(rec) == 4000
    x9 = np.ascontiguousarray(rec[:, 0:9])
    _, lobes = g.evaluate(x9)
    assert np.array_equal(lobes.view(np.uint32), rec[:, 13:53].view(np.uint32))
    # against the FP64 network on a subset
    O = oracle()
    p = g.params().astype(np.float64)
    out64 = oracle_forward(p, x9[:300])
    for j in range(0, 300, 7):
        L = SpoLobes()
        O.spo_guide_lobes(np.ascontiguousarray(out64[j]), L)
        assert rec[j, 13:21] == pytest.approx(list(L.pi), abs=2e-2)
    # the scatter of every record
    n_guide = 0
    for r in rec:
        wi = np.zeros(3, np.float32)
        pdf = C.c_float()
        O.spo_guided_scatter(np.ascontiguousarray(r[13:53]), np.ascontiguousarray(r[3:6]),
                             np.ascontiguousarray(r[6:9]), float(r[9]), 0.5, float(r[10]),
                             float(r[11]), float(r[12]), SPO_TX_CR, wi, C.byref(pdf))
        assert np.array_equal(wi.view(np.uint32), r[53:56].view(np.uint32))
        assert np.float32(pdf.value).view(np.uint32) == r[56].view(np.uint32)
        n_guide += r[10] >= 0.5
    assert 1000 < n_guide < 3000   # about half the events sample the mixture
    g.close()
    s.close()


@pytest.mark.parametrize("name", ["config4", "config5"])
def test_full_size_guided_workload_properties(device, name):
    """The full-size guided workloads with an online-trained network: every
    pixel finite, every ray ends as a hit or a miss, the image mean equals
    the plain tracer's within the Monte Carlo error."""
    desc = getattr(scenes, name)()
    w, h = desc["camera"][0], desc["camera"][1]
    s, cam = scenes.load_gpu(device, desc)
    tiles = V.ComputeTiles(w, h, 64, 64)
    dimg = device.alloc(w * h * 16)
    g = V.Guide(device, seed=1, max_weight=10.0)
    for k in range(12):
        V.path_trace_tiles_device(V.make_context(cam, s, spp=1, bounces=6, clamp=0.0,
                                                 estimator=V.EST_GUIDED, rng=V.RNG_PCG,
                                                 seed=5, sample_offset=k, guide=g,
                                                 train=True, transcendentals=V.TX_FAST),
                                  tiles, dimg)
    assert g.stats()["steps"] == 48

    def means(est, seed, reps, spp):
        out = []
        for k in range(reps):
            ctx = V.make_context(cam, s, spp=spp, bounces=6, clamp=0.0, estimator=est,
                                 rng=V.RNG_PCG, seed=seed + k, guide=g,
                                 transcendentals=V.TX_FAST, collect=True)
            m = V.path_trace_tiles_device(ctx, tiles, dimg)
            rays, hits, misses = (m.values[i] for i in (V.METRIC_RAYS, V.METRIC_HITS,
                                                        V.METRIC_MISSES))
            assert rays == hits + misses and misses <= w * h * spp
            img = dimg.download(np.float32, w * h * 4).reshape(h, w, 4)[:, :, :3]
            assert np.isfinite(img).all()
            out.append(img.mean())
        return np.array(out)

    spp = 4 if name == "config4" else 2
    gm = means(V.EST_GUIDED, 100, 4, spp)
    pm = means(V.EST_UNBIASED, 200, 4, spp)
    se = np.sqrt(gm.var(ddof=1) / 4 + pm.var(ddof=1) / 4)
    print(f"\n{name}: guided {gm.mean():.5f} +- {gm.std():.5f}, plain {pm.mean():.5f} "
          f"+- {pm.std():.5f}")
    assert abs(gm.mean() - pm.mean()) < 5 * se + 2e-3 * pm.mean()
    g.close()
    dimg.free()
    s.close()
