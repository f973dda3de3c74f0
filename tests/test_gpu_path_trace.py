"""Wavefront path tracer on the GPU against the oracle and the reference.

* Against the oracle with the same per-pixel random streams and
  double-evaluated transcendentals: every pixel bit-exact, counters equal.
* Against the reference's own output (golden vectors, libm
  transcendentals): per-pixel agreement within a stated tolerance.
* sp_PathTraceTile semantics: only the tile's pixels are written; magenta
  for a missing material; metrics.
"""
import ctypes as C

import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes
from _libs import (SPO_HIT_BRUTE, SPO_HIT_REFERENCE, SPO_RNG_PCG,
                   SPO_RNG_XORSHIFT_PIXEL, SPO_TX_CR, SpoConfig, oracle)
from _scenes import load_oracle

pytestmark = pytest.mark.gpu

MODES = [(V.RNG_XORSHIFT_PIXEL, SPO_RNG_XORSHIFT_PIXEL, V.EST_PARITY),
         (V.RNG_PCG, SPO_RNG_PCG, V.EST_PARITY),
         (V.RNG_XORSHIFT_PIXEL, SPO_RNG_XORSHIFT_PIXEL, V.EST_UNBIASED),
         (V.RNG_PCG, SPO_RNG_PCG, V.EST_UNBIASED)]


def gpu_render(device, desc, spp, bounces, est, rng, seed=0x5EED, clamp=10.0,
               tiles=None, collect=True):
    s, cam = scenes.load_gpu(device, desc)
    w, h = desc["camera"][0], desc["camera"][1]
    tiles = tiles or V.ComputeTiles(w, h, 64, 64)
    ctx = V.make_context(cam, s, spp=spp, bounces=bounces, clamp=clamp, estimator=est,
                         rng=rng, seed=seed, collect=collect)
    img = np.zeros(w * h * 4, np.float32)
    m = V.path_trace_tiles(ctx, tiles, img)
    return img, m


def oracle_render(desc, spp, bounces, est, orng, seed=0x5EED, clamp=10.0):
    O = oracle()
    s, cam = load_oracle(desc)
    w, h = desc["camera"][0], desc["camera"][1]
    cfg = SpoConfig(rng=orng, tx=SPO_TX_CR, hit=SPO_HIT_BRUTE, estimator=est,
                    samplesPerPixel=spp, bounces=bounces, radianceClamp=clamp,
                    jitterYFirst=1, seed=seed)
    px = np.zeros(w * h * 4, np.float32)
    met = np.zeros(12, np.uint64)
    assert O.spo_path_trace_tile(s, C.byref(cam), C.byref(cfg), 0, 0, w, h, None, px,
                                 met) == 0
    O.spo_scene_destroy(s)
    return px, met


@pytest.mark.parametrize("rng,orng,est", MODES)
@pytest.mark.parametrize("env", ["cube", "equirect"])
def test_config1_matches_oracle_bit_exact(device, rng, orng, est, env):
    desc = scenes.config1(64, 48, env=env, env_size=32)
    img, m = gpu_render(device, desc, 4, 3, est, rng)
    px, met = oracle_render(desc, 4, 3, est, orng)
    assert np.array_equal(img.view(np.uint32), px.view(np.uint32))
    assert [m.values[k] for k in (1, 2, 3, 4)] == [int(met[k]) for k in (1, 2, 3, 4)]


def test_config1_full_size_matches_oracle(device):
    """The config-1 workload at its full size: 256x256, 4 samples, 3 bounces,
    per-pixel PCG streams, parity estimator."""
    desc = scenes.config1(256, 256, env="cube")
    img, m = gpu_render(device, desc, 4, 3, V.EST_PARITY, V.RNG_PCG)
    px, met = oracle_render(desc, 4, 3, V.EST_PARITY, SPO_RNG_PCG)
    same = (img.view(np.uint32) == px.view(np.uint32)).reshape(-1, 4).all(1)
    assert same.all(), f"{(~same).sum()} pixels differ"
    assert m.values[1] == 256 * 256 * 4
    assert [m.values[k] for k in (2, 3, 4)] == [int(met[k]) for k in (2, 3, 4)]


def test_config1_against_reference_golden(device, golden):
    """The reference's own output (libm transcendentals, first-visited tie
    rule) against the GPU (double-evaluated transcendentals, lower-index tie
    rule) at equal per-pixel xorshift streams.  Tolerance: the two runs differ
    by the last bit of a few sin/cos/atan2/pow results, which moves a path by
    an ulp; at least 90% of the pixels are bit-identical, 99% agree to 1e-4
    relative, and every pixel to 5% relative or 1e-3 absolute unless its
    path changed what it hit (at most 1% of pixels)."""
    g = golden("c1_render")
    W, H = int(g["width"]), int(g["height"])
    desc = scenes.config1(W, H, env="equirect", env_size=32)
    for spp in (1, 2):
        img, m = gpu_render(device, desc, spp, 3, V.EST_PARITY, V.RNG_XORSHIFT_PIXEL,
                            seed=int(g["seed_base"]))
        ref = g[f"pixels_spp{spp}"]
        a = img.reshape(-1, 4)[:, :3].astype(np.float64)
        b = ref.reshape(-1, 4)[:, :3].astype(np.float64)
        exact = (img.view(np.uint32) == ref.view(np.uint32)).reshape(-1, 4).all(1)
        rel = (np.abs(a - b) / np.maximum(np.abs(b), 1e-3)).max(1)
        print(f"spp {spp}: exact {exact.mean():.4f}, rel<1e-4 {(rel < 1e-4).mean():.4f}, "
              f"max rel {rel.max():.3e}")
        assert exact.mean() >= 0.90
        assert (rel < 1e-4).mean() >= 0.99
        assert (rel > 0.05).mean() <= 0.01
        # the counters see the same paths
        refm = g[f"metrics_spp{spp}"]
        for k in (1, 2, 3, 4):
            assert abs(int(m.values[k]) - int(refm[k])) <= 0.01 * int(refm[k]) + 2


def test_scene_intersect_against_reference_golden(device, golden):
    """sp_RayIntersectScene through the GPU traversal (no transcendentals on
    this path): t, material, normal and uv bit-exact against the reference."""
    g = golden("scene_rays")
    desc = scenes.config1(32, 32, env="equirect", env_size=32)
    s, _ = scenes.load_gpu(device, desc)
    r = s.intersect_scene(g["o"], g["d"])
    out = g["out"]
    got = np.concatenate([r["t"][:, None], r["material"].view(np.float32)[:, None],
                          r["normal"], r["uv"]], 1)
    hit = out[:, 0] > 0
    assert np.array_equal((r["t"] > 0), hit)
    assert np.array_equal(got[hit].view(np.uint32), out[hit].view(np.uint32))


def test_ray_intersect_scene_two_objects(device):
    """TestRayIntersectScene, unit_tests/test_simd_path_tracer.cpp:216-276"""
    s = V.Scene(device)
    v = np.array([[-.5, -.5, 0, 0, 0, 1, 0, 0], [.5, -.5, 0, 0, 0, 1, 0, 0],
                  [0, .5, 0, 0, 0, 1, 0, 0]], np.float32)
    m0 = s.create_mesh(v, [0, 1, 2], smooth=False)
    m1 = s.create_mesh(v, [0, 1, 2], smooth=False)
    q = scenes.quat_axis_angle((0, 1, 0), np.float32(3.14159265359) * np.float32(0.25))
    s.add_object(m0, 53, (0, 2, -5))
    s.add_object(m1, 53, (0, 2, -15), q, (2, 2, 2))
    s.build()
    r = s.intersect_scene([[0, 2, 0]], [[0, 0, -1]])
    assert r["t"][0] >= 0 and r["material"][0] == 53 and r["object"][0] == 0


def test_tile_writes_only_its_pixels_and_magenta(device):
    """TestPathTraceSingleColor / TestPathTraceTile / TestMetrics,
    unit_tests/test_simd_path_tracer.cpp:49-135 and :333-368: an empty scene
    with no background material gives magenta, only inside the tile."""
    s = V.Scene(device)
    s.build()
    cam = V.sp_ConfigureCamera(4, 4, (0, 0, 0), (0, 0, 0, 1), 1.0)
    ctx = V.make_context(cam, s, spp=1, bounces=3, rng=V.RNG_XORSHIFT_PIXEL)
    img = np.full(4 * 4 * 4, 7.0, np.float32)
    m = V.sp_PathTraceTile(ctx, V.Tile(1, 1, 3, 3), img)
    px = img.reshape(4, 4, 4)
    inside = np.zeros((4, 4), bool)
    inside[1:3, 1:3] = True
    assert np.all(px[inside] == np.array([1, 0, 1, 1], np.float32))
    assert np.all(px[~inside] == 7.0)
    assert m.values[V.METRIC_ELAPSED_NS] > 0 and m.values[V.METRIC_PATHS] == 4
    assert m.values[V.METRIC_RAYS] == 4 and m.values[V.METRIC_MISSES] == 4
    # the whole image
    img = np.zeros(64, np.float32)
    V.sp_PathTraceTile(ctx, V.Tile(0, 0, 4, 4), img)
    assert np.all(img.reshape(-1, 4) == np.array([1, 0, 1, 1], np.float32))
    # tiles past the image edge are clamped (src/simd_path_tracer.cpp:186-189)
    img = np.zeros(64, np.float32)
    V.sp_PathTraceTile(ctx, V.Tile(2, 2, 64, 64), img)
    assert np.all(img.reshape(4, 4, 4)[2:, 2:] == np.array([1, 0, 1, 1], np.float32))
    assert np.all(img.reshape(4, 4, 4)[:2] == 0)


def test_errors_instead_of_aborts(device):
    s = V.Scene(device)
    cam = V.sp_ConfigureCamera(4, 4, (0, 0, 0), (0, 0, 0, 1), 1.0)
    ctx = V.make_context(cam, s, spp=1, bounces=3)
    img = np.zeros(64, np.float32)
    with pytest.raises(V.SpgError) as e:
        V.sp_PathTraceTile(ctx, V.Tile(0, 0, 4, 4), img)
    assert e.value.code == -4  # not built
    s.build()
    ctx = V.make_context(cam, s, spp=1, bounces=0)
    with pytest.raises(V.SpgError):
        V.sp_PathTraceTile(ctx, V.Tile(0, 0, 4, 4), img)
    with pytest.raises(V.SpgError):
        s.add_object(5, 0, (0, 0, 0))


def test_unbiased_furnace_and_constant_sky(device):
    """Closed forms of the unbiased estimator on the GPU.  (1) Under a
    constant sky, a black convex object: every sample is exactly 0 or L_e.
    (2) Inside a closed sphere of albedo a and emission E the mean lies
    between E and E(1-q^N)/(1-q), q = a(1-0.04), and near the upper bound."""
    v, i = V.gen_icosphere(3)
    desc = {"meshes": [(v, i, True)],
            "objects": [(0, 1, (0, 0, 0), (0, 0, 0, 1), (1, 1, 1))],
            "materials": [(1, (0, 0, 0), (0, 0, 0), 0.0, V.SPG_NO_TEXTURE,
                           V.SPG_NO_TEXTURE),
                          (9, (0, 0, 0), (2, 2, 2), 0.0, V.SPG_NO_TEXTURE,
                           V.SPG_NO_TEXTURE)],
            "background": 9, "camera": (32, 32, (0, 0, 3), (0, 0, 0, 1), 0.8)}
    img, _ = gpu_render(device, desc, 4, 3, V.EST_UNBIASED, V.RNG_PCG, clamp=0.0)
    r = img.reshape(-1, 4)[:, 0]
    assert np.all(np.isin(r, np.float32([0, .5, 1, 1.5, 2])))
    vin = v.copy()
    vin[:, 3:6] *= -1
    iin = i.reshape(-1, 3)[:, ::-1].ravel().copy()
    a, E, N = 0.5, 1.0, 4
    desc = {"meshes": [(vin, iin, True)],
            "objects": [(0, 1, (0, 0, 0), (0, 0, 0, 1), (5, 5, 5))],
            "materials": [(1, (a, a, a), (E, E, E), 0.0, V.SPG_NO_TEXTURE,
                           V.SPG_NO_TEXTURE)],
            "background": 9, "camera": (32, 32, (0, 0, 0), (0, 0, 0, 1), 0.8)}
    img, _ = gpu_render(device, desc, 64, N, V.EST_UNBIASED, V.RNG_PCG, clamp=0.0)
    mean = img.reshape(-1, 4)[:, 0].mean()
    q = a * (1 - 0.04)
    hi = E * (1 - q ** N) / (1 - q)
    assert E < mean < hi and mean == pytest.approx(hi, rel=0.05)


def test_unbiased_mean_is_stable_across_seeds(device):
    """Independent PCG seeds give image means within their Monte Carlo error
    (the estimator is a proper average, not seed-dependent)."""
    desc = scenes.config1(64, 64, env="cube", env_size=32)
    means = []
    for seed in (1, 2, 3, 4):
        img, _ = gpu_render(device, desc, 16, 3, V.EST_UNBIASED, V.RNG_PCG, seed=seed,
                            clamp=0.0, collect=False)
        means.append(img.reshape(-1, 4)[:, :3].mean())
    means = np.array(means)
    assert means.std() / means.mean() < 0.05


def _oracle_tile_sah(s, cam, desc, tile, spp, bounces, est, seed=0x5EED, clamp=0.0):
    """The oracle over the product's SAH hierarchy, one tile of the image."""
    O = oracle()
    w, h = desc["camera"][0], desc["camera"][1]
    cfg = SpoConfig(rng=SPO_RNG_PCG, tx=SPO_TX_CR, hit=SPO_HIT_REFERENCE,
                    estimator=est, samplesPerPixel=spp, bounces=bounces,
                    radianceClamp=clamp, jitterYFirst=1, seed=seed)
    px = np.zeros(w * h * 4, np.float32)
    met = np.zeros(12, np.uint64)
    assert O.spo_path_trace_tile(s, C.byref(cam), C.byref(cfg), tile.minX, tile.minY,
                                 tile.maxX, tile.maxY, None, px, met) == 0
    return px, met


@pytest.mark.parametrize("name,tiles", [
    ("config4", [(224, 224, 288, 288), (0, 0, 64, 64), (448, 448, 512, 512)]),
    ("config5", [(928, 508, 992, 572), (928, 0, 992, 48), (0, 1040, 64, 1080)]),
])
def test_box_scene_tiles_match_oracle_bit_exact(device, name, tiles):
    """The million-triangle workloads (many rotated, scaled instances, a
    two-level hierarchy, 6 bounces, the plain unbiased tracer): tiles of the
    full-size image equal the oracle's bit for bit, as do the ray, hit and
    miss counts."""
    desc = getattr(scenes, name)()
    w, h = desc["camera"][0], desc["camera"][1]
    s, cam = scenes.load_gpu(device, desc)
    ctx = V.make_context(cam, s, spp=2, bounces=6, clamp=0.0, estimator=V.EST_UNBIASED,
                         rng=V.RNG_PCG, seed=0x5EED, collect=True)
    os_, ocam = load_oracle(desc, sah=True)
    for t in tiles:
        tile = V.Tile(*t)
        img = np.zeros(w * h * 4, np.float32)
        m = V.sp_PathTraceTile(ctx, tile, img)
        px, met = _oracle_tile_sah(os_, ocam, desc, tile, 2, 6, V.EST_UNBIASED)
        a = img.reshape(h, w, 4)[t[1]:t[3], t[0]:t[2]]
        b = px.reshape(h, w, 4)[t[1]:t[3], t[0]:t[2]]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (name, t)
        assert [m.values[k] for k in (2, 3, 4)] == [int(met[k]) for k in (2, 3, 4)]
        assert m.values[V.METRIC_RAYS] > (t[2] - t[0]) * (t[3] - t[1]) * 2
    oracle().spo_scene_destroy(os_)
    s.close()


@pytest.mark.parametrize("est", [V.EST_PARITY, V.EST_UNBIASED, V.EST_GUIDED])
def test_wavefront_conserves_rays(device, est):
    """Queue bookkeeping: every ray traced ends as a hit or a miss, every
    path starts one camera ray, and no path exceeds the bounce count."""
    desc = scenes.box_scene(3, 2, 96, 64, seed=3)
    s, cam = scenes.load_gpu(device, desc)
    g = V.Guide(device, seed=2, batch_size=4096) if est == V.EST_GUIDED else None
    spp, bounces = 3, 5
    ctx = V.make_context(cam, s, spp=spp, bounces=bounces, clamp=0.0, estimator=est,
                         rng=V.RNG_PCG, guide=g, train=g is not None, collect=True)
    img = np.zeros(96 * 64 * 4, np.float32)
    m = V.path_trace_tiles(ctx, V.ComputeTiles(96, 64, 64, 64), img)
    paths = 96 * 64 * spp
    rays, hits, misses = (m.values[k] for k in (V.METRIC_RAYS, V.METRIC_HITS,
                                                 V.METRIC_MISSES))
    assert m.values[V.METRIC_PATHS] == paths
    assert rays == hits + misses
    assert paths <= rays <= paths * bounces
    assert misses <= paths
    assert np.all(np.isfinite(img))
    if g is not None:
        g.close()
    s.close()


def test_slot_list_cache_invalidation(device):
    """The device keeps the last call's tile-to-pixel slot list and reuses
    it for a call with the same tiles and image size.  Interleaved calls
    that change the tiles, return to earlier tiles, keep the tiles but change
    the image size, trace more paths than fit in flight (slot chunks) and
    then fewer again, each equal a fresh render bit for bit."""
    big = scenes.config1(2048, 1088, env="cube", env_size=16)
    small = scenes.config1(256, 192, env="cube", env_size=16)
    s_big, cam_big = scenes.load_gpu(device, big)
    s_small, cam_small = scenes.load_gpu(device, small)
    cam_small2 = V.sp_ConfigureCamera(192, 256, (0.0, 1.0, 4.0), (0, 0, 0, 1), 0.8)

    def render(cam, s, tiles, spp, w, h):
        ctx = V.make_context(cam, s, spp=spp, bounces=3, estimator=V.EST_PARITY,
                             rng=V.RNG_PCG, seed=0x5EED)
        img = np.zeros(w * h * 4, np.float32)
        V.path_trace_tiles(ctx, tiles, img)
        return img

    t1 = V.ComputeTiles(256, 192, 64, 64)
    t2 = V.ComputeTiles(256, 192, 32, 32)[5:40]
    t_all_big = V.ComputeTiles(2048, 1088, 64, 64)
    seq = [(cam_small, s_small, t1, 2, 256, 192), (cam_small, s_small, t2, 2, 256, 192),
           (cam_small, s_small, t1, 2, 256, 192), (cam_small2, s_small, t1, 2, 192, 256),
           (cam_big, s_big, t_all_big, 1, 2048, 1088),   # 2.2M slots: two chunks
           (cam_small, s_small, t1, 2, 256, 192), (cam_small, s_small, t1, 2, 256, 192)]
    results = [render(*a) for a in seq]
    fresh = {}
    for a, r in zip(seq, results):
        key = (id(a[0]), id(a[2]))
        if key not in fresh:
            # a fresh render: a different device, so no cache is shared
            d2 = V.Device(0)
            s2, _ = scenes.load_gpu(d2, big if a[0] is cam_big else small)
            fresh[key] = render(a[0], s2, a[2], a[3], a[4], a[5])
            s2.close()
            d2.close()
        assert np.array_equal(r.view(np.uint32), fresh[key].view(np.uint32))
    s_big.close()
    s_small.close()
