"""Next-event estimation toward a cube-map environment (SPG_EST_NEE; no
reference counterpart): the importance-sampling table, its sampler and
density, and the oracle's estimator against closed forms and against the
plain unbiased estimator (CPU only)."""
import ctypes as C
import math

import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes
from _libs import (SPO_EST_NEE, SPO_EST_UNBIASED, SPO_HIT_BRUTE, SPO_RNG_PCG,
                   SPO_TX_CR, SpoCamera, SpoConfig, f32, oracle)
from _scenes import load_oracle


def _tables(cube, size):
    O = oracle()
    n = 6 * size * size
    a = [np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.uint32)]
    assert O.spo_env_table_build(np.ascontiguousarray(cube, np.float32), size, size,
                                 *a) == 0
    return a


def _rects(size):
    """Face-plane rectangles of the lookup's texels (face-major, rows,
    columns), from the lookup's own rounding boundaries."""
    i = np.arange(size)
    x0 = np.where(i == 0, -1.0, (2 * i - 1) / size - 1.0)
    x1 = np.where(i == size - 1, 1.0, (2 * i + 1) / size - 1.0)
    y1 = np.where(i == 0, 1.0, 1.0 - (2 * i - 1) / size)
    y0 = np.where(i == size - 1, -1.0, 1.0 - (2 * i + 1) / size)
    X0, Y0 = np.meshgrid(x0, y0)
    X1, Y1 = np.meshgrid(x1, y1)
    area = ((X1 - X0) * (Y1 - Y0)).ravel()
    return np.tile(area, 6)


@pytest.mark.parametrize("kind", ["sky", "random"])
def test_table_product_equals_oracle_and_is_a_distribution(kind):
    """The product's table equals the oracle's bit for bit; the texel
    probabilities P_i = pdf_i * area_i sum to 1, and the alias table
    reproduces them."""
    size = 32
    if kind == "sky":
        cube = V.gen_sky_cubemap(size, scenes.SUN_DIR, scenes.SUN_RADIUS,
                                 scenes.SUN_RADIANCE)
    else:
        cube = np.random.default_rng(3).exponential(1.0, 6 * size * size * 4)
        cube = cube.astype(np.float32)
    pdf, prob, alias = V.env_table_build(cube, size)
    opdf, oprob, oalias = _tables(cube, size)
    assert np.array_equal(pdf.view(np.uint32), opdf.view(np.uint32))
    assert np.array_equal(prob.view(np.uint32), oprob.view(np.uint32))
    assert np.array_equal(alias, oalias)
    n = pdf.size
    P = pdf.astype(np.float64) * _rects(size)
    assert P.sum() == pytest.approx(1.0, abs=1e-5)
    rec = prob.astype(np.float64).copy()
    np.add.at(rec, alias, 1.0 - prob.astype(np.float64))
    rec -= np.where(alias == np.arange(n), 1.0 - prob, 0.0)
    assert np.allclose(rec / n, P, atol=2e-7, rtol=1e-5)


def test_density_of_a_constant_map_is_uniform_on_the_sphere():
    """Constant environment: the density is 1/(4 pi) up to the variation of
    (1 + fx^2 + fy^2)^(3/2) across a texel (within 15% at 32x32 faces),
    and its average over uniform directions is 1/(4 pi) (the density
    integrates to 1)."""
    O = oracle()
    size = 32
    cube = np.ones(6 * size * size * 4, np.float32)
    pdf, _, _ = _tables(cube, size)
    rng = np.random.default_rng(5)
    d = rng.normal(size=(200000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    p = np.array([O.spo_env_pdf(pdf, size, size, np.ascontiguousarray(x)) for x in d[:20000]])
    assert np.abs(p * 4 * math.pi - 1).max() < 0.15
    assert (p * 4 * math.pi).mean() == pytest.approx(1.0, abs=2e-3)


def test_sampler_follows_the_table():
    """Directions drawn by the sampler land in texel i with frequency P_i
    (chi-square over the texels of the sun and a band of sky), and the
    density the estimator evaluates at a drawn direction equals pdf_i times
    (1 + fx^2 + fy^2)^(3/2) of its texel."""
    O = oracle()
    size = 16
    cube = V.gen_sky_cubemap(size, scenes.SUN_DIR, scenes.SUN_RADIUS, scenes.SUN_RADIANCE)
    pdf, prob, alias = _tables(cube, size)
    P = pdf.astype(np.float64) * _rects(size)
    rng = np.random.default_rng(7)
    N = 200000
    r = rng.integers(0, 2 ** 32, N, dtype=np.uint64).astype(np.uint32)
    e = (rng.integers(0, 2 ** 24, (N, 3)) / 2.0 ** 24).astype(np.float32)
    out = np.zeros(3, np.float32)
    counts = np.zeros(P.size)
    look = np.zeros(4, np.float32)
    ids = np.arange(6 * size * size, dtype=np.float32)
    idcube = np.repeat(ids, 4).astype(np.float32)
    for k in range(N):
        O.spo_env_direction(prob, alias, size, size, int(r[k]), float(e[k, 0]),
                            float(e[k, 1]), float(e[k, 2]), out)
        O.spo_cubemap_lookup(idcube, size, size, out, look)
        counts[int(look[0])] += 1
    big = P * N > 50
    chi2 = (((counts[big] - P[big] * N) ** 2) / (P[big] * N)).sum()
    dof = big.sum()
    assert chi2 < dof + 5 * math.sqrt(2 * dof)
    assert counts[~big].sum() == pytest.approx(P[~big].sum() * N, rel=0.1, abs=50)


def _sphere_scene(albedo, Le, env_cube=None):
    O = oracle()
    s = O.spo_scene_create()
    v = np.zeros(2048 * 8, np.float32)
    idx = np.zeros(4096, np.uint32)
    nv, ni = C.c_uint32(), C.c_uint32()
    O.spo_icosphere(2, v, idx, C.byref(nv), C.byref(ni))
    O.spo_mesh_create(s, v, nv.value, idx, ni.value, 1, 1)
    O.spo_scene_add_object(s, 0, 1, f32(0, 0, 0), f32(0, 0, 0, 1), f32(1, 1, 1))
    O.spo_scene_build(s)
    O.spo_material_register(s, f32(albedo, albedo, albedo), 0xFFFFFFFF, f32(0, 0, 0),
                            0xFFFFFFFF, 0.0, 1)
    size = 8
    cube = np.full(6 * size * size * 4, Le, np.float32) if env_cube is None else env_cube
    O.spo_texture_register(s, np.ascontiguousarray(cube), size, size, 5, 1)
    O.spo_material_register(s, f32(0, 0, 0), 0xFFFFFFFF, f32(0, 0, 0), 5, 0.0, 9)
    O.spo_set_background(s, 9)
    assert O.spo_scene_build_env(s) == 1
    return s


def test_nee_closed_form_convex_object_under_constant_sky():
    """A convex Lambert sphere (albedo a, the reference model's Schlick
    Fresnel with F0 = 0.04) under a constant sky L_e, two path vertices: the
    pixel expectation lies between a (1 - 0.04) L_e (F at its minimum
    everywhere) and a L_e, and the NEE and BSDF-only estimators agree within
    their standard errors.  (Here cosine sampling is already nearly
    zero-variance -- f cos / pdf is constant but for F -- so NEE, which
    spreads samples over the whole sky, is not expected to beat it.)"""
    O = oracle()
    a, Le = 0.5, 2.0
    s = _sphere_scene(a, Le)
    cam = SpoCamera()
    O.spo_camera_configure(C.byref(cam), 8, 8, f32(0, 0, 1.6), f32(0, 0, 0, 1), 0.8)
    res = {}
    for est in (SPO_EST_UNBIASED, SPO_EST_NEE):
        vals = []
        for seed in range(24):
            cfg = SpoConfig(rng=SPO_RNG_PCG, tx=SPO_TX_CR, hit=SPO_HIT_BRUTE,
                            estimator=est, samplesPerPixel=16, bounces=2,
                            radianceClamp=0.0, jitterYFirst=1, seed=seed)
            px = np.zeros(8 * 8 * 4, np.float32)
            assert O.spo_path_trace_tile(s, C.byref(cam), C.byref(cfg), 0, 0, 8, 8, None,
                                         px, np.zeros(12, np.uint64)) == 0
            vals.append(px.reshape(-1, 4)[:, 0])
        res[est] = np.array(vals)  # [seeds, pixels]; the camera sees only the sphere
    for est in res:
        m = res[est].mean()
        assert a * (1 - 0.04) * Le * 0.9 < m < a * Le * 1.001, (est, m)
    mu, mn = res[SPO_EST_UNBIASED].mean(), res[SPO_EST_NEE].mean()
    se = math.sqrt(res[SPO_EST_UNBIASED].var() / res[SPO_EST_UNBIASED].size +
                   res[SPO_EST_NEE].var() / res[SPO_EST_NEE].size)
    assert abs(mu - mn) < 4 * se + 1e-6
    O.spo_scene_destroy(s)


def test_nee_agrees_with_unbiased_and_cuts_error_on_config1():
    """Config 1 (sun of radiance 50 in the sky cube): per-pixel means of the
    NEE and the plain unbiased estimators agree (the image-mean difference
    within 4 standard errors), and at equal samples NEE's squared error
    against a long plain render is several times lower."""
    O = oracle()
    desc = scenes.config1(24, 16, env="cube", env_size=32)
    s, cam = load_oracle(desc)

    def render(est, spp, seed):
        cfg = SpoConfig(rng=SPO_RNG_PCG, tx=SPO_TX_CR, hit=SPO_HIT_BRUTE, estimator=est,
                        samplesPerPixel=spp, bounces=3, radianceClamp=0.0,
                        jitterYFirst=1, seed=seed)
        px = np.zeros(24 * 16 * 4, np.float32)
        assert O.spo_path_trace_tile(s, C.byref(cam), C.byref(cfg), 0, 0, 24, 16, None,
                                     px, np.zeros(12, np.uint64)) == 0
        return px.reshape(-1, 4)[:, :3].astype(np.float64)

    ref = render(SPO_EST_UNBIASED, 2048, 99)
    u = [render(SPO_EST_UNBIASED, 16, k) for k in range(8)]
    n = [render(SPO_EST_NEE, 16, k) for k in range(8)]
    mu, mn = np.mean(u), np.mean(n)
    se = math.sqrt(np.var([x.mean() for x in u]) / 8 + np.var([x.mean() for x in n]) / 8)
    assert abs(mu - mn) < 4 * se + 1e-4
    eu = np.mean([((x - ref) ** 2).mean() for x in u])
    en = np.mean([((x - ref) ** 2).mean() for x in n])
    assert en < eu / 3, (en, eu)
    O.spo_scene_destroy(s)


def _lamp_scene():
    """A 6x6 grey floor lit by an emissive sphere (radius 0.3, emission 20)
    one unit above it, black background; camera above, looking down."""
    desc = {
        "meshes": [V.gen_icosphere(2) + (True,), V.gen_plane() + (True,)],
        "objects": [
            (0, 7, (0.0, 1.0, 0.0), (0.0, 0.0, 0.0, 1.0), (0.3, 0.3, 0.3)),
            (1, 8, (0.0, 0.0, 0.0), tuple(scenes.quat_axis_angle((1.0, 0.0, 0.0),
                                                                 -math.pi / 2)),
             (6.0, 6.0, 6.0)),
        ],
        "materials": [
            (7, (0.0, 0.0, 0.0), (20.0, 20.0, 20.0), 0.0, 0xFFFFFFFF, 0xFFFFFFFF),
            (8, (0.5, 0.5, 0.5), (0.0, 0.0, 0.0), 0.3, 0xFFFFFFFF, 0xFFFFFFFF),
            (9, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.0, 0xFFFFFFFF, 0xFFFFFFFF),
        ],
        "background": 9,
        "camera": (16, 12, (0.0, 2.5, 2.5), tuple(scenes.quat_axis_angle(
            (1.0, 0.0, 0.0), -0.8)), 0.8),
        "cubemaps": [], "textures": [],
    }
    return desc


def test_emitter_table_is_a_distribution_over_the_emissive_triangles():
    """Only the emissive object's triangles are in the table (pairs name
    them in index order), P = pdfA x world area sums to 1, and the alias
    table reproduces P."""
    O = oracle()
    desc = _lamp_scene()
    s, _ = load_oracle(desc)
    nt = len(desc["meshes"][0][1]) // 3
    pdfA, prob = np.zeros(4096, np.float32), np.zeros(4096, np.float32)
    alias, pair = np.zeros(4096, np.uint32), np.zeros(8192, np.uint32)
    base = np.zeros(2, np.int32)
    n = O.spo_scene_emitter_table(s, pdfA, prob, alias, pair, base)
    assert n == nt and list(base) == [0, -1]
    pr = pair[: 2 * n].reshape(-1, 2)
    assert (pr[:, 0] == 0).all() and (pr[:, 1] == np.arange(n)).all()
    v, idx = desc["meshes"][0][0], np.asarray(desc["meshes"][0][1]).reshape(-1, 3)
    p = np.asarray(v, np.float64).reshape(-1, 8)[:, :3] * 0.3 + [0.0, 1.0, 0.0]
    t = p[idx]
    area = 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)
    P = pdfA[:n].astype(np.float64) * area
    assert P.sum() == pytest.approx(1.0, abs=1e-5)
    assert np.allclose(P, area / area.sum(), rtol=1e-5)
    rec = prob[:n].astype(np.float64).copy()
    np.add.at(rec, alias[:n], 1.0 - prob[:n].astype(np.float64))
    rec -= np.where(alias[:n] == np.arange(n), 1.0 - prob[:n], 0.0)
    assert np.allclose(rec / n, P, atol=2e-7, rtol=1e-5)
    O.spo_scene_destroy(s)


def test_emitter_nee_agrees_with_unbiased_and_cuts_error():
    """A floor lit only by a small emissive sphere: the NEE estimator's image
    mean equals the plain unbiased estimator's within 4 standard errors,
    and at equal samples its squared error against a long plain render is
    several times lower."""
    O = oracle()
    desc = _lamp_scene()
    s, cam = load_oracle(desc)

    def render(est, spp, seed):
        cfg = SpoConfig(rng=SPO_RNG_PCG, tx=SPO_TX_CR, hit=SPO_HIT_BRUTE, estimator=est,
                        samplesPerPixel=spp, bounces=3, radianceClamp=0.0,
                        jitterYFirst=1, seed=seed)
        px = np.zeros(16 * 12 * 4, np.float32)
        assert O.spo_path_trace_tile(s, C.byref(cam), C.byref(cfg), 0, 0, 16, 12, None,
                                     px, np.zeros(12, np.uint64)) == 0
        return px.reshape(-1, 4)[:, :3].astype(np.float64)

    ref = render(SPO_EST_UNBIASED, 4096, 99)
    u = [render(SPO_EST_UNBIASED, 16, k) for k in range(8)]
    n = [render(SPO_EST_NEE, 16, k) for k in range(8)]
    mu, mn = np.mean(u), np.mean(n)
    se = math.sqrt(np.var([x.mean() for x in u]) / 8 + np.var([x.mean() for x in n]) / 8)
    assert abs(mu - mn) < 4 * se + 1e-4
    eu = np.mean([((x - ref) ** 2).mean() for x in u])
    en = np.mean([((x - ref) ** 2).mean() for x in n])
    assert en < eu / 3, (en, eu)
    O.spo_scene_destroy(s)


def test_sky_occlusion_query_is_any_positive_hit():
    """The sky shadow test (spo_scene_occluded): occluded iff some object's
    closest mesh hit lies at a positive world t, whatever the order in
    which objects are visited -- the broadphase walk and the brute-force
    object loop agree, a ray whose closest hit is at t > 0 is occluded, and
    a ray with no hit is not.  Rays start on the ground next to the
    sphere's contact point (the concave joint) and at random points."""
    from _libs import SPO_HIT_REFERENCE, SpoSceneHit

    desc = scenes.config1(16, 16, env="cube", env_size=8)
    s, _ = load_oracle(desc)
    O = oracle()
    rng = np.random.default_rng(5)
    n_occ = 0
    for k in range(3000):
        if k % 2:
            r = rng.uniform(0.0, 0.3)
            a = rng.uniform(0, 2 * math.pi)
            o = f32(r * math.cos(a), 1e-4, r * math.sin(a))
        else:
            o = f32(*rng.uniform(-3, 3, 3))
            o[1] = abs(o[1]) + 1e-4
        d = rng.normal(size=3)
        d[1] = abs(d[1])
        d = f32(*(d / np.linalg.norm(d)))
        met = np.zeros(12, np.uint64)
        occ_ref = O.spo_scene_occluded(s, o, d, SPO_HIT_REFERENCE, met)
        occ_brute = O.spo_scene_occluded(s, o, d, SPO_HIT_BRUTE, met)
        assert occ_ref == occ_brute
        h = SpoSceneHit()
        assert O.spo_scene_intersect(s, o, d, SPO_HIT_BRUTE, SPO_TX_CR, C.byref(h),
                                     met) == 0
        if h.t > 0.0:
            assert occ_ref == 1
        if h.object == 0xFFFFFFFF:
            assert occ_ref == 0
        n_occ += occ_ref
    assert 300 < n_occ < 2700
    O.spo_scene_destroy(s)


def _lamps_scene(n=5, seed=3):
    """A 10x10 grey floor with n emissive spheres of different radii,
    heights and emissions scattered over it, black background."""
    r = np.random.default_rng(seed)
    objects = [(1, 8, (0.0, 0.0, 0.0), tuple(scenes.quat_axis_angle((1.0, 0.0, 0.0),
                                                                      -math.pi / 2)),
                (10.0, 10.0, 10.0))]
    materials = [(8, (0.5, 0.5, 0.5), (0.0, 0.0, 0.0), 0.3, 0xFFFFFFFF, 0xFFFFFFFF),
                 (9, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.0, 0xFFFFFFFF, 0xFFFFFFFF)]
    for k in range(n):
        rad = float(r.uniform(0.1, 0.4))
        e = float(r.uniform(2.0, 30.0))
        materials.append((100 + k, (0.0, 0.0, 0.0), (e, e * 0.8, e * 0.6), 0.0,
                          0xFFFFFFFF, 0xFFFFFFFF))
        objects.append((0, 100 + k, (float(r.uniform(-4, 4)), float(r.uniform(0.5, 2.0)),
                                     float(r.uniform(-4, 4))), (0.0, 0.0, 0.0, 1.0),
                        (rad, rad, rad)))
    return {
        "meshes": [V.gen_icosphere(2) + (True,), V.gen_plane() + (True,)],
        "objects": objects, "materials": materials, "background": 9,
        "camera": (16, 12, (0.0, 5.0, 7.0), tuple(scenes.quat_axis_angle(
            (1.0, 0.0, 0.0), -0.6)), 0.8),
        "cubemaps": [], "textures": [],
    }


def test_spatial_emitter_tables():
    """The spatial emitter choice's records: one per emissive sphere, in
    object order, with the bounding sphere of its world vertices (centre and
    half diagonal of their box), power = mean emission x world area and
    1 / area (FP64 areas); each sphere's alias table reproduces its
    triangles' area shares; objects that do not emit have no record."""
    O = oracle()
    desc = _lamps_scene()
    s, _ = load_oracle(desc)
    no = len(desc["objects"])
    nt = len(desc["meshes"][0][1]) // 3
    N = 5 * nt
    pdfA, prob = np.zeros(N, np.float32), np.zeros(N, np.float32)
    alias, pair = np.zeros(N, np.uint32), np.zeros(2 * N, np.uint32)
    base = np.zeros(no, np.int32)
    n = O.spo_scene_emitter_table(s, pdfA, prob, alias, pair, base)
    assert n == N and base[0] == -1
    rec = np.zeros(8 * 256, np.float32)
    slot = np.zeros(no, np.int32)
    probL, aliasL = np.zeros(N, np.float32), np.zeros(N, np.uint32)
    ne = O.spo_scene_emitter_instances(s, base, n, rec, slot, probL, aliasL)
    assert ne == 5 and list(slot) == [-1, 0, 1, 2, 3, 4]
    v = np.asarray(desc["meshes"][0][0], np.float64).reshape(-1, 8)[:, :3]
    idx = np.asarray(desc["meshes"][0][1]).reshape(-1, 3)
    for k in range(5):
        _, mid, pos, _, scl = desc["objects"][k + 1]
        em = [m for m in desc["materials"] if m[0] == mid][0][2]
        p = v * scl[0] + np.asarray(pos, np.float64)
        lo, hi = p.min(axis=0), p.max(axis=0)
        t = p[idx]
        area = 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)
        r8 = rec[8 * k: 8 * k + 8]
        assert np.allclose(r8[:3], (lo + hi) / 2, rtol=1e-6, atol=1e-6)
        assert r8[3] == pytest.approx(np.linalg.norm(hi - lo) / 2, rel=1e-6)
        assert r8[4] == pytest.approx(np.mean(em) * area.sum(), rel=1e-6)
        assert r8[5] == pytest.approx(1.0 / area.sum(), rel=1e-6)
        first, count = rec[8 * k + 6:8 * k + 8].view(np.uint32)
        assert first == base[k + 1] and count == nt
        pr = probL[first:first + nt].astype(np.float64)
        al = aliasL[first:first + nt]
        got = pr.copy()
        np.add.at(got, al, 1.0 - pr)
        got -= np.where(al == np.arange(nt), 1.0 - pr, 0.0)
        assert np.allclose(got / nt, area / area.sum(), atol=2e-7, rtol=1e-5)
    O.spo_scene_destroy(s)


def test_spatial_emitter_nee_agrees_with_unbiased_and_cuts_error():
    """Five emissive spheres of different size, height and power over a
    floor: the NEE estimator (the spatial emitter choice) has the plain
    unbiased estimator's image mean within 4 standard errors and, at equal
    samples, a squared error against a long plain render well below the
    plain estimator's."""
    O = oracle()
    desc = _lamps_scene()
    s, cam = load_oracle(desc)

    def render(est, spp, seed):
        cfg = SpoConfig(rng=SPO_RNG_PCG, tx=SPO_TX_CR, hit=SPO_HIT_BRUTE, estimator=est,
                        samplesPerPixel=spp, bounces=3, radianceClamp=0.0,
                        jitterYFirst=1, seed=seed)
        px = np.zeros(16 * 12 * 4, np.float32)
        assert O.spo_path_trace_tile(s, C.byref(cam), C.byref(cfg), 0, 0, 16, 12, None,
                                     px, np.zeros(12, np.uint64)) == 0
        return px.reshape(-1, 4)[:, :3].astype(np.float64)

    ref = render(SPO_EST_UNBIASED, 4096, 99)
    u = [render(SPO_EST_UNBIASED, 16, k) for k in range(8)]
    n = [render(SPO_EST_NEE, 16, k) for k in range(8)]
    mu, mn = np.mean(u), np.mean(n)
    se = math.sqrt(np.var([x.mean() for x in u]) / 8 + np.var([x.mean() for x in n]) / 8)
    assert abs(mu - mn) < 4 * se + 1e-4
    eu = np.mean([((x - ref) ** 2).mean() for x in u])
    en = np.mean([((x - ref) ** 2).mean() for x in n])
    assert en < eu / 3, (en, eu)
    O.spo_scene_destroy(s)


def _buried_lamps_scene():
    """The five-lamp floor with every lamp moved below the floor (y < 0):
    from any point of the floor's upper side all emitter bounding spheres
    lie behind the tangent plane, so the spatial choice has total weight 0
    there and no emitter is sampled; one lamp stays above, seen from the
    lamps' own surfaces."""
    desc = _lamps_scene()
    objs = []
    for k, o in enumerate(desc["objects"]):
        if k >= 2:  # object 0 is the floor, lamp 1 stays above it
            o = (o[0], o[1], (o[2][0], -1.5, o[2][2]), o[3], o[4])
        objs.append(o)
    desc["objects"] = objs
    return desc


def _many_lamps_scene(n=300, seed=11):
    """n tiny emissive icospheres (level 0, 20 triangles) over the floor:
    more emitters than the spatial choice scans, so the global table is
    used."""
    r = np.random.default_rng(seed)
    desc = _lamps_scene(n=0)
    desc["meshes"] = [V.gen_icosphere(0) + (True,), V.gen_plane() + (True,)]
    for k in range(8):  # eight emissions shared by the lamps
        e = float(r.uniform(1.0, 20.0))
        desc["materials"].append((1000 + k, (0.0, 0.0, 0.0), (e, e, e), 0.0,
                                  0xFFFFFFFF, 0xFFFFFFFF))
    for k in range(n):
        desc["objects"].append((0, 1000 + k % 8, (float(r.uniform(-4, 4)),
                                              float(r.uniform(0.3, 2.0)),
                                              float(r.uniform(-4, 4))),
                                (0.0, 0.0, 0.0, 1.0), (0.05, 0.05, 0.05)))
    return desc


def test_spatial_choice_is_off_above_256_emitters():
    """300 emissive objects: no spatial records (the global table is used)
    and every object's slot is -1."""
    O = oracle()
    desc = _many_lamps_scene()
    s, _ = load_oracle(desc)
    no = len(desc["objects"])
    N = 300 * 20
    pdfA, prob = np.zeros(N, np.float32), np.zeros(N, np.float32)
    alias, pair = np.zeros(N, np.uint32), np.zeros(2 * N, np.uint32)
    base = np.zeros(no, np.int32)
    n = O.spo_scene_emitter_table(s, pdfA, prob, alias, pair, base)
    assert n == N
    rec = np.zeros(8 * 256, np.float32)
    slot = np.zeros(no, np.int32)
    probL, aliasL = np.zeros(N, np.float32), np.zeros(N, np.uint32)
    assert O.spo_scene_emitter_instances(s, base, n, rec, slot, probL, aliasL) == 0
    assert (slot == -1).all()
    O.spo_scene_destroy(s)
