"""The oracle against the golden vectors recorded from the reference's own
sources (tests/golden/make_golden.py).  With libm transcendentals the
restatement must reproduce the reference bit for bit."""
import ctypes as C

import numpy as np

from _libs import (SPO_HIT_REFERENCE, SPO_RNG_XORSHIFT_PIXEL,
                   SPO_RNG_XORSHIFT_TILE, SPO_TX_LIBM, SpoConfig, SpoSceneHit,
                   f32, oracle)
from _scenes import load_oracle


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_moller_trumbore_vectors(golden):
    g = golden("prim_vectors")
    O = oracle()
    out = np.zeros((len(g["mt_o"]), 6), np.float32)
    slow = np.zeros_like(out)
    for i in range(len(out)):
        a = [np.ascontiguousarray(x) for x in (g["mt_o"][i], g["mt_d"][i],
                                               g["mt_tri"][i, 0], g["mt_tri"][i, 1],
                                               g["mt_tri"][i, 2])]
        O.spo_rt_mt(*a, out[i])
        O.spo_rt_triangle_slow(*a, slow[i])
    assert np.array_equal(bits(out), bits(g["mt_out"]))
    assert np.array_equal(bits(slow), bits(g["slow_out"]))
    # the vectors exercise both outcomes (half the aimed rays hit back faces)
    assert (g["mt_out"][:, 0] > 0).sum() > 500 and (g["mt_out"][:, 0] < 0).sum() > 1000


def test_slab_vectors(golden):
    g = golden("prim_vectors")
    O = oracle()
    m = len(g["slab_o"])
    mask = np.zeros(m, np.uint32)
    scal = np.zeros(m, np.float32)
    simd1 = np.zeros(m, np.float32)
    with np.errstate(divide="ignore"):
        for i in range(m):
            o = np.ascontiguousarray(g["slab_o"][i])
            d = np.ascontiguousarray(g["slab_d"][i])
            mask[i] = O.spo_rt_aabb4(np.ascontiguousarray(g["slab_lo"][i].ravel()),
                                     np.ascontiguousarray(g["slab_hi"][i].ravel()), o,
                                     np.ascontiguousarray(g["slab_inv"][i]))
            lo0 = np.ascontiguousarray(g["slab_lo"][i, 0])
            hi0 = np.ascontiguousarray(g["slab_hi"][i, 0])
            scal[i] = O.spo_rt_aabb_scalar(lo0, hi0, o, d)
            simd1[i] = O.spo_rt_aabb_simd1(lo0, hi0, o, d)
    assert np.array_equal(mask, g["slab_mask"])
    assert np.array_equal(bits(scal), bits(g["slab_scalar"]))
    assert np.array_equal(bits(scal), bits(g["slab_scalar_test"]))
    assert np.array_equal(bits(simd1), bits(g["slab_simd1"]))
    # axis-parallel rays (infinite inverse direction) are part of the set
    assert np.isinf(g["slab_inv"]).any()


def test_rng_vectors(golden):
    g = golden("prim_vectors")
    O = oracle()
    for k, s in enumerate(g["xs_seeds"]):
        st = C.c_uint32(int(s))
        got = [O.spo_xorshift32(C.byref(st)) for _ in range(64)]
        assert np.array_equal(np.array(got, np.uint32), g["xs_out"][k])
        st = C.c_uint32(int(s))
        uni = np.array([O.spo_random_unilateral(C.byref(st)) for _ in range(64)],
                       np.float32)
        assert np.array_equal(bits(uni), bits(g["uni_out"][k]))


def test_hemisphere_vectors(golden):
    g = golden("prim_vectors")
    O = oracle()
    out = np.zeros((len(g["hemi_n"]), 3), np.float32)
    for i in range(len(out)):
        st = C.c_uint32(int(g["hemi_seed"][i]))
        O.spo_hemisphere(np.ascontiguousarray(g["hemi_n"][i]), C.byref(st),
                         SPO_TX_LIBM, out[i])
        assert st.value == g["hemi_state"][i]
    assert np.array_equal(bits(out), bits(g["hemi_out"]))


def test_map_vectors(golden):
    g = golden("prim_vectors")
    O = oracle()
    sph = np.zeros((len(g["map_dirs"]), 2), np.float32)
    eq = np.zeros_like(sph)
    for i in range(len(sph)):
        O.spo_to_spherical(np.ascontiguousarray(g["map_dirs"][i]), SPO_TX_LIBM, sph[i])
        O.spo_map_to_equirect(np.ascontiguousarray(sph[i]), SPO_TX_LIBM, eq[i])
    assert np.array_equal(bits(sph), bits(g["map_sph"]))
    assert np.array_equal(bits(eq), bits(g["map_eq"]))


def test_sampler_vectors(golden):
    g = golden("prim_vectors")
    O = oracle()
    img = np.ascontiguousarray(g["samp_img"].ravel())
    near = np.zeros((len(g["samp_uv"]), 4), np.float32)
    bil = np.zeros_like(near)
    for i, uv in enumerate(g["samp_uv"]):
        O.spo_sample_nearest(img, 7, 5, np.ascontiguousarray(uv), near[i])
        O.spo_sample_bilinear(img, 7, 5, np.ascontiguousarray(uv), bil[i])
    assert np.array_equal(bits(near), bits(g["samp_nearest"]))
    assert np.array_equal(bits(bil), bits(g["samp_bilinear"]))


def test_cubemap_vectors(golden):
    g = golden("prim_vectors")
    O = oracle()
    out = np.zeros(6 * 8 * 8 * 4, np.float32)
    O.spo_create_cubemap(np.ascontiguousarray(g["cube_src"].ravel()), 32, 16, 8, 8, out)
    assert np.array_equal(bits(out), bits(g["cube_out"]))
    for f in range(6):
        b = np.zeros(9, np.float32)
        O.spo_cubemap_basis(f, b)
        assert np.array_equal(b, g["cube_bases"][f])


def test_transform_vectors(golden):
    g = golden("prim_vectors")
    O = oracle()
    for i in range(len(g["xf_pos"])):
        q = np.zeros(4, np.float32)
        O.spo_quat_axis_angle(np.ascontiguousarray(g["xf_axis"][i]),
                              float(g["xf_angle"][i]), SPO_TX_LIBM, q)
        assert np.array_equal(bits(q), bits(g["xf_quat"][i]))
        m = np.zeros(16, np.float32)
        mi = np.zeros(16, np.float32)
        pos, qq, sc = (np.ascontiguousarray(g[k][i]) for k in ("xf_pos", "xf_quat", "xf_scale"))
        O.spo_object_matrices(pos, qq, sc, m, mi)
        assert np.array_equal(bits(m), bits(g["xf_model"][i]))
        assert np.array_equal(bits(mi), bits(g["xf_inv"][i]))
        box = np.zeros(6, np.float32)
        O.spo_transform_aabb(np.ascontiguousarray(g["xf_boxlo"][i]),
                             np.ascontiguousarray(g["xf_boxhi"][i]), pos, qq, sc, box)
        assert np.array_equal(bits(box), bits(g["xf_box"][i]))
        r = np.zeros(3, np.float32)
        O.spo_rotate_vector(pos, qq, r)
        assert np.array_equal(bits(r), bits(g["xf_rot"][i]))


def test_tile_vectors(golden):
    g = golden("prim_vectors")
    O = oracle()
    for i, c in enumerate(g["tile_cases"]):
        buf = np.zeros(4 * 4096, np.uint32)
        n = O.spo_compute_tiles(*[int(x) for x in c], buf)
        assert n == g["tile_counts"][i]
        assert np.array_equal(buf.reshape(-1, 4)[:n], g["tiles"][i][:n])


def test_icosphere_and_greedy_bvh(golden):
    g = golden("meshes")
    O = oracle()
    v = np.zeros(2048 * 8, np.float32)
    idx = np.zeros(4096, np.uint32)
    nv, ni = C.c_uint32(), C.c_uint32()
    O.spo_icosphere(3, v, idx, C.byref(nv), C.byref(ni))
    assert (nv.value, ni.value) == (1272, 3840)
    assert np.array_equal(bits(v[: 8 * nv.value]), bits(g["ico_verts"]))
    assert np.array_equal(idx[: ni.value], g["ico_idx"])
    tris = g["ico_verts"].reshape(-1, 8)[g["ico_idx"].reshape(-1, 3), :3]
    lo = np.ascontiguousarray(tris.min(axis=1)).ravel()
    hi = np.ascontiguousarray(tris.max(axis=1)).ravel()
    t = O.spo_bvh_create(lo, hi, len(tris))
    cap = 4096
    boxes = np.zeros(6 * cap, np.float32)
    kids = np.zeros(4 * cap, np.int32)
    leaf = np.zeros(cap, np.uint32)
    n = O.spo_bvh_dump(t, boxes, kids, leaf, cap)
    O.spo_bvh_destroy(t)
    assert n == len(g["bvh_leaf"])
    assert np.array_equal(bits(boxes[: 6 * n]), bits(g["bvh_boxes"]))
    assert np.array_equal(kids[: 4 * n], g["bvh_children"])
    assert np.array_equal(leaf[:n], g["bvh_leaf"])
    pv = np.zeros(32, np.float32)
    pi = np.zeros(6, np.uint32)
    O.spo_plane(pv, pi, C.byref(nv), C.byref(ni))
    assert np.array_equal(bits(pv), bits(g["plane_verts"]))
    assert np.array_equal(pi, g["plane_idx"])


def _c1_desc(w, h):
    from vk_cinematic_amd import scenes
    return scenes.config1(w, h, env="equirect", env_size=32)


def test_scene_intersect_vectors(golden):
    g = golden("scene_rays")
    O = oracle()
    s, _ = load_oracle(_c1_desc(32, 32))
    out = np.zeros((len(g["o"]), 7), np.float32)
    met = np.zeros(12, np.uint64)
    for i in range(len(out)):
        h = SpoSceneHit()
        assert O.spo_scene_intersect(s, np.ascontiguousarray(g["o"][i]),
                                     np.ascontiguousarray(g["d"][i]),
                                     SPO_HIT_REFERENCE, SPO_TX_LIBM, C.byref(h), met) == 0
        out[i] = [h.t, np.array([h.materialId], np.uint32).view(np.float32)[0],
                  *h.normal, *h.uv]
    O.spo_scene_destroy(s)
    assert np.array_equal(bits(out), bits(g["out"]))
    # counters (indices 5-9 are the reference's rdtsc cycle sums)
    assert np.array_equal(met[10:], g["metrics"][10:])
    assert (g["out"][:, 0] > 0).sum() > 1000


def test_radiance_for_path_vectors(golden):
    g = golden("radiance")
    O = oracle()
    s = O.spo_scene_create()
    O.spo_material_register(s, f32(0, 0, 0), 0xFFFFFFFF, f32(1, 1, 1), 0xFFFFFFFF, 0.0, 0)
    O.spo_material_register(s, f32(.18, .18, .18), 0xFFFFFFFF, f32(0, 0, 0), 0xFFFFFFFF,
                            0.0, 1)
    O.spo_material_register(s, f32(.5, .3, .2), 0xFFFFFFFF, f32(.1, .2, .3), 0xFFFFFFFF,
                            0.45, 2)
    out = np.zeros(3, np.float32)
    O.spo_compute_radiance_for_path(s, np.ascontiguousarray(g["stale_path"].ravel()), 2,
                                    SPO_TX_LIBM, 10.0, out)
    # TestEvaluateLightPath (unit_tests/test_simd_path_tracer.cpp:278-306)
    # predates the GGX term: with roughness 0 and N.H = 1 the reference
    # computes 0/0 and the path radiance is NaN, not the 0.18 it asserts.
    assert np.isnan(g["stale_out"]).all() and np.isnan(out).all()
    res = np.zeros((len(g["lens"]), 3), np.float32)
    for i in range(len(res)):
        O.spo_compute_radiance_for_path(s, np.ascontiguousarray(g["paths"][i].ravel()),
                                        int(g["lens"][i]), SPO_TX_LIBM, 10.0, res[i])
    O.spo_scene_destroy(s)
    assert np.array_equal(bits(res), bits(g["outs"]))


def test_path_trace_tile_vectors(golden):
    g = golden("c1_render")
    O = oracle()
    W, H = int(g["width"]), int(g["height"])
    s, cam = load_oracle(_c1_desc(W, H))
    for spp in (1, 2):
        cfg = SpoConfig(rng=SPO_RNG_XORSHIFT_PIXEL, tx=SPO_TX_LIBM, hit=SPO_HIT_REFERENCE,
                        estimator=0, samplesPerPixel=spp, bounces=3, radianceClamp=10.0,
                        jitterYFirst=1, seed=int(g["seed_base"]))
        px = np.zeros(W * H * 4, np.float32)
        met = np.zeros(12, np.uint64)
        assert O.spo_path_trace_tile(s, C.byref(cam), C.byref(cfg), 0, 0, W, H, None,
                                     px, met) == 0
        assert np.array_equal(bits(px), bits(g[f"pixels_spp{spp}"]))
        counters = [1, 2, 3, 4, 10, 11]  # 0 and 5-9 are rdtsc cycle sums
        assert np.array_equal(met[counters], g[f"metrics_spp{spp}"][counters])
    cfg = SpoConfig(rng=SPO_RNG_XORSHIFT_TILE, tx=SPO_TX_LIBM, hit=SPO_HIT_REFERENCE,
                    estimator=0, samplesPerPixel=2, bounces=3, radianceClamp=10.0,
                    jitterYFirst=1, seed=0)
    st = C.c_uint32(0xF51C0E49)
    px = np.zeros(W * H * 4, np.float32)
    met = np.zeros(12, np.uint64)
    O.spo_path_trace_tile(s, C.byref(cam), C.byref(cfg), 8, 8, 24, 24, C.byref(st), px, met)
    assert st.value == int(g["tile_state"])
    assert np.array_equal(bits(px.reshape(H, W, 4)[8:24, 8:24]), bits(g["tile_pixels"]))
    assert np.array_equal(met[[1, 2, 3, 4, 10, 11]], g["tile_metrics"][[1, 2, 3, 4, 10, 11]])
    O.spo_scene_destroy(s)
