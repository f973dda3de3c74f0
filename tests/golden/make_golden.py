"""Writes the golden fixtures of tests/golden/ from the reference's own
sources, compiled by oracle/Makefile into oracle/_ref/ (run `make -C oracle`
first; the reference checkout must be present).

Each fixture holds inputs and the reference's outputs for them:

  prim_vectors.npz   Moller-Trumbore, slab tests (4-wide production and
                     scalar test variants), RNG, hemisphere sampling,
                     spherical/equirect maps, samplers, cube-map build,
                     transforms, tiles
  meshes.npz         icosphere(3) arrays and the greedy BVH built over it
  scene_rays.npz     sp_RayIntersectScene on the config-1 scene
  radiance.npz       ComputeRadianceForPath, including the unit test
                     TestEvaluateLightPath case (unit_tests/
                     test_simd_path_tracer.cpp:278-306)
  c1_render.npz      sp_PathTraceTile on the config-1 scene (equirect sky):
                     per-pixel xorshift streams (1x1 tiles) and one 16x16
                     tile with the app's per-tile seed 0xF51C0E49

Run: python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from _libs import ref, ref_test, f32, NO_TEXTURE  # noqa: E402
from _scenes import load_reference  # noqa: E402

SEED = 20241202


def prim_vectors(R, RT, rng):
    out = {}
    # --- Moller-Trumbore: rays aimed at random triangles, plus misses,
    # vertex/edge hits and grazing and degenerate triangles
    n = 4096
    tri = rng.normal(size=(n, 3, 3)).astype(np.float32)
    tri[: n // 8, 2] = tri[: n // 8, 0] + 1e-7 * rng.normal(size=(n // 8, 3)).astype(np.float32)
    o = rng.uniform(-4, 4, size=(n, 3)).astype(np.float32)
    bary = rng.dirichlet([1, 1, 1], size=n)
    bary[n // 4: n // 4 + 256] = np.eye(3)[rng.integers(0, 3, 256)]
    target = np.einsum("nk,nkj->nj", bary, tri.astype(np.float64))
    target[n // 2:] += rng.normal(size=(n - n // 2, 3)) * 0.5
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    d[n - 256:, rng.integers(0, 3, 256)] = 0.0
    mt = np.zeros((n, 6), np.float32)
    slow = np.zeros((n, 6), np.float32)
    for i in range(n):
        R.ref_mt(o[i].copy(), d[i].copy(), tri[i, 0].copy(), tri[i, 1].copy(),
                 tri[i, 2].copy(), mt[i])
        RT.ref_triangle(o[i].copy(), d[i].copy(), tri[i, 0].copy(),
                        tri[i, 1].copy(), tri[i, 2].copy(), slow[i])
    out.update(mt_o=o, mt_d=d, mt_tri=tri, mt_out=mt, slow_out=slow)

    # --- 4-wide slab test with axis-parallel rays and origins on box planes
    m = 4096
    lo = rng.uniform(-2, 1, size=(m, 4, 3)).astype(np.float32)
    hi = (lo + rng.uniform(0, 2, size=(m, 4, 3))).astype(np.float32)
    so = rng.uniform(-3, 3, size=(m, 3)).astype(np.float32)
    sd = rng.normal(size=(m, 3))
    sd /= np.linalg.norm(sd, axis=1, keepdims=True)
    sd = sd.astype(np.float32)
    sd[: m // 8, rng.integers(0, 3, m // 8)] = 0.0
    so[m // 8: m // 4, 0] = lo[m // 8: m // 4, 0, 0]
    inv = np.zeros((m, 3), np.float32)
    mask = np.zeros(m, np.uint32)
    scal = np.zeros(m, np.float32)
    scal_test = np.zeros(m, np.float32)
    simd1 = np.zeros(m, np.float32)
    with np.errstate(divide="ignore"):
        for i in range(m):
            R.ref_inverse(sd[i].copy(), inv[i])
            mask[i] = R.ref_aabb4(lo[i].ravel().copy(), hi[i].ravel().copy(),
                                  so[i].copy(), inv[i].copy())
            scal[i] = R.ref_aabb_scalar(lo[i, 0].copy(), hi[i, 0].copy(),
                                        so[i].copy(), sd[i].copy())
            scal_test[i] = RT.ref_aabb_scalar(lo[i, 0].copy(), hi[i, 0].copy(),
                                              so[i].copy(), sd[i].copy())
            simd1[i] = R.ref_aabb_simd1(lo[i, 0].copy(), hi[i, 0].copy(),
                                        so[i].copy(), sd[i].copy())
    out.update(slab_lo=lo, slab_hi=hi, slab_o=so, slab_d=sd, slab_inv=inv,
               slab_mask=mask, slab_scalar=scal, slab_scalar_test=scal_test,
               slab_simd1=simd1)

    # --- xorshift streams and unilateral/bilateral conversion
    seeds = np.array([1, 0xF51C0E49, 0x1A34C249, 123456789, 0xFFFFFFFF, 2],
                     np.uint32)
    xs = np.zeros((len(seeds), 64), np.uint32)
    uni = np.zeros((len(seeds), 64), np.float32)
    for k, sd0 in enumerate(seeds):
        st = C.c_uint32(int(sd0))
        for j in range(64):
            xs[k, j] = R.ref_xorshift(C.byref(st))
        st = C.c_uint32(int(sd0))
        for j in range(64):
            uni[k, j] = R.ref_random_unilateral(C.byref(st))
    out.update(xs_seeds=seeds, xs_out=xs, uni_out=uni)

    # --- hemisphere sampling
    h = 2048
    hn = rng.normal(size=(h, 3))
    hn /= np.linalg.norm(hn, axis=1, keepdims=True)
    hn = hn.astype(np.float32)
    hs = rng.integers(1, 2**32 - 1, size=h, dtype=np.uint64).astype(np.uint32)
    hout = np.zeros((h, 3), np.float32)
    hstate = np.zeros(h, np.uint32)
    for i in range(h):
        st = C.c_uint32(int(hs[i]))
        R.ref_random_direction_on_hemisphere(hn[i].copy(), C.byref(st), hout[i])
        hstate[i] = st.value
    out.update(hemi_n=hn, hemi_seed=hs, hemi_out=hout, hemi_state=hstate)

    # --- spherical / equirect maps
    dirs = rng.normal(size=(2048, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    dirs = dirs.astype(np.float32)
    dirs[:6] = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0], [-1, 0, 0],
                         [0, -1, 0], [0, 0, 1]], np.float32)
    sph = np.zeros((len(dirs), 2), np.float32)
    eq = np.zeros((len(dirs), 2), np.float32)
    for i in range(len(dirs)):
        R.ref_to_spherical(dirs[i].copy(), sph[i])
        R.ref_map_to_equirect(sph[i].copy(), eq[i])
    out.update(map_dirs=dirs, map_sph=sph, map_eq=eq)

    # --- samplers
    img = rng.uniform(0, 4, size=(5, 7, 4)).astype(np.float32)
    uv = rng.uniform(0, 1, size=(1024, 2)).astype(np.float32)
    uv[:4] = np.array([[0, 0], [0.5, 0.5], [0.999, 0.999], [0.2, 0.9]], np.float32)
    near = np.zeros((len(uv), 4), np.float32)
    bil = np.zeros((len(uv), 4), np.float32)
    for i in range(len(uv)):
        R.ref_sample_nearest(img.ravel().copy(), 7, 5, uv[i].copy(), near[i])
        R.ref_sample_bilinear(img.ravel().copy(), 7, 5, uv[i].copy(), bil[i])
    out.update(samp_img=img, samp_uv=uv, samp_nearest=near, samp_bilinear=bil)

    # --- cube-map resampling of an equirect image
    eqi = rng.uniform(0, 2, size=(16, 32, 4)).astype(np.float32)
    cube = np.zeros(6 * 8 * 8 * 4, np.float32)
    R.ref_create_cubemap(eqi.ravel().copy(), 32, 16, 8, 8, cube)
    bases = np.zeros((6, 9), np.float32)
    for f in range(6):
        R.ref_cubemap_basis(f, bases[f])
    out.update(cube_src=eqi, cube_out=cube, cube_bases=bases)

    # --- object transforms and AABB transform
    k = 256
    pos = rng.uniform(-5, 5, size=(k, 3)).astype(np.float32)
    axis = rng.normal(size=(k, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    axis = axis.astype(np.float32)
    ang = rng.uniform(-3, 3, size=k).astype(np.float32)
    quat = np.zeros((k, 4), np.float32)
    for i in range(k):
        R.ref_quat_axis_angle(axis[i].copy(), ang[i], quat[i])
    scale = rng.uniform(0.2, 3, size=(k, 1)).repeat(3, 1).astype(np.float32)
    model = np.zeros((k, 16), np.float32)
    invm = np.zeros((k, 16), np.float32)
    box = np.zeros((k, 6), np.float32)
    bl = rng.uniform(-1, 0, size=(k, 3)).astype(np.float32)
    bh = rng.uniform(0, 1, size=(k, 3)).astype(np.float32)
    rot = np.zeros((k, 3), np.float32)
    for i in range(k):
        R.ref_object_matrices(pos[i].copy(), quat[i].copy(), scale[i].copy(),
                              model[i], invm[i])
        R.ref_transform_aabb(bl[i].copy(), bh[i].copy(), pos[i].copy(),
                             quat[i].copy(), scale[i].copy(), box[i])
        R.ref_rotate_vector(pos[i].copy(), quat[i].copy(), rot[i])
    out.update(xf_pos=pos, xf_axis=axis, xf_angle=ang, xf_quat=quat,
               xf_scale=scale, xf_model=model, xf_inv=invm, xf_boxlo=bl,
               xf_boxhi=bh, xf_box=box, xf_rot=rot)

    # --- tiles
    cases = np.array([[10, 10, 2, 2, 64], [9, 9, 2, 2, 64], [10, 10, 2, 2, 10],
                      [1920, 1080, 64, 64, 4096], [256, 256, 64, 64, 4096],
                      [1024, 768, 64, 64, 4096]], np.uint32)
    tiles = np.zeros((len(cases), 4096, 4), np.uint32)
    counts = np.zeros(len(cases), np.uint32)
    for i, c in enumerate(cases):
        buf = np.zeros(4 * 4096, np.uint32)
        counts[i] = R.ref_compute_tiles(*[int(x) for x in c], buf)
        tiles[i] = buf.reshape(4096, 4)
    out.update(tile_cases=cases, tile_counts=counts, tiles=tiles)
    return out


def meshes(R):
    v = np.zeros(2048 * 8, np.float32)
    i = np.zeros(4096, np.uint32)
    nv, ni = C.c_uint32(), C.c_uint32()
    R.ref_icosphere(3, v, i, C.byref(nv), C.byref(ni))
    v, i = v[: 8 * nv.value], i[: ni.value]
    tris = v.reshape(-1, 8)[i.reshape(-1, 3), :3]
    lo = np.ascontiguousarray(tris.min(axis=1)).ravel()
    hi = np.ascontiguousarray(tris.max(axis=1)).ravel()
    R.ref_reset(64 << 20)
    h = R.ref_bvh_create(lo, hi, len(tris))
    cap = 4096
    boxes = np.zeros(6 * cap, np.float32)
    kids = np.zeros(4 * cap, np.int32)
    leaf = np.zeros(cap, np.uint32)
    nn = R.ref_bvh_dump(h, boxes, kids, leaf, cap)
    pv = np.zeros(32, np.float32)
    pi = np.zeros(6, np.uint32)
    R.ref_plane(pv, pi, C.byref(nv), C.byref(ni))
    return dict(ico_verts=v, ico_idx=i, bvh_boxes=boxes[: 6 * nn],
                bvh_children=kids[: 4 * nn], bvh_leaf=leaf[:nn],
                plane_verts=pv, plane_idx=pi)


def c1_desc(width, height):
    from vk_cinematic_amd import scenes
    return scenes.config1(width, height, env="equirect", env_size=32)


def scene_rays(R, rng):
    desc = c1_desc(32, 32)
    load_reference(desc)
    n = 4096
    o = np.zeros((n, 3), np.float32)
    o[:, 0] = rng.uniform(-3, 3, n)
    o[:, 1] = rng.uniform(0.05, 3, n)
    o[:, 2] = rng.uniform(-3, 5, n)
    target = np.stack([rng.uniform(-1.2, 1.2, n), rng.uniform(-0.5, 2.2, n),
                       rng.uniform(-1.2, 1.2, n)], 1)
    d = target - o
    d[n // 2:] = rng.normal(size=(n - n // 2, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    res = np.zeros((n, 7), np.float32)
    met = np.zeros(12, np.uint64)
    for i in range(n):
        R.ref_scene_intersect(o[i].copy(), d[i].copy(), res[i], met)
    return dict(o=o, d=d, out=res, metrics=met)


def radiance(R, rng):
    R.ref_reset(16 << 20)
    # TestEvaluateLightPath's materials and path
    R.ref_material_register(f32(0, 0, 0), NO_TEXTURE, f32(1, 1, 1), NO_TEXTURE, 0.0, 0)
    R.ref_material_register(f32(.18, .18, .18), NO_TEXTURE, f32(0, 0, 0), NO_TEXTURE,
                            0.0, 1)
    R.ref_material_register(f32(.5, .3, .2), NO_TEXTURE, f32(.1, .2, .3), NO_TEXTURE,
                            0.45, 2)
    stale = np.zeros((2, 15), np.float32)
    stale[0, 0] = np.array([1], np.uint32).view(np.float32)[0]
    stale[0, 10:13] = (0, 1, 0)
    stale[0, 7:10] = (0, 1, 0)
    stale_out = np.zeros(3, np.float32)
    R.ref_compute_radiance_for_path(stale.ravel().copy(), 2, stale_out)
    n = 512
    paths = np.zeros((n, 3, 15), np.float32)
    lens = rng.integers(1, 4, n).astype(np.uint32)
    outs = np.zeros((n, 3), np.float32)
    for i in range(n):
        for k in range(3):
            ids = np.array([rng.integers(0, 4)], np.uint32)
            paths[i, k, 0] = ids.view(np.float32)[0]
            nrm = rng.normal(size=3)
            nrm /= np.linalg.norm(nrm)
            inc = rng.normal(size=3)
            inc /= np.linalg.norm(inc)
            out_ = rng.normal(size=3)
            out_ /= np.linalg.norm(out_)
            paths[i, k, 1:4] = rng.normal(size=3)
            paths[i, k, 4:7] = out_
            paths[i, k, 7:10] = inc
            paths[i, k, 10:13] = nrm
            paths[i, k, 13:15] = rng.uniform(0, 1, 2)
        R.ref_compute_radiance_for_path(paths[i].ravel().copy(), int(lens[i]), outs[i])
    return dict(stale_path=stale, stale_out=stale_out, paths=paths, lens=lens,
                outs=outs)


def c1_render(R):
    from _libs import oracle
    W = H = 32
    desc = c1_desc(W, H)
    load_reference(desc)
    seeds = np.array([oracle().spo_pixel_seed(0x5EED, i) for i in range(W * H)],
                     np.uint32)
    out = {}
    for spp in (1, 2):
        px = np.zeros(W * H * 4, np.float32)
        met = np.zeros(12, np.uint64)
        R.ref_path_trace_pixels(0, 0, W, H, seeds, spp, met, px)
        out[f"pixels_spp{spp}"] = px
        out[f"metrics_spp{spp}"] = met
    st = C.c_uint32(0xF51C0E49)
    px = np.zeros(W * H * 4, np.float32)
    met = np.zeros(12, np.uint64)
    R.ref_path_trace_tile(8, 8, 24, 24, C.byref(st), 2, met, px)
    out.update(tile_pixels=px.reshape(H, W, 4)[8:24, 8:24].copy(), tile_metrics=met,
               tile_state=np.uint32(st.value), seed_base=np.uint32(0x5EED),
               width=np.uint32(W), height=np.uint32(H))
    return out


def main():
    R, RT = ref(), ref_test()
    if R is None or RT is None:
        sys.exit("oracle/_ref is not built: run `make -C oracle` with the "
                 "reference checkout present")
    rng = np.random.default_rng(SEED)
    R.ref_reset(64 << 20)
    np.savez_compressed(os.path.join(HERE, "prim_vectors.npz"), **prim_vectors(R, RT, rng))
    np.savez_compressed(os.path.join(HERE, "meshes.npz"), **meshes(R))
    np.savez_compressed(os.path.join(HERE, "scene_rays.npz"), **scene_rays(R, rng))
    np.savez_compressed(os.path.join(HERE, "radiance.npz"), **radiance(R, rng))
    np.savez_compressed(os.path.join(HERE, "c1_render.npz"), **c1_render(R))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
