"""Benchmark workloads as scene descriptions.

A description is plain data (meshes, objects, materials, textures, camera)
that can be loaded into the GPU library (:func:`load_gpu`) or handed to any
other consumer with the same meaning, so that every consumer traces exactly
the same inputs.

* ``config1``: one icosphere (subdivision 3, 1280 triangles, radius 1) on a
  20x20 ground quad, Lambert albedo 0.18 (``sp_Material{albedo 0.18,
  roughness 0}`` in the reference's material model), sky environment with a
  sun disk of radiance 50 as a 6x64x64 RGBA32F cube map, 256x256 pixels,
  4 samples, 3 bounces.
* ``soup``: the closest-hit workload, 2^22 rays against a 1M-triangle soup.
* ``config4`` / ``config5``: 7x7 / 14x14 displaced icospheres of 20480
  triangles each (all distinct geometry, about 1M / 4M triangles) inside a
  closed box with a 1x1 ceiling opening lit by the sky, materials drawn from
  the reference app's material types, 512x512 / 1920x1080, 6 bounces.
"""
import functools
import math

import numpy as np

from . import (Scene, SPG_NO_TEXTURE, gen_icosphere, gen_plane, gen_sky_cubemap,
               gen_sky_equirect, gen_triangle_soup, sp_ConfigureCamera)

SUN_DIR = (0.4, 0.75, -0.5)
SUN_RADIUS = math.radians(3.0)
SUN_RADIANCE = 50.0
ENV_TEXTURE_ID = 100
BACKGROUND_ID = 8
ALBEDO_ID = 1
GROUND_ID = 2


def quat_axis_angle(axis, angle):
    """Quat(axis, angle), src/math_lib.h:556-563, evaluated in float32."""
    a = np.float32(angle) * np.float32(0.5)
    s = np.float32(math.sin(float(a)))
    c = np.float32(math.cos(float(a)))
    ax = np.asarray(axis, np.float32)
    return np.array([ax[0] * s, ax[1] * s, ax[2] * s, c], np.float32)


def config1(width=256, height=256, env="cube", env_size=64):
    """The plain path tracer workload (icosphere on a ground quad)."""
    iv, ii = gen_icosphere(3)
    pv, pi = gen_plane()
    ground_q = quat_axis_angle((1.0, 0.0, 0.0), -math.pi / 2)
    desc = {
        "meshes": [(iv, ii, True), (pv, pi, True)],
        "objects": [
            (0, ALBEDO_ID, (0.0, 1.0, 0.0), (0.0, 0.0, 0.0, 1.0), (1.0, 1.0, 1.0)),
            (1, GROUND_ID, (0.0, 0.0, 0.0), tuple(ground_q), (20.0, 20.0, 20.0)),
        ],
        "materials": [
            (ALBEDO_ID, (0.18, 0.18, 0.18), (0.0, 0.0, 0.0), 0.0,
             SPG_NO_TEXTURE, SPG_NO_TEXTURE),
            (GROUND_ID, (0.18, 0.18, 0.18), (0.0, 0.0, 0.0), 0.0,
             SPG_NO_TEXTURE, SPG_NO_TEXTURE),
            (BACKGROUND_ID, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.0,
             SPG_NO_TEXTURE, ENV_TEXTURE_ID),
        ],
        "background": BACKGROUND_ID,
        "camera": (width, height, (0.0, 1.0, 4.0), (0.0, 0.0, 0.0, 1.0), 0.8),
    }
    if env == "cube":
        desc["cubemaps"] = [(ENV_TEXTURE_ID, gen_sky_cubemap(
            env_size, SUN_DIR, SUN_RADIUS, SUN_RADIANCE), env_size)]
        desc["textures"] = []
    else:
        w, h = 4 * env_size, 2 * env_size
        desc["textures"] = [(ENV_TEXTURE_ID, gen_sky_equirect(
            w, h, SUN_DIR, SUN_RADIUS, SUN_RADIANCE), w, h)]
        desc["cubemaps"] = []
    return desc


def load_gpu(device, desc, max_leaf=4):
    """Builds the description into a GPU Scene; returns (scene, camera)."""
    s = Scene(device)
    for v, i, smooth in desc["meshes"]:
        s.create_mesh(v, i, smooth)
    for mesh, mat, pos, q, sc in desc["objects"]:
        s.add_object(mesh, mat, pos, q, sc)
    for tid, rgba, w, h in desc.get("textures", []):
        s.register_texture(tid, rgba, w, h)
    for tid, rgba, size in desc.get("cubemaps", []):
        s.register_cubemap(tid, rgba, size)
    for mid, albedo, emission, rough, atex, etex in desc["materials"]:
        s.register_material(mid, albedo, emission, rough, atex, etex)
    s.set_background(desc["background"])
    s.build(max_leaf=max_leaf)
    w, h, pos, q, fd = desc["camera"]
    cam = sp_ConfigureCamera(w, h, pos, q, fd)
    return s, cam


@functools.lru_cache(maxsize=8)
def _welded_icosphere(level):
    v, idx = gen_icosphere(level)
    pos = v[:, :3].astype(np.float64)
    key = np.round(pos * 1e6).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inv = inv.ravel()
    return pos[first], inv[idx.astype(np.int64)]


def displaced_icosphere(level, seed, amplitude=0.12, waves=6):
    """An icosphere (the reference's topology, vertices welded) displaced
    along the radius by a seeded sum of plane waves; smooth normals from
    area-weighted face normals; uv = spherical angles."""
    p, idx = _welded_icosphere(level)
    rng = np.random.default_rng(seed)
    freq = rng.normal(size=(waves, 3)) * rng.uniform(2.0, 6.0, size=(waves, 1))
    phase = rng.uniform(0, 2 * np.pi, size=waves)
    amp = rng.uniform(0.3, 1.0, size=waves)
    amp = amp / amp.sum() * amplitude
    r = 1.0 + (np.sin(p @ freq.T + phase) * amp).sum(axis=1)
    p = p * r[:, None]
    tri = idx.reshape(-1, 3)
    fn = np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]])
    n = np.stack([sum(np.bincount(tri[:, k], fn[:, c], len(p)) for k in range(3))
                  for c in range(3)], axis=1)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    out = np.zeros((len(p), 8), np.float32)
    out[:, 0:3] = p
    out[:, 3:6] = n
    out[:, 6] = (np.arctan2(p[:, 2], p[:, 0]) / (2 * np.pi) + 0.5)
    out[:, 7] = np.arccos(np.clip(p[:, 1] / np.linalg.norm(p, axis=1), -1, 1)) / np.pi
    return out, idx.astype(np.uint32).ravel()


def box_with_opening(half=8.0, height=6.0, hole=1.0):
    """A closed box [-half,half] x [0,height] x [-half,half] whose faces
    point inward (the Moller-Trumbore test culls back faces), with a
    hole x hole opening in the middle of the ceiling."""
    quads = []
    h, H, e = half, height, hole / 2

    def quad(a, b, c, d):
        quads.append((a, b, c, d))

    quad((-h, 0, -h), (-h, 0, h), (h, 0, h), (h, 0, -h))          # floor, +y
    quad((-h, 0, -h), (h, 0, -h), (h, H, -h), (-h, H, -h))        # back, +z
    quad((-h, 0, h), (-h, H, h), (h, H, h), (h, 0, h))            # front, -z
    quad((-h, 0, -h), (-h, H, -h), (-h, H, h), (-h, 0, h))        # left, +x
    quad((h, 0, -h), (h, 0, h), (h, H, h), (h, H, -h))            # right, -x
    # ceiling (normal -y) around the opening
    for x0, x1, z0, z1 in ((-h, h, -h, -e), (-h, h, e, h), (-h, -e, -e, e),
                           (e, h, -e, e)):
        quad((x0, H, z0), (x1, H, z0), (x1, H, z1), (x0, H, z1))
    verts, idx = [], []
    for a, b, c, d in quads:
        a, b, c, d = (np.array(x, np.float64) for x in (a, b, c, d))
        n = np.cross(b - a, c - a)
        n /= np.linalg.norm(n)
        base = len(verts)
        for p, uv in ((a, (0, 0)), (b, (1, 0)), (c, (1, 1)), (d, (0, 1))):
            verts.append([*p, *n, *uv])
        idx += [base, base + 1, base + 2, base, base + 2, base + 3]
    return np.array(verts, np.float32), np.array(idx, np.uint32)


# Material types of the reference app (src/main.cpp:1318-1336): diffuse red,
# blue, white with roughness 0.6, black, whites with roughness
# 0.1/0.3/0.6/0.8, and three emitters.
REFERENCE_MATERIALS = [
    ("red", (0.18, 0.1, 0.1), (0, 0, 0), 0.0),
    ("blue", (0.1, 0.1, 0.18), (0, 0, 0), 0.0),
    ("white", (1, 1, 1), (0, 0, 0), 0.6),
    ("black", (0, 0, 0), (0, 0, 0), 0.0),
    ("white_r10", (1, 1, 1), (0, 0, 0), 0.1),
    ("white_r30", (1, 1, 1), (0, 0, 0), 0.3),
    ("white_r60", (1, 1, 1), (0, 0, 0), 0.6),
    ("white_r80", (1, 1, 1), (0, 0, 0), 0.8),
    ("blue_light", (0, 0, 0), (1.6, 2.4, 4.0), 0.0),
    ("white_light", (0, 0, 0), (10, 10, 10), 0.0),
    ("orange_light", (0, 0, 0), (10, 8, 7.5), 0.0),
]


def box_scene(grid, level, width, height, seed=0, unique_meshes=None,
              light_fraction=None, half=8.0, room_height=6.0):
    """Displaced-icosphere instances on a grid x grid layout inside the box
    with a 1x1 ceiling opening, lit by the sky through the opening.  Each
    instance draws a reference material type; non-emitters get albedo
    U[0.1, 0.9] (the type's colour normalized, times the draw).  By default
    every instance has its own displaced mesh (unique_meshes = grid^2), so
    the triangle count is all distinct geometry."""
    rng = np.random.default_rng(seed)
    unique_meshes = unique_meshes or grid * grid
    meshes = [displaced_icosphere(level, seed * 1009 + k) + (True,)
              for k in range(unique_meshes)]
    bv, bi = box_with_opening(half, room_height, 1.0)
    meshes.append((bv, bi, False))
    box_mesh = len(meshes) - 1
    materials = [(1000, (0.7, 0.7, 0.7), (0, 0, 0), 0.0, SPG_NO_TEXTURE,
                  SPG_NO_TEXTURE),
                 (BACKGROUND_ID, (0, 0, 0), (0, 0, 0), 0.0, SPG_NO_TEXTURE,
                  ENV_TEXTURE_ID)]
    objects = [(box_mesh, 1000, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0),
                (1.0, 1.0, 1.0))]
    spacing = 2.0 * (half - 0.8) / grid
    radius = 0.42 * spacing
    mid = 2000
    for gz in range(grid):
        for gx in range(grid):
            x = -half + 0.8 + spacing * (gx + 0.5) + rng.uniform(-0.05, 0.05) * spacing
            z = -half + 0.8 + spacing * (gz + 0.5) + rng.uniform(-0.05, 0.05) * spacing
            s = radius * rng.uniform(0.8, 1.0)
            q = quat_axis_angle(rng.normal(size=3) / np.sqrt(3), rng.uniform(0, np.pi))
            q = (q / np.linalg.norm(q.astype(np.float64))).astype(np.float32)
            if light_fraction is None:
                kind = int(rng.integers(0, len(REFERENCE_MATERIALS)))
            elif rng.uniform() < light_fraction:
                kind = 8 + int(rng.integers(0, 3))
            else:
                kind = int(rng.integers(0, 8))
            name, alb, em, rough = REFERENCE_MATERIALS[kind]
            if kind >= 8:
                albedo = (0.0, 0.0, 0.0)
            else:
                c = np.array(alb, np.float64)
                c = c / c.max() if c.max() > 0 else np.ones(3)
                albedo = tuple(float(v) for v in c * rng.uniform(0.1, 0.9))
            materials.append((mid, albedo, tuple(float(v) for v in em), float(rough),
                              SPG_NO_TEXTURE, SPG_NO_TEXTURE))
            mesh = gz * grid + gx if unique_meshes == grid * grid else \
                int(rng.integers(0, unique_meshes))
            objects.append((mesh, mid,
                            (float(x), float(s * 1.05), float(z)), tuple(q),
                            (float(s), float(s), float(s))))
            mid += 1
    cam_q = quat_axis_angle((1.0, 0.0, 0.0), -0.25)
    desc = {
        "meshes": meshes,
        "objects": objects,
        "materials": materials,
        "background": BACKGROUND_ID,
        "camera": (width, height, (0.0, 3.2, half - 0.6), tuple(cam_q), 0.8),
        "cubemaps": [(ENV_TEXTURE_ID, gen_sky_cubemap(64, SUN_DIR, SUN_RADIUS,
                                                      SUN_RADIANCE), 64)],
        "textures": [],
    }
    tris = sum(len(meshes[o[0]][1]) // 3 for o in objects)
    desc["triangles"] = tris
    return desc


def config4(width=512, height=512, seed=4):
    """About 1M triangles: 7x7 displaced icospheres of 20480 triangles in
    the box with the opening, 512x512, lit only by the sky through the
    opening (surface material types only)."""
    return box_scene(7, 5, width, height, seed=seed, light_fraction=0.0)


def config5(width=1920, height=1080, seed=5):
    """About 4M triangles: 14x14 displaced icospheres of 20480 triangles in
    the box with the opening, 1920x1080."""
    return box_scene(14, 5, width, height, seed=seed)


def soup(n_tris=1_000_000, n_rays=1 << 22, seed=7):
    """Closest-hit workload: origins uniform in [-2,2]^3, directions uniform
    on the sphere, against n_tris triangles with centroids uniform in the unit
    cube and lognormal(mu=-5, sigma=0.5) edge lengths."""
    tris = gen_triangle_soup(n_tris, seed=seed, mu=-5.0, sigma=0.5)
    rng = np.random.default_rng(seed + 1)
    o = rng.uniform(-2.0, 2.0, size=(n_rays, 3)).astype(np.float32)
    z = rng.uniform(-1.0, 1.0, size=n_rays)
    ph = rng.uniform(0.0, 2.0 * np.pi, size=n_rays)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([r * np.cos(ph), r * np.sin(ph), z], axis=1).astype(np.float32)
    return tris, o, d


def soup_mesh(tris):
    """Non-indexed mesh arrays (VertexPNT, indices) of a triangle soup."""
    n = tris.shape[0]
    v = np.zeros((3 * n, 8), np.float32)
    v[:, 0:3] = tris.reshape(-1, 3)
    return v, np.arange(3 * n, dtype=np.uint32)


@functools.lru_cache(maxsize=4)
def _sky_cube(size):
    return gen_sky_cubemap(size, SUN_DIR, SUN_RADIUS, SUN_RADIANCE)


def grid_scene(seed, triangles, side, unique_meshes=4):
    """One batched-workload scene: noise-displaced icospheres (subdivision
    3, 4 or 5 by size) instanced on a jittered square grid over a ground
    quad, about `triangles` triangles in all, Lambert materials (roughness
    0) with per-instance albedo U[0.1, 0.9], the sky cube map with the sun,
    a side x side image looking down on the grid."""
    r = np.random.default_rng(seed)
    level = 3 if triangles < 1e5 else (4 if triangles < 5e5 else 5)
    per = 20 * 4 ** level
    count = max(1, int(round(triangles / per)))
    g = int(math.ceil(math.sqrt(count)))
    meshes = [displaced_icosphere(level, seed * 7919 + j, amplitude=0.15) + (True,)
              for j in range(unique_meshes)]
    pv, pi = gen_plane()
    meshes.append((pv, pi, True))
    ground = len(meshes) - 1
    spacing = 2.4
    extent = g * spacing
    materials = [(1, (0.5, 0.5, 0.5), (0, 0, 0), 0.0, SPG_NO_TEXTURE, SPG_NO_TEXTURE),
                 (BACKGROUND_ID, (0, 0, 0), (0, 0, 0), 0.0, SPG_NO_TEXTURE,
                  ENV_TEXTURE_ID)]
    objects = [(ground, 1, (0.0, 0.0, 0.0),
                tuple(quat_axis_angle((1.0, 0.0, 0.0), -math.pi / 2)),
                (extent * 2.0, extent * 2.0, extent * 2.0))]
    mid = 1000
    for k in range(count):
        gx, gz = k % g, k // g
        x = (gx + 0.5) * spacing - extent / 2 + r.uniform(-0.2, 0.2)
        z = (gz + 0.5) * spacing - extent / 2 + r.uniform(-0.2, 0.2)
        s = float(r.uniform(0.7, 1.0))
        q = quat_axis_angle(r.normal(size=3) / np.sqrt(3), r.uniform(0, np.pi))
        q = (q / np.linalg.norm(q.astype(np.float64))).astype(np.float32)
        a = float(r.uniform(0.1, 0.9))
        materials.append((mid, (a, a, a), (0, 0, 0), 0.0, SPG_NO_TEXTURE,
                          SPG_NO_TEXTURE))
        objects.append((int(r.integers(0, unique_meshes)), mid, (float(x), s, float(z)),
                        tuple(q), (s, s, s)))
        mid += 1
    dist = 0.75 * extent + 4.0
    cam_q = quat_axis_angle((1.0, 0.0, 0.0), -0.6)
    desc = {"meshes": meshes, "objects": objects, "materials": materials,
            "background": BACKGROUND_ID,
            "camera": (side, side, (0.0, 0.68 * dist, 0.82 * dist), tuple(cam_q), 0.8),
            "cubemaps": [(ENV_TEXTURE_ID, _sky_cube(64), 64)], "textures": []}
    desc["triangles"] = sum(len(meshes[o[0]][1]) // 3 for o in objects)
    return desc


def config3(count=64, seed=3, min_side=128, max_side=1024, min_tris=1e4,
            max_tris=2e6):
    """The batched workload: `count` seeded grid scenes, triangle counts
    log-uniform in [min_tris, max_tris], square images with sides uniform
    in [min_side, max_side]; rendered at 16 spp, 5 bounces."""
    out = []
    for k in range(count):
        r = np.random.default_rng([seed, k])
        tris = float(np.exp(r.uniform(np.log(min_tris), np.log(max_tris))))
        side = int(r.integers(min_side, max_side + 1))
        out.append(grid_scene(seed * 1000003 + k, tris, side))
    return out
