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
"""
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
