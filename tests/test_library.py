"""The C-ABI library on the CPU: it loads, exports every entry point that
include/spgpu.h declares, and its host-only entry points (generators, camera,
tiles, the SAH builder) agree with the oracle.  No kernel is launched."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import vk_cinematic_amd as V
from _libs import SpoCamera, f32, oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "spgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(V.LIB_PATH)
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_declared_symbols():
    """Every declared entry point has a typed Python binding."""
    from vk_cinematic_amd import _SIGS
    assert set(declared_symbols()) <= set(_SIGS)


def test_errors_are_reported_not_aborted():
    lib = V.lib()
    assert lib.spg_device_create(0, None) == -1
    assert b"NULL" in lib.spg_last_error()
    assert lib.spg_compute_tiles(4, 4, 0, 2, None, 0, None) == -1
    with pytest.raises(V.SpgError):
        V.ComputeTiles(4, 4, 0, 2)


def test_icosphere_matches_reference_topology(golden):
    g = golden("meshes")
    v, i = V.gen_icosphere(3)
    assert np.array_equal(v.ravel().view(np.uint32), g["ico_verts"].view(np.uint32))
    assert np.array_equal(i, g["ico_idx"])
    pv, pi = V.gen_plane()
    assert np.array_equal(pv.ravel(), g["plane_verts"])
    assert np.array_equal(pi, g["plane_idx"])


@pytest.mark.parametrize("w,h,pos,q,fd", [
    (4, 2, (0, 2, 0), (0, 0, 0, 1), 0.1),
    (256, 256, (0, 1, 4), (0, 0, 0, 1), 0.8),
    (1920, 1080, (1, 2, 3), (0.1, 0.2, 0.3, 0.927362), 0.8),
    (200, 480, (-1, 0.5, 7), (0, 0.3826834, 0, 0.9238795), 0.3),
])
def test_camera_matches_oracle(w, h, pos, q, fd):
    """sp_ConfigureCamera, src/simd_path_tracer.cpp:1-36"""
    cam = V.sp_ConfigureCamera(w, h, pos, q, fd)
    oc = SpoCamera()
    oracle().spo_camera_configure(C.byref(oc), w, h, f32(*pos), f32(*q), fd)
    assert bytes(cam) == bytes(oc)


def test_compute_tiles_matches_reference(golden):
    g = golden("prim_vectors")
    for i, c in enumerate(g["tile_cases"]):
        tiles = V.ComputeTiles(*[int(x) for x in c])
        assert len(tiles) == g["tile_counts"][i]
        got = np.array([[t.minX, t.minY, t.maxX, t.maxY] for t in tiles], np.uint32)
        assert np.array_equal(got, g["tiles"][i][: len(tiles)])


def test_pixel_seed_matches_oracle():
    O = oracle()
    for base in (0, 1, 0x5EED, 0xFFFFFFFF):
        for idx in (0, 1, 17, 65535, 2073599):
            assert V.pixel_seed(base, idx) == O.spo_pixel_seed(base, idx)
            assert V.pixel_seed(base, idx) != 0


def _xorshift_right_inverse(y, s):
    """x with x ^ (x >> s) == y on 32 bits (fixed point of x = y ^ (x >> s))."""
    x = y
    for _ in range(32 // s + 1):
        x = y ^ (x >> s)
    return x


def test_pixel_seed_is_the_published_finalizer_of_a_golden_ratio_sequence():
    """The per-pixel seed is C. Wellons' "lowbias32" integer finalizer
    (hash-prospector: xorshift 16, multiply 0x7FEB352D, xorshift 15,
    multiply 0x846CA68B, xorshift 16) of base + floor(2^32 / phi) * index,
    with 0 remapped because a xorshift32 state of 0 never leaves it.  Pinned
    without restating the forward steps: running the inverse (xorshift
    inverses, modular inverses of the multipliers) on the seed recovers the
    golden-ratio sequence, the one input that the finalizer sends to 0 gets
    the remap, and one flipped input bit flips every output bit with
    probability 1/2 (avalanche, within sampling error)."""
    import math
    import random
    O = oracle()
    M = 1 << 32
    golden = int((math.sqrt(5) - 1) / 2 * M)
    ia, ib = pow(0x7FEB352D, -1, M), pow(0x846CA68B, -1, M)

    def inverse(y):
        x = _xorshift_right_inverse(y, 16)
        x = _xorshift_right_inverse(x * ib % M, 15)
        return _xorshift_right_inverse(x * ia % M, 16)

    r = random.Random(5)
    cases = [(r.getrandbits(32), r.getrandbits(32)) for _ in range(3000)]
    cases += [(b, i) for b in (0, 1, 0xFFFFFFFF) for i in (1, 2, 1920 * 1080 - 1)]
    for base, idx in cases:
        y = O.spo_pixel_seed(base, idx)
        assert V.pixel_seed(base, idx) == y
        assert inverse(y) == (base + golden * idx) % M, (base, idx)
    # the finalizer fixes 0: the index whose sequence value is 0 gets the remap
    gi = pow(golden, -1, M)
    for base in (0, 5, 0xDEADBEEF):
        idx = (-base * gi) % M
        assert O.spo_pixel_seed(base, idx) == V.pixel_seed(base, idx) == 0x6D2B79F5
    n = 2048
    flips = np.zeros((32, 32))
    for _ in range(n):
        x = r.getrandbits(32)
        y = O.spo_pixel_seed(x, 0)
        for i in range(32):
            d = y ^ O.spo_pixel_seed(x ^ (1 << i), 0)
            flips[i] += [(d >> j) & 1 for j in range(32)]
    p = flips / n
    # sampling error of one probability: 0.5 / sqrt(2048) = 0.011
    assert np.abs(p - 0.5).max() < 0.06 and np.abs(p - 0.5).mean() < 0.015


def test_sky_generators():
    sky = V.gen_sky_cubemap(16, sun_radius=0.2).reshape(6, 16, 16, 4)
    assert np.all(sky[..., 3] == 1) and np.isclose(sky[..., :3].max(), 50.0)
    assert (sky[..., 0] == 50).sum() > 0
    eq = V.gen_sky_equirect(64, 32, sun_radius=0.2).reshape(32, 64, 4)
    assert (eq[..., 0] == 50).sum() > 0
    # the upper hemisphere (image top rows) is sky, the bottom ground
    assert eq[0, :, 2].mean() > eq[-1, :, 2].mean()


def test_triangle_soup_statistics():
    t = V.gen_triangle_soup(20000, seed=3).reshape(-1, 3, 3).astype(np.float64)
    c = t.mean(axis=1)
    assert c.min() > -0.05 and c.max() < 1.05
    e = np.linalg.norm(t[:, 1] - t[:, 0], axis=1)
    assert np.median(e) == pytest.approx(np.exp(-5.0), rel=0.03)
    assert np.std(np.log(e)) == pytest.approx(0.5, rel=0.05)
    # equilateral
    e2 = np.linalg.norm(t[:, 2] - t[:, 1], axis=1)
    assert np.allclose(e, e2, rtol=1e-3)


@pytest.mark.parametrize("leaf", [1, 4, 8])
def test_sah_bvh_invariants(leaf):
    """Every triangle referenced by exactly one leaf; child boxes contain
    their subtrees; depth within the traversal stack bound."""
    tris = V.gen_triangle_soup(30000, seed=5)
    nodes, order, depth = V.bvh_build_host(tris, max_leaf=leaf)
    assert depth <= 42
    rc = oracle().spo_bvh2_check(np.ascontiguousarray(nodes.ravel()), len(nodes), order,
                                 np.ascontiguousarray(tris.ravel()), len(tris))
    assert rc == 0


def test_sah_bvh_depth_bound_on_degenerate_input():
    """All centroids coincide: the builder falls back to balanced splits and
    stays within the depth bound."""
    tris = np.tile(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32), (5000, 1))
    nodes, order, depth = V.bvh_build_host(tris, max_leaf=2)
    assert depth <= 42
    rc = oracle().spo_bvh2_check(np.ascontiguousarray(nodes.ravel()), len(nodes), order,
                                 np.ascontiguousarray(tris.ravel()), len(tris))
    assert rc == 0


def test_traversal_twin_equals_brute_force():
    """The GPU traversal algorithm, restated on the CPU over the product's
    BVH, returns the brute-force closest hit (smallest t > 0, ties to the
    lower triangle index) for every ray."""
    from vk_cinematic_amd import scenes
    _, o, d = scenes.soup(16, 4096, seed=9)
    # larger triangles than the benchmark soup so that many rays hit
    tris = np.ascontiguousarray(V.gen_triangle_soup(20000, seed=9, mu=-2.5, sigma=0.5))
    rays = np.ascontiguousarray(np.concatenate([o, d], 1).ravel())
    nodes, order, _ = V.bvh_build_host(tris, max_leaf=4)
    O = oracle()
    p1 = np.zeros(len(o), np.uint32)
    t1 = np.zeros(len(o), np.float32)
    box, tri = C.c_uint64(), C.c_uint64()
    O.spo_bvh2_closest(np.ascontiguousarray(nodes.ravel()), order, tris.ravel(), rays,
                       len(o), 8, p1, t1, C.byref(box), C.byref(tri))
    p2 = np.zeros(len(o), np.uint32)
    t2 = np.zeros(len(o), np.float32)
    O.spo_closest_hit_brute(tris.ravel(), len(tris), rays, len(o), 8, p2, t2)
    assert np.array_equal(p1, p2)
    assert np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
    assert (p2 != 0xFFFFFFFF).sum() > 50
    assert box.value > len(o) and tri.value > 0
