"""ctypes bindings for the test-side libraries.

* ``oracle()``  -- oracle/liboracle.so, the C restatement of the reference
  tracer (the checker; see oracle/spo.h).
* ``ref()`` / ``ref_test()`` -- oracle/_ref/libref*.so, the reference's own
  sources compiled by oracle/Makefile where the reference checkout exists.
  The built files travel with the repository's snapshot like the other
  built libraries, but only CPU tests (not marked ``gpu``) load them, and
  only when they run (never at collection); tests that need them skip when
  the build is absent.

Nothing in the product imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")
REF_TEST_SO = os.path.join(ROOT, "oracle", "_ref", "libref_test.so")

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
u32ref = C.POINTER(C.c_uint32)

SPO_RNG_XORSHIFT_TILE, SPO_RNG_XORSHIFT_PIXEL, SPO_RNG_PCG = 0, 1, 2
SPO_TX_LIBM, SPO_TX_CR = 0, 1
SPO_HIT_REFERENCE, SPO_HIT_BRUTE = 0, 1
SPO_EST_PARITY, SPO_EST_UNBIASED, SPO_EST_GUIDED, SPO_EST_NEE, SPO_EST_GUIDED_NEE = 0, 1, 2, 3, 4
NO_TEXTURE = 0xFFFFFFFF


class SpoConfig(C.Structure):
    _fields_ = [("rng", C.c_int), ("tx", C.c_int), ("hit", C.c_int),
                ("estimator", C.c_int), ("samplesPerPixel", C.c_uint32),
                ("bounces", C.c_uint32), ("radianceClamp", C.c_float),
                ("jitterYFirst", C.c_int), ("seed", C.c_uint32),
                ("bsdfFraction", C.c_float)]


class SpoCamera(C.Structure):
    _fields_ = [("right", C.c_float * 3), ("up", C.c_float * 3),
                ("forward", C.c_float * 3), ("position", C.c_float * 3),
                ("filmCenter", C.c_float * 3), ("width", C.c_uint32),
                ("height", C.c_uint32), ("halfPixelWidth", C.c_float),
                ("halfPixelHeight", C.c_float), ("halfFilmWidth", C.c_float),
                ("halfFilmHeight", C.c_float)]


class SpoSceneHit(C.Structure):
    _fields_ = [("t", C.c_float), ("materialId", C.c_uint32),
                ("normal", C.c_float * 3), ("uv", C.c_float * 2),
                ("object", C.c_uint32), ("prim", C.c_uint32)]


GUIDE_PARAMS = 4 * (64 * 64 + 64) + 48 * 64 + 48
GUIDE_LOBES = 8


class SpoLobes(C.Structure):
    _fields_ = [("pi", C.c_double * 8), ("kappa", C.c_double * 8),
                ("mu", (C.c_double * 3) * 8)]


def _sig(lib, name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)


_oracle = None


def oracle():
    global _oracle
    if _oracle is not None:
        return _oracle
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError("oracle/liboracle.so is not built (run build())")
    L = C.CDLL(ORACLE_SO)
    vp = C.c_void_p
    _sig(L, "spo_rt_mt", C.c_float, f32p, f32p, f32p, f32p, f32p, f32p)
    _sig(L, "spo_rt_triangle_slow", C.c_float, f32p, f32p, f32p, f32p, f32p, f32p)
    _sig(L, "spo_rt_aabb_scalar", C.c_float, f32p, f32p, f32p, f32p)
    _sig(L, "spo_rt_aabb_simd1", C.c_float, f32p, f32p, f32p, f32p)
    _sig(L, "spo_rt_aabb4", C.c_uint32, f32p, f32p, f32p, f32p)
    _sig(L, "spo_xorshift32", C.c_uint32, u32ref)
    _sig(L, "spo_random_unilateral", C.c_float, u32ref)
    _sig(L, "spo_random_bilateral", C.c_float, u32ref)
    _sig(L, "spo_pixel_seed", C.c_uint32, C.c_uint32, C.c_uint32)
    _sig(L, "spo_pcg32_seed", None, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
         C.c_uint64, C.c_uint64)
    _sig(L, "spo_pcg32_next", C.c_uint32, C.POINTER(C.c_uint64), C.c_uint64)
    _sig(L, "spo_pcg32_advance", None, C.POINTER(C.c_uint64), C.c_uint64, C.c_uint64)
    _sig(L, "spo_hemisphere", None, f32p, u32ref, C.c_int, f32p)
    _sig(L, "spo_to_spherical", None, f32p, C.c_int, f32p)
    _sig(L, "spo_map_to_equirect", None, f32p, C.c_int, f32p)
    _sig(L, "spo_spherical_to_cartesian", None, f32p, C.c_int, f32p)
    _sig(L, "spo_sample_nearest", None, f32p, C.c_uint32, C.c_uint32, f32p, f32p)
    _sig(L, "spo_sample_bilinear", None, f32p, C.c_uint32, C.c_uint32, f32p, f32p)
    _sig(L, "spo_cubemap_basis", None, C.c_uint32, f32p)
    _sig(L, "spo_cubemap_lookup", None, f32p, C.c_uint32, C.c_uint32, f32p, f32p)
    _sig(L, "spo_env_table_build", C.c_int, f32p, C.c_uint32, C.c_uint32, f32p, f32p,
         u32p)
    _sig(L, "spo_env_pdf", C.c_float, f32p, C.c_uint32, C.c_uint32, f32p)
    _sig(L, "spo_env_direction", None, f32p, u32p, C.c_uint32, C.c_uint32, C.c_uint32,
         C.c_float, C.c_float, C.c_float, f32p)
    _sig(L, "spo_scene_build_env", C.c_int, vp)
    _sig(L, "spo_sincos_cr_array", None, f32p, C.c_uint32, f32p, f32p)
    _sig(L, "spo_scene_emitter_table", C.c_uint32, vp, f32p, f32p, u32p, u32p, i32p)
    _sig(L, "spo_scene_set_light_split", None, vp, C.c_float)
    _sig(L, "spo_scene_emitter_instances", C.c_uint32, vp, i32p, C.c_uint32, f32p, i32p,
         f32p, u32p)
    _sig(L, "spo_create_cubemap", None, f32p, C.c_uint32, C.c_uint32, C.c_uint32,
         C.c_uint32, f32p)
    _sig(L, "spo_create_irradiance_cubemap", None, f32p, C.c_uint32, C.c_uint32,
         C.c_uint32, C.c_uint32, C.c_float, C.c_int, f32p)
    _sig(L, "spo_tonemap", None, f32p, C.c_uint32, C.c_float, C.c_int,
         np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS"))
    _sig(L, "spo_accumulate", None, f32p, f32p, C.c_uint32, C.c_uint32, C.c_uint32)
    _sig(L, "spo_compute_tiles", C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
         C.c_uint32, C.c_uint32, u32p)
    _sig(L, "spo_transform_aabb", None, f32p, f32p, f32p, f32p, f32p, f32p)
    _sig(L, "spo_object_matrices", None, f32p, f32p, f32p, f32p, f32p)
    _sig(L, "spo_rotate_vector", None, f32p, f32p, f32p)
    _sig(L, "spo_quat_axis_angle", None, f32p, C.c_float, C.c_int, f32p)
    _sig(L, "spo_quat_mul", None, f32p, f32p, f32p)
    _sig(L, "spo_icosphere", None, C.c_uint32, f32p, u32p, u32ref, u32ref)
    _sig(L, "spo_plane", None, f32p, u32p, u32ref, u32ref)
    _sig(L, "spo_bvh_create", vp, f32p, f32p, C.c_uint32)
    _sig(L, "spo_bvh_destroy", None, vp)
    _sig(L, "spo_bvh_intersect", C.c_uint32, vp, f32p, f32p, C.c_uint32, u32p,
         u32ref, u32ref)
    _sig(L, "spo_bvh_dump", C.c_uint32, vp, f32p, i32p, u32p, C.c_uint32)
    _sig(L, "spo_scene_create", vp)
    _sig(L, "spo_scene_destroy", None, vp)
    _sig(L, "spo_mesh_create", C.c_int, vp, f32p, C.c_uint32, u32p, C.c_uint32,
         C.c_int, C.c_int)
    _sig(L, "spo_scene_add_object", C.c_int, vp, C.c_uint32, C.c_uint32, f32p,
         f32p, f32p)
    _sig(L, "spo_mesh_set_bvh2", C.c_int, vp, C.c_uint32, f32p, C.c_uint32, u32p)
    _sig(L, "spo_scene_set_limits", None, vp, C.c_uint32, C.c_uint32, C.c_uint32)
    _sig(L, "spo_scene_build", None, vp)
    _sig(L, "spo_material_register", C.c_int, vp, f32p, C.c_uint32, f32p,
         C.c_uint32, C.c_float, C.c_uint32)
    _sig(L, "spo_texture_register", C.c_int, vp, f32p, C.c_uint32, C.c_uint32,
         C.c_uint32, C.c_int)
    _sig(L, "spo_set_background", None, vp, C.c_uint32)
    _sig(L, "spo_evaluate_material", None, vp, C.c_uint32, f32p, f32p, C.c_int, f32p)
    _sig(L, "spo_compute_radiance_for_path", None, vp, f32p, C.c_uint32, C.c_int,
         C.c_float, f32p)
    _sig(L, "spo_camera_configure", None, C.POINTER(SpoCamera), C.c_uint32,
         C.c_uint32, f32p, f32p, C.c_float)
    _sig(L, "spo_film_position", None, C.POINTER(SpoCamera), C.c_float, C.c_float,
         f32p)
    _sig(L, "spo_scene_intersect", C.c_int, vp, f32p, f32p, C.c_int, C.c_int,
         C.POINTER(SpoSceneHit), u64p)
    _sig(L, "spo_mesh_intersect", C.c_float, vp, C.c_uint32, f32p, f32p, C.c_int,
         f32p, u32ref, u64p)
    _sig(L, "spo_scene_occluded", C.c_int, vp, f32p, f32p, C.c_int, u64p)
    _sig(L, "spo_guided_scatter", None, f32p, f32p, f32p, C.c_float, C.c_float,
         C.c_float, C.c_float, C.c_float, C.c_int, f32p, C.POINTER(C.c_float))
    _sig(L, "spo_guide_analytic_lobes", None, f32p, f32p, f32p, f32p)
    _sig(L, "spo_guided_sink", None, C.c_void_p, C.c_uint32)
    _sig(L, "spo_guided_sink_count", C.c_uint32)
    _sig(L, "spo_compute_tracer_tile", None, vp, f32p, C.c_uint32, f32p)
    _sig(L, "spo_checker_bilinear", C.c_float, C.c_float, C.c_float)
    _sig(L, "spo_closest_hit_brute", None, f32p, C.c_uint32, f32p, C.c_uint32,
         C.c_uint32, u32p, f32p)
    _sig(L, "spo_closest_hit_brute_f64", None, f32p, C.c_uint32, f32p, C.c_uint32,
         C.c_uint32, u32p, f64p)
    _sig(L, "spo_mt_f64_pairs", None, f32p, f32p, u32p, C.c_uint32, f64p)
    _sig(L, "spo_bvh2_closest", None, f32p, u32p, f32p, f32p, C.c_uint32,
         C.c_uint32, u32p, f32p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64))
    _sig(L, "spo_bvh2_check", C.c_int, f32p, C.c_uint32, u32p, f32p, C.c_uint32)
    _sig(L, "spo_wide_closest", None, u8p, u32p, f32p, f32p, C.c_uint32,
         C.c_uint32, u32p, f32p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64))
    _sig(L, "spo_wide_check", C.c_int, u8p, C.c_uint32, u32p, f32p, C.c_uint32,
         C.POINTER(C.c_uint32))
    _sig(L, "spo_path_trace_tile", C.c_int, vp, C.POINTER(SpoCamera),
         C.POINTER(SpoConfig), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
         u32ref, f32p, u64p)
    _sig(L, "spo_render_threaded", C.c_double, vp, C.POINTER(SpoCamera),
         C.POINTER(SpoConfig), C.c_uint32, C.c_uint32, C.c_uint32, f32p, u64p)
    # learned guide (oracle/spo_guide.h)
    _sig(L, "spo_guide_encode", None, f32p, f32p, f32p, f64p)
    _sig(L, "spo_guide_forward", None, f64p, f64p, f64p, C.c_void_p)
    _sig(L, "spo_guide_lobes", None, f64p, C.POINTER(SpoLobes))
    _sig(L, "spo_guide_pdf", C.c_double, C.POINTER(SpoLobes), f64p)
    _sig(L, "spo_guide_sample", None, C.POINTER(SpoLobes), C.c_double, C.c_double,
         C.c_double, f64p)
    _sig(L, "spo_vmf_pdf", C.c_double, C.c_double, f64p, f64p)
    _sig(L, "spo_vmf_sample", None, C.c_double, f64p, C.c_double, C.c_double, f64p)
    _sig(L, "spo_guide_loss", C.c_double, f64p, f64p, f64p, f64p, C.c_uint32,
         C.c_void_p)
    _sig(L, "spo_guide_adam", None, f32p, f32p, f32p, f64p, C.c_uint32, C.c_double,
         C.c_double, C.c_double, C.c_double, C.c_uint32)
    _oracle = L
    return L


_refs = {}


def _load_ref(path):
    if path in _refs:
        return _refs[path]
    if not os.path.exists(path):
        _refs[path] = None
        return None
    L = C.CDLL(path)
    vp = C.c_void_p
    _sig(L, "ref_is_test_config", C.c_int)
    _sig(L, "ref_reset", None, C.c_ulonglong)
    _sig(L, "ref_mt", None, f32p, f32p, f32p, f32p, f32p, f32p)
    _sig(L, "ref_triangle", None, f32p, f32p, f32p, f32p, f32p, f32p)
    _sig(L, "ref_aabb4", C.c_uint32, f32p, f32p, f32p, f32p)
    _sig(L, "ref_aabb_scalar", C.c_float, f32p, f32p, f32p, f32p)
    _sig(L, "ref_aabb_simd1", C.c_float, f32p, f32p, f32p, f32p)
    _sig(L, "ref_inverse", None, f32p, f32p)
    _sig(L, "ref_xorshift", C.c_uint32, u32ref)
    _sig(L, "ref_random_unilateral", C.c_float, u32ref)
    _sig(L, "ref_random_bilateral", C.c_float, u32ref)
    _sig(L, "ref_random_direction_on_hemisphere", None, f32p, u32ref, f32p)
    _sig(L, "ref_to_spherical", None, f32p, f32p)
    _sig(L, "ref_map_to_equirect", None, f32p, f32p)
    _sig(L, "ref_spherical_to_cartesian", None, f32p, f32p)
    _sig(L, "ref_normalize", None, f32p, f32p)
    _sig(L, "ref_sample_nearest", None, f32p, C.c_uint32, C.c_uint32, f32p, f32p)
    _sig(L, "ref_sample_bilinear", None, f32p, C.c_uint32, C.c_uint32, f32p, f32p)
    if not L.ref_is_test_config():
        _sig(L, "ref_cubemap_basis", None, C.c_uint32, f32p)
        _sig(L, "ref_create_cubemap", None, f32p, C.c_uint32, C.c_uint32, C.c_uint32,
             C.c_uint32, f32p)
        _sig(L, "ref_create_irradiance_cubemap", None, f32p, C.c_uint32, C.c_uint32,
             C.c_uint32, C.c_uint32, f32p)
    _sig(L, "ref_compute_tiles", C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
         C.c_uint32, C.c_uint32, u32p)
    _sig(L, "ref_transform_aabb", None, f32p, f32p, f32p, f32p, f32p, f32p)
    _sig(L, "ref_rotate_vector", None, f32p, f32p, f32p)
    _sig(L, "ref_quat_axis_angle", None, f32p, C.c_float, f32p)
    _sig(L, "ref_quat_mul", None, f32p, f32p, f32p)
    _sig(L, "ref_object_matrices", None, f32p, f32p, f32p, f32p, f32p)
    _sig(L, "ref_icosphere", None, C.c_uint32, f32p, u32p, u32ref, u32ref)
    _sig(L, "ref_plane", None, f32p, u32p, u32ref, u32ref)
    _sig(L, "ref_bvh_create", vp, f32p, f32p, C.c_uint32)
    _sig(L, "ref_bvh_intersect", C.c_uint32, vp, f32p, f32p, C.c_uint32, u32p,
         u32ref, u32ref)
    _sig(L, "ref_bvh_dump", C.c_uint32, vp, f32p, i32p, u32p, C.c_uint32)
    _sig(L, "ref_mesh_create", C.c_uint32, f32p, C.c_uint32, u32p, C.c_uint32, C.c_int)
    _sig(L, "ref_scene_add_object", None, C.c_uint32, C.c_uint32, f32p, f32p, f32p)
    _sig(L, "ref_scene_build", None)
    _sig(L, "ref_scene_intersect", None, f32p, f32p, f32p, u64p)
    _sig(L, "ref_mesh_intersect", None, C.c_uint32, f32p, f32p, f32p, u64p)
    _sig(L, "ref_material_register", C.c_int, f32p, C.c_uint32, f32p, C.c_uint32,
         C.c_float, C.c_uint32)
    _sig(L, "ref_texture_register", C.c_int, f32p, C.c_uint32, C.c_uint32, C.c_uint32)
    _sig(L, "ref_set_background", None, C.c_uint32)
    _sig(L, "ref_compute_radiance_for_path", None, f32p, C.c_uint32, f32p)
    _sig(L, "ref_evaluate_material", None, C.c_uint32, f32p, f32p, f32p)
    _sig(L, "ref_camera_configure", None, C.c_uint32, C.c_uint32, f32p, f32p,
         C.c_float)
    _sig(L, "ref_camera_state", None, f32p)
    _sig(L, "ref_film_position", None, C.c_float, C.c_float, f32p)
    _sig(L, "ref_path_trace_tile", None, C.c_uint32, C.c_uint32, C.c_uint32,
         C.c_uint32, u32ref, C.c_uint32, u64p, f32p)
    _sig(L, "ref_path_trace_pixels", None, C.c_uint32, C.c_uint32, C.c_uint32,
         C.c_uint32, u32p, C.c_uint32, u64p, f32p)
    _refs[path] = L
    return L


def ref():
    """The reference built with its production switches (or None)."""
    return _load_ref(REF_SO)


def ref_test():
    """The reference built with its unit-test switches (or None)."""
    return _load_ref(REF_TEST_SO)


def f32(*a):
    return np.ascontiguousarray(np.array(a, dtype=np.float32).ravel())


def u32(x):
    return C.c_uint32(x)


def wide_boxes(nodes64):
    """Decoded child boxes of exported four-wide nodes ([n, 64] bytes):
    (lo [n,4,3], hi [n,4,3], present [n,4]), each coordinate p + q * s in
    float32 (the product is exact, the sum rounds once)."""
    w = np.ascontiguousarray(nodes64, np.uint8)
    n = len(w)
    p = w[:, 0:12].copy().view(np.float32).reshape(n, 3)
    s = (w[:, 12:15].astype(np.uint32) << 23).view(np.float32)
    refs = w[:, 16:32].copy().view(np.int32).reshape(n, 4)
    q = w[:, 32:56].reshape(n, 6, 4).transpose(0, 2, 1).astype(np.float32)
    lo = p[:, None, :] + q[:, :, 0:3] * s[:, None, :]
    hi = p[:, None, :] + q[:, :, 3:6] * s[:, None, :]
    return lo, hi, refs != 0x7FFFFFFF
