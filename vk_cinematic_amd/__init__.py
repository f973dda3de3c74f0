"""vk_cinematic_amd -- MI355X (gfx950) wavefront path tracer.

Python host bindings over the C ABI of ``libspgpu.so`` (declared in
``include/spgpu.h``).  The names mirror the reference tracer's interface
(deadVertex/vk_cinematic, ``sp_`` prefix): ``sp_ConfigureCamera``,
``sp_PathTraceTile``, ``ComputeTiles`` and a :class:`Scene` that carries the
``sp_CreateMesh`` / ``sp_AddObjectToScene`` / ``sp_BuildSceneBroadphase`` /
``sp_RegisterMaterial`` / ``sp_RegisterTexture`` calls.

All compute runs in the HIP library; there is no CPU fallback.  Importing
raises if the library has not been built (``__graft_entry__.build()``).
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspgpu.so")

SPG_OK = 0
SPG_NO_TEXTURE = 0xFFFFFFFF
SPG_MISS = 0xFFFFFFFF
RNG_XORSHIFT_PIXEL, RNG_PCG = 1, 2
EST_PARITY, EST_UNBIASED, EST_GUIDED, EST_NEE, EST_GUIDED_NEE = 0, 1, 2, 3, 4
TX_DOUBLE, TX_FAST = 1, 2
METRIC_ELAPSED_NS, METRIC_PATHS, METRIC_RAYS, METRIC_HITS, METRIC_MISSES = 0, 1, 2, 3, 4
METRIC_AABB_TESTS, METRIC_MESH_TESTS = 10, 11


class SpgError(RuntimeError):
    """A negative status from the C ABI, with spg_last_error()'s message."""

    def __init__(self, code, msg):
        super().__init__(f"spgpu error {code}: {msg}")
        self.code = code


class Tile(C.Structure):
    """Tile, src/tile.h:3-9."""
    _fields_ = [("minX", C.c_uint32), ("minY", C.c_uint32),
                ("maxX", C.c_uint32), ("maxY", C.c_uint32)]


class Metrics(C.Structure):
    """sp_Metrics, src/sp_metrics.h:45-48."""
    _fields_ = [("values", C.c_uint64 * 12)]


class Material(C.Structure):
    """sp_Material, src/sp_material_system.h:3-12."""
    _fields_ = [("albedo", C.c_float * 3), ("albedoTexture", C.c_uint32),
                ("emission", C.c_float * 3), ("emissionTexture", C.c_uint32),
                ("roughness", C.c_float)]


class Camera(C.Structure):
    """sp_Camera, src/simd_path_tracer.h:17-32."""
    _fields_ = [("right", C.c_float * 3), ("up", C.c_float * 3),
                ("forward", C.c_float * 3), ("position", C.c_float * 3),
                ("filmCenter", C.c_float * 3), ("width", C.c_uint32),
                ("height", C.c_uint32), ("halfPixelWidth", C.c_float),
                ("halfPixelHeight", C.c_float), ("halfFilmWidth", C.c_float),
                ("halfFilmHeight", C.c_float)]


class TraceConfig(C.Structure):
    _fields_ = [("samplesPerPixel", C.c_uint32), ("bounces", C.c_uint32),
                ("radianceClamp", C.c_float), ("estimator", C.c_uint32),
                ("rngMode", C.c_uint32), ("seed", C.c_uint32),
                ("transcendentals", C.c_uint32), ("sampleOffset", C.c_uint32),
                ("collectMetrics", C.c_uint32), ("guideTrain", C.c_uint32),
                ("guide", C.c_void_p)]


class BatchItem(C.Structure):
    """spg_BatchItem (include/spgpu.h)."""
    _fields_ = [("camera", C.POINTER(Camera)), ("scene", C.c_void_p),
                ("deviceImage", C.c_void_p)]


class BatchTuning(C.Structure):
    """spg_BatchTuning (include/spgpu.h)."""
    _fields_ = [("pathsInFlight", C.c_uint32), ("extendBlocks", C.c_uint32),
                ("candidates", C.c_uint32), ("tunedMsPerSample", C.c_float),
                ("pixels", C.c_uint64), ("graphNodes", C.c_uint32),
                ("refill", C.c_uint32), ("cached", C.c_uint32)]


class GuideConfig(C.Structure):
    """spg_GuideConfig (include/spgpu.h)."""
    _fields_ = [("seed", C.c_uint32), ("learningRate", C.c_float),
                ("batchSize", C.c_uint32), ("stepsPerPass", C.c_uint32),
                ("maxWeight", C.c_float), ("bsdfFraction", C.c_float)]


class Context(C.Structure):
    """sp_Context, src/simd_path_tracer.h:34-46."""
    _fields_ = [("camera", C.POINTER(Camera)), ("scene", C.c_void_p),
                ("config", TraceConfig)]


class BuildOpts(C.Structure):
    _fields_ = [("maxLeafSize", C.c_uint32), ("sahBins", C.c_uint32),
                ("threads", C.c_uint32)]


class TraversalStats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("boxTests", C.c_uint64),
                ("triTests", C.c_uint64), ("nodeFetches", C.c_uint64),
                ("hits", C.c_uint64)]


_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_int = C.c_int
_u32 = C.c_uint32

_SIGS = {
    "spg_last_error": (C.c_char_p, []),
    "spg_device_create": (_int, [_int, C.POINTER(_vp)]),
    "spg_device_destroy": (_int, [_vp]),
    "spg_device_synchronize": (_int, [_vp]),
    "spg_device_alloc": (_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "spg_device_free": (_int, [_vp, _vp]),
    "spg_copy_to_device": (_int, [_vp, _vp, _vp, C.c_size_t]),
    "spg_copy_to_host": (_int, [_vp, _vp, _vp, C.c_size_t]),
    "spg_scene_create": (_int, [_vp, C.POINTER(_vp)]),
    "spg_scene_destroy": (_int, [_vp]),
    "spg_mesh_create": (_int, [_vp, _f32p, _u32, _u32p, _u32, _int, C.POINTER(_u32)]),
    "spg_scene_add_object": (_int, [_vp, _u32, _u32, _f32p, _f32p, _f32p,
                                    C.POINTER(_u32)]),
    "spg_scene_build": (_int, [_vp, C.POINTER(BuildOpts)]),
    "spg_bvh_build_host": (_int, [_f32p, _u32, C.POINTER(BuildOpts), _f32p, _u32,
                                  _u32p, C.POINTER(_u32), C.POINTER(_u32)]),
    "spg_bvh_build_wide_host": (_int, [_f32p, _u32, C.POINTER(BuildOpts), _u8p, _u32,
                                       _u32p, C.POINTER(_u32), C.POINTER(_u32)]),
    "spg_mesh_wide_export": (_int, [_vp, _u32, _vp, _u32, _vp, _u32, C.POINTER(_u32),
                                    C.POINTER(_u32), C.POINTER(_u32)]),
    "spg_mesh_bvh_info": (_int, [_vp, _u32, C.POINTER(_u32), C.POINTER(_u32),
                                 C.POINTER(_u32), C.POINTER(C.c_double)]),
    "spg_mesh_bvh_export": (_int, [_vp, _u32, _f32p, _u32, _u32p, _u32,
                                   C.POINTER(C.c_int32)]),
    "spg_material_register": (_int, [_vp, C.POINTER(Material), _u32]),
    "spg_texture_register": (_int, [_vp, _f32p, _u32, _u32, _u32]),
    "spg_cubemap_register": (_int, [_vp, _f32p, _u32, _u32, _u32]),
    "spg_set_background_material": (_int, [_vp, _u32]),
    "spg_camera_configure": (_int, [C.POINTER(Camera), _u32, _u32, _f32p, _f32p,
                                    C.c_float]),
    "spg_compute_tiles": (_int, [_u32, _u32, _u32, _u32, C.POINTER(Tile), _u32,
                                 C.POINTER(_u32)]),
    "spg_path_trace_tile": (_int, [C.POINTER(Context), Tile, C.POINTER(Metrics),
                                   _f32p]),
    "spg_path_trace_tiles": (_int, [C.POINTER(Context), C.POINTER(Tile), _u32,
                                    C.POINTER(Metrics), _f32p]),
    "spg_path_trace_tiles_device": (_int, [C.POINTER(Context), C.POINTER(Tile), _u32,
                                           C.POINTER(Metrics), _vp]),
    "spg_intersect_closest": (_int, [_vp, _u32, _f32p, _f32p, _f32p, _f32p, _f32p,
                                     _f32p, _u32, _u32p, _f32p, _vp, _vp]),
    "spg_intersect_closest_device": (_int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp,
                                            _u32, _vp, _vp, _vp, _vp,
                                            C.POINTER(TraversalStats)]),
    "spg_intersect_closest_device_noalloc": (_int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp,
                                                    _vp, _u32, _vp, _vp, _vp, _vp]),
    "spg_intersect_scene": (_int, [_vp, _f32p, _u32, _f32p, _vp, _vp]),
    "spg_time_intersect_closest": (_int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp,
                                          _u32, _vp, _vp, _u32, C.POINTER(C.c_double)]),
    "spg_last_trace_timings": (_int, [_vp, C.POINTER(C.c_double), C.POINTER(_u32)]),
    "spg_last_trace_counters": (_int, [_vp, C.POINTER(C.c_uint64)]),
    "spg_gen_icosphere": (_int, [_u32, _f32p, _u32, _u32p, _u32, C.POINTER(_u32),
                                 C.POINTER(_u32)]),
    "spg_gen_plane": (_int, [_f32p, _u32p]),
    "spg_gen_triangle_soup": (_int, [C.c_uint64, _u32, C.c_float, C.c_float, _f32p]),
    "spg_gen_sky_cubemap": (_int, [_u32, _f32p, C.c_float, C.c_float, _f32p]),
    "spg_gen_sky_equirect": (_int, [_u32, _u32, _f32p, C.c_float, C.c_float, _f32p]),
    "spg_accumulate_device": (_int, [_vp, _vp, _vp, _u32, _u32, _u32]),
    "spg_tonemap_device": (_int, [_vp, _vp, _u32, C.c_float, _int, _vp]),
    "spg_env_table_build": (_int, [_f32p, _u32, _u32, _f32p, _f32p, _u32p]),
    "spg_sincos_host": (_int, [_f32p, _u32, _f32p, _f32p]),
    "spg_time_post": (_int, [_vp, _vp, _vp, _u32, C.c_float, _int, _vp, _u32,
                             C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "spg_write_pfm": (_int, [C.c_char_p, _u32, _u32, _f32p]),
    "spg_obj_parse": (_int, [C.c_char_p, C.c_size_t, _vp, _u32, _vp, _u32,
                             C.POINTER(_u32), C.POINTER(_u32)]),
    "spg_create_irradiance_cubemap": (_int, [_vp, _f32p, _u32, _u32, _u32, _u32,
                                             C.c_float, _int, _f32p,
                                             C.POINTER(C.c_double)]),
    "spg_pixel_seed": (_u32, [_u32, _u32]),
    "spg_batch_create": (_int, [_vp, _vp, _u32, _vp, _int, C.POINTER(_vp)]),
    "spg_batch_run": (_int, [_vp, _vp]),
    "spg_batch_tuning": (_int, [_vp, _vp]),
    "spg_batch_last_ms": (_int, [_vp, C.POINTER(C.c_double)]),
    "spg_batch_destroy": (_int, [_vp]),
    "spg_batch_profile": (_int, [_vp, _int, C.POINTER(C.c_double), C.POINTER(_u32),
                                 C.POINTER(C.c_uint64)]),
    "spg_guide_default_config": (None, [_vp]),
    "spg_guide_create": (_int, [_vp, _vp, C.POINTER(_vp)]),
    "spg_guide_destroy": (_int, [_vp]),
    "spg_guide_get_params": (_int, [_vp, _f32p, _u32]),
    "spg_guide_set_params": (_int, [_vp, _f32p, _u32]),
    "spg_guide_evaluate": (_int, [_vp, _f32p, _u32, _vp, _vp]),
    "spg_guide_gradient": (_int, [_vp, _f32p, _u32, _int, C.c_float, _f32p,
                                  C.POINTER(C.c_float)]),
    "spg_guide_train": (_int, [_vp, _f32p, _u32, _u32, C.POINTER(C.c_float)]),
    "spg_guide_stats": (_int, [_vp, C.POINTER(_u32), C.POINTER(_u32),
                               C.POINTER(C.c_float)]),
    "spg_compute_tracer_tiles": (_int, [_vp, _f32p, _u32p, _u32, _f32p]),
    "spg_check_arithmetic": (_int, [_vp, C.POINTER(C.c_uint64)]),
    "spg_guide_samples": (_int, [_vp, _vp, _u32, C.POINTER(_u32)]),
    "spg_guide_debug": (_int, [_vp, _u32, _u32]),
    "spg_guide_debug_records": (_int, [_vp, _vp, _u32, C.POINTER(_u32)]),
    "spg_guide_save": (_int, [_vp, C.c_char_p]),
    "spg_guide_load": (_int, [_vp, C.c_char_p]),
    "spg_guide_light_split": (_int, [_vp, C.POINTER(C.c_float)]),
    "spg_scene_learn_light_split": (_int, [_vp, C.c_uint32]),
    "spg_scene_light_split": (_int, [_vp, C.POINTER(C.c_float)]),
    "spg_guide_set_light_split": (_int, [_vp, C.c_float]),
}

_lib = None
_live_scenes = weakref.WeakSet()
_live_devices = weakref.WeakSet()


@atexit.register
def _release_all():
    # Release GPU objects while the HIP runtime is still up: scenes and
    # guides first, then devices (interpreter teardown would otherwise order
    # them freely).
    live = list(_live_scenes)
    for s in [x for x in live if isinstance(x, Batch)]:  # batches pin guides
        s.close()
    for s in live:
        s.close()
    for d in list(_live_devices):
        d.close()


def lib():
    """The loaded libspgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(code):
    if code != SPG_OK:
        raise SpgError(code, lib().spg_last_error().decode(errors="replace"))
    return code


def _f32(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel())
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} floats, got {a.size}")
    return a


def _u32a(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32).ravel())


class Device:
    """One GPU: a HIP stream and the wavefront path buffers."""

    def __init__(self, ordinal=0):
        h = C.c_void_p()
        check(lib().spg_device_create(ordinal, C.byref(h)))
        self.handle = h
        _live_devices.add(self)

    def synchronize(self):
        check(lib().spg_device_synchronize(self.handle))

    def alloc(self, nbytes):
        p = C.c_void_p()
        check(lib().spg_device_alloc(self.handle, nbytes, C.byref(p)))
        return DeviceBuffer(self, p, nbytes)

    def upload(self, array):
        a = np.ascontiguousarray(array)
        buf = self.alloc(a.nbytes)
        check(lib().spg_copy_to_device(self.handle, buf.ptr, a.ctypes.data, a.nbytes))
        return buf

    def create_irradiance_cubemap(self, equirect, face_width, face_height,
                                  clamp=10.0, transcendentals=TX_DOUBLE):
        """CreateIrradianceCubeMap (src/cubemap.cpp:108-233) of an RGBA32F
        equirect image [h][w][4]; returns the faces [6][fh][fw][4] and the
        kernel's milliseconds."""
        img = np.ascontiguousarray(equirect, np.float32)
        eh, ew = img.shape[0], img.shape[1]
        out = np.zeros(6 * face_width * face_height * 4, np.float32)
        ms = C.c_double()
        check(lib().spg_create_irradiance_cubemap(
            self.handle, img.ravel(), ew, eh, face_width, face_height, clamp,
            transcendentals, out, C.byref(ms)))
        return out.reshape(6, face_height, face_width, 4), ms.value

    def accumulate(self, dst, src, pixels, before, added):
        """dst = (dst * before + src * added) / (before + added) on device
        RGBA32F buffers (progressive rendering across trace calls)."""
        check(lib().spg_accumulate_device(self.handle, dst.ptr, src.ptr, pixels,
                                          before, added))

    def tonemap(self, src, pixels, exposure=1.0, transcendentals=TX_DOUBLE):
        """The viewer's tone mapping of a device RGBA32F image; returns
        RGBA8 [pixels][4] on the host."""
        out = self.alloc(4 * pixels)
        check(lib().spg_tonemap_device(self.handle, src.ptr, pixels, exposure,
                                       transcendentals, out.ptr))
        px = out.download(np.uint8, 4 * pixels).reshape(-1, 4)
        out.free()
        return px

    def last_timings(self):
        ms = (C.c_double * 6)()
        n = (C.c_uint32 * 6)()
        check(lib().spg_last_trace_timings(self.handle, ms, n))
        names = ["raygen", "extend", "shade", "finalize", "guide", "total"]
        return {k: (ms[i], n[i]) for i, k in enumerate(names)}

    def last_counters(self):
        """rays, hits, misses, mesh-hierarchy slab tests, mesh (instance)
        entries, triangle tests, instance-hierarchy slab tests and guide
        network evaluations of the last trace call (test counts only when it
        collected metrics)."""
        c = (C.c_uint64 * 8)()
        check(lib().spg_last_trace_counters(self.handle, c))
        return dict(zip(["rays", "hits", "misses", "slab_tests", "mesh_tests",
                         "tri_tests", "tlas_slab_tests", "guide_evals"], list(c)))

    def close(self):
        if self.handle:
            lib().spg_device_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    def __init__(self, dev, ptr, nbytes):
        self.dev, self.ptr, self.nbytes = dev, ptr, nbytes

    def upload(self, array):
        """Copy a host array into the start of this buffer."""
        a = np.ascontiguousarray(array)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        check(lib().spg_copy_to_device(self.dev.handle, self.ptr, a.ctypes.data,
                                       a.nbytes))

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        check(lib().spg_copy_to_host(self.dev.handle, out.ctypes.data, self.ptr,
                                     out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().spg_device_free(self.dev.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Scene:
    """sp_Scene + sp_MaterialSystem (src/sp_scene.h:16-29,
    src/sp_material_system.h:35-46) resident in HBM."""

    def __init__(self, device):
        self.device = device
        h = C.c_void_p()
        check(lib().spg_scene_create(device.handle, C.byref(h)))
        self.handle = h
        _live_scenes.add(self)

    # sp_CreateMesh + sp_BuildMeshMidphase (src/sp_scene.cpp:8-75)
    def create_mesh(self, verts8, indices, smooth=True):
        v = _f32(verts8)
        i = _u32a(indices)
        mid = C.c_uint32()
        check(lib().spg_mesh_create(self.handle, v, v.size // 8, i, i.size,
                                    int(bool(smooth)), C.byref(mid)))
        return mid.value

    # sp_AddObjectToScene (src/sp_scene.cpp:77-117)
    def add_object(self, mesh, material, position, orientation=(0, 0, 0, 1),
                   scale=(1, 1, 1)):
        oid = C.c_uint32()
        check(lib().spg_scene_add_object(self.handle, mesh, material,
                                         _f32(position, 3), _f32(orientation, 4),
                                         _f32(scale, 3), C.byref(oid)))
        return oid.value

    # sp_BuildSceneBroadphase (src/sp_scene.cpp:119-125) + midphase builds
    def build(self, max_leaf=4, bins=16, threads=0):
        opts = BuildOpts(max_leaf, bins, threads)
        check(lib().spg_scene_build(self.handle, C.byref(opts)))

    def learn_light_split(self, enable=True):
        """NEE calls on this scene learn the sky share of their light
        samples (enable=False stops and forgets it)."""
        check(lib().spg_scene_learn_light_split(self.handle, int(bool(enable))))

    def light_split(self):
        f = C.c_float()
        check(lib().spg_scene_light_split(self.handle, C.byref(f)))
        return f.value

    def bvh_info(self, mesh):
        n, l, d, c = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_double()
        check(lib().spg_mesh_bvh_info(self.handle, mesh, C.byref(n), C.byref(l),
                                      C.byref(d), C.byref(c)))
        return {"nodes": n.value, "leaves": l.value, "max_depth": d.value,
                "sah_cost": c.value}

    def bvh_export(self, mesh, n_nodes, n_prims):
        nodes = np.zeros(16 * n_nodes, np.float32)
        prims = np.zeros(n_prims, np.uint32)
        root = C.c_int32()
        check(lib().spg_mesh_bvh_export(self.handle, mesh, nodes, n_nodes, prims,
                                        n_prims, C.byref(root)))
        return nodes.reshape(-1, 16), prims

    def wide_counts(self, mesh):
        """(four-wide nodes, triangles) of a built mesh."""
        cnt, nprim = C.c_uint32(), C.c_uint32()
        check(lib().spg_mesh_wide_export(self.handle, mesh, None, 0, None, 0,
                                         C.byref(cnt), C.byref(nprim), None))
        return cnt.value, nprim.value

    def wide_export(self, mesh):
        """The four-wide hierarchy the traversal walks: (nodes [n,64]
        uint8, primitive order, stack bound)."""
        cnt, nprim, pend = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().spg_mesh_wide_export(self.handle, mesh, None, 0, None, 0,
                                         C.byref(cnt), C.byref(nprim), C.byref(pend)))
        nodes = np.zeros(64 * cnt.value, np.uint8)
        prims = np.zeros(nprim.value, np.uint32)
        check(lib().spg_mesh_wide_export(self.handle, mesh, nodes.ctypes.data, cnt.value,
                                         prims.ctypes.data, nprim.value, None, None, None))
        return nodes.reshape(-1, 64), prims, pend.value

    # sp_RegisterMaterial (src/sp_material_system.cpp:1-13)
    def register_material(self, id, albedo=(0, 0, 0), emission=(0, 0, 0),
                          roughness=0.0, albedo_texture=SPG_NO_TEXTURE,
                          emission_texture=SPG_NO_TEXTURE):
        m = Material((C.c_float * 3)(*albedo), albedo_texture,
                     (C.c_float * 3)(*emission), emission_texture, roughness)
        check(lib().spg_material_register(self.handle, C.byref(m), id))

    # sp_RegisterTexture (src/sp_material_system.cpp:43-57)
    def register_texture(self, id, rgba, width, height):
        check(lib().spg_texture_register(self.handle, _f32(rgba, width * height * 4),
                                         width, height, id))

    def register_cubemap(self, id, rgba6, size):
        check(lib().spg_cubemap_register(self.handle, _f32(rgba6, 6 * size * size * 4),
                                         size, size, id))

    def set_background(self, id):
        check(lib().spg_set_background_material(self.handle, id))

    def intersect_closest(self, mesh, origins, directions, want_uv=False):
        """Closest hit of rays in mesh space; returns (prim, t[, u, v])."""
        o = np.asarray(origins, np.float32).reshape(-1, 3)
        d = np.asarray(directions, np.float32).reshape(-1, 3)
        n = o.shape[0]
        cols = [np.ascontiguousarray(x) for x in (o[:, 0], o[:, 1], o[:, 2],
                                                 d[:, 0], d[:, 1], d[:, 2])]
        prim = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        check(lib().spg_intersect_closest(self.handle, mesh, *cols, n, prim, t,
                                          u.ctypes.data, v.ctypes.data))
        return (prim, t, u, v) if want_uv else (prim, t)

    def intersect_scene(self, origins, directions):
        """sp_RayIntersectScene for a batch: dict of t, material, normal, uv,
        object, prim."""
        o = np.asarray(origins, np.float32).reshape(-1, 3)
        d = np.asarray(directions, np.float32).reshape(-1, 3)
        rays = np.ascontiguousarray(np.concatenate([o, d], axis=1).ravel())
        n = o.shape[0]
        out = np.zeros(7 * n, np.float32)
        obj = np.zeros(n, np.uint32)
        prim = np.zeros(n, np.uint32)
        check(lib().spg_intersect_scene(self.handle, rays, n, out, obj.ctypes.data,
                                        prim.ctypes.data))
        out = out.reshape(n, 7)
        return {"t": out[:, 0].copy(), "material": out[:, 1].view(np.uint32).copy(),
                "normal": out[:, 2:5].copy(), "uv": out[:, 5:7].copy(),
                "object": obj, "prim": prim}

    def close(self):
        if self.handle:
            lib().spg_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- reference-named entry points -------------------------------------
def sp_ConfigureCamera(width, height, position, rotation, film_distance):
    """sp_ConfigureCamera (src/simd_path_tracer.cpp:1-36); returns a Camera."""
    cam = Camera()
    check(lib().spg_camera_configure(C.byref(cam), width, height,
                                     _f32(position, 3), _f32(rotation, 4),
                                     float(film_distance)))
    return cam


def ComputeTiles(width, height, tile_width, tile_height, max_tiles=4096):
    """ComputeTiles (src/tile.h:11-42); returns a list of Tile."""
    arr = (Tile * max(max_tiles, 1))()
    n = C.c_uint32()
    check(lib().spg_compute_tiles(width, height, tile_width, tile_height, arr,
                                  max_tiles, C.byref(n)))
    return [Tile(t.minX, t.minY, t.maxX, t.maxY) for t in arr[: n.value]]


def make_context(camera, scene, spp=1, bounces=3, clamp=10.0,
                 estimator=EST_PARITY, rng=RNG_PCG, seed=0x5EED,
                 transcendentals=TX_DOUBLE, sample_offset=0, collect=False,
                 guide=None, train=False):
    cfg = TraceConfig(spp, bounces, clamp, estimator, rng, seed, transcendentals,
                      sample_offset, int(bool(collect)), int(bool(train)),
                      guide.handle if guide is not None else None)
    ctx = Context(C.pointer(camera), scene.handle, cfg)
    ctx._keep = (camera, scene, guide)
    return ctx


class Batch:
    """Many scenes rendered as one wavefront from a captured HIP graph
    (spg_batch_*).  items: [(camera, scene, DeviceBuffer image)]; the
    remaining arguments as make_context's."""

    def __init__(self, device, items, spp=16, bounces=5, clamp=0.0,
                 estimator=EST_UNBIASED, rng=RNG_PCG, seed=0x5EED,
                 transcendentals=TX_FAST, sample_offset=0, collect=False,
                 autotune=True, guide=None, train=False):
        self.device = device
        arr = (BatchItem * len(items))()
        for k, (cam, scene, img) in enumerate(items):
            arr[k] = BatchItem(C.pointer(cam), scene.handle, img.ptr)
        if guide is not None and estimator == EST_UNBIASED:
            estimator = EST_GUIDED
        cfg = TraceConfig(spp, bounces, clamp, estimator, rng, seed, transcendentals,
                          sample_offset, int(bool(collect)), int(bool(train)),
                          guide.handle if guide is not None else None)
        h = C.c_void_p()
        check(lib().spg_batch_create(device.handle, C.byref(arr), len(items),
                                     C.byref(cfg), int(bool(autotune)), C.byref(h)))
        self.handle = h
        self._keep = (items, guide)
        _live_scenes.add(self)

    def run(self, metrics=None):
        m = metrics if metrics is not None else Metrics()
        check(lib().spg_batch_run(self.handle, C.byref(m)))
        return m

    def tuning(self):
        t = BatchTuning()
        check(lib().spg_batch_tuning(self.handle, C.byref(t)))
        return {f: getattr(t, f) for f, _ in BatchTuning._fields_}

    def last_ms(self):
        ms = C.c_double()
        check(lib().spg_batch_last_ms(self.handle, C.byref(ms)))
        return ms.value

    def profile(self, collect=False):
        """Eager render with per-class launch timings and counters."""
        ms = (C.c_double * 6)()
        n = (C.c_uint32 * 6)()
        c = (C.c_uint64 * 8)()
        check(lib().spg_batch_profile(self.handle, int(bool(collect)), ms, n, c))
        names = ["raygen", "extend", "shade", "finalize", "guide", "total"]
        return ({k: (ms[i], n[i]) for i, k in enumerate(names)},
                dict(zip(["rays", "hits", "misses", "slab_tests", "mesh_tests",
                          "tri_tests", "tlas_slab_tests"], list(c))))

    def close(self):
        if self.handle:
            lib().spg_batch_destroy(self.handle)
            self.handle = None
        _live_scenes.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


GUIDE_PARAMS = 19760
GUIDE_OUTPUTS = 48
GUIDE_LOBE_FLOATS = 40
GUIDE_SAMPLE_FLOATS = 13
GUIDE_ANALYTIC_MIXTURE = 1
GUIDE_RECORD_FLOATS = 64


class Guide:
    """A learned path guide (spg_guide_*): MLP -> 8-lobe vMF mixture,
    trained online by the tracer when a context sets train=True."""

    def __init__(self, device, seed=1, learning_rate=3e-3, batch_size=16384,
                 steps_per_pass=4, max_weight=50.0, bsdf_fraction=0.5):
        self.device = device
        cfg = GuideConfig(seed, learning_rate, batch_size, steps_per_pass,
                          max_weight, bsdf_fraction)
        h = C.c_void_p()
        check(lib().spg_guide_create(device.handle, C.byref(cfg), C.byref(h)))
        self.handle = h
        _live_scenes.add(self)

    def params(self):
        p = np.zeros(GUIDE_PARAMS, np.float32)
        check(lib().spg_guide_get_params(self.handle, p, GUIDE_PARAMS))
        return p

    def set_params(self, p):
        check(lib().spg_guide_set_params(self.handle, _f32(p, GUIDE_PARAMS),
                                         GUIDE_PARAMS))

    def evaluate(self, inputs9):
        """Raw outputs (n x 48) and mixtures (n x 40) for rows pn, n, v."""
        x = _f32(inputs9)
        n = x.size // 9
        out = np.zeros((n, GUIDE_OUTPUTS), np.float32)
        lobes = np.zeros((n, GUIDE_LOBE_FLOATS), np.float32)
        check(lib().spg_guide_evaluate(self.handle, x, n, out.ctypes.data,
                                       lobes.ctypes.data))
        return out, lobes

    def gradient(self, samples13, normalize=False, max_weight=0.0):
        x = _f32(samples13)
        g = np.zeros(GUIDE_PARAMS, np.float32)
        loss = C.c_float()
        check(lib().spg_guide_gradient(self.handle, x, x.size // GUIDE_SAMPLE_FLOATS,
                                       int(bool(normalize)), max_weight, g,
                                       C.byref(loss)))
        return g, loss.value

    def train(self, samples13, steps=1):
        x = _f32(samples13)
        loss = C.c_float()
        check(lib().spg_guide_train(self.handle, x, x.size // GUIDE_SAMPLE_FLOATS,
                                    steps, C.byref(loss)))
        return loss.value

    def samples(self):
        """The training samples of the last traced pass, in training order
        (n x 13: pn, n, v, wi, weight)."""
        n = _u32()
        check(lib().spg_guide_samples(self.handle, None, 0, C.byref(n)))
        out = np.zeros((n.value, GUIDE_SAMPLE_FLOATS), np.float32)
        if n.value:
            check(lib().spg_guide_samples(self.handle, out.ctypes.data, n.value,
                                          C.byref(n)))
        return out

    def debug(self, analytic=False, record=0):
        """Verification hooks: the analytic mixture instead of the network;
        record the first `record` scatter events of the next traces."""
        check(lib().spg_guide_debug(self.handle, GUIDE_ANALYTIC_MIXTURE if analytic
                                    else 0, record))

    def records(self):
        """Recorded scatter events (n x 64, see spg_guide_debug)."""
        n = _u32()
        check(lib().spg_guide_debug_records(self.handle, None, 0, C.byref(n)))
        out = np.zeros((n.value, GUIDE_RECORD_FLOATS), np.float32)
        if n.value:
            check(lib().spg_guide_debug_records(self.handle, out.ctypes.data, n.value,
                                                C.byref(n)))
        return out

    def save(self, path):
        """Checkpoint: configuration, parameters, Adam moments and steps."""
        check(lib().spg_guide_save(self.handle, os.fsencode(path)))

    def light_split(self):
        """The learned sky share of guided + NEE light samples (-1 before
        any training pass delivered light)."""
        f = C.c_float()
        check(lib().spg_guide_light_split(self.handle, C.byref(f)))
        return f.value

    def set_light_split(self, sky_fraction):
        check(lib().spg_guide_set_light_split(self.handle, float(sky_fraction)))

    def load(self, path):
        """Restores a checkpoint written by save (its configuration too)."""
        check(lib().spg_guide_load(self.handle, os.fsencode(path)))

    def stats(self):
        c, st, loss = C.c_uint32(), C.c_uint32(), C.c_float()
        check(lib().spg_guide_stats(self.handle, C.byref(c), C.byref(st),
                                    C.byref(loss)))
        return {"collected": c.value, "steps": st.value, "loss": loss.value}

    def close(self):
        if self.handle:
            lib().spg_guide_destroy(self.handle)
            self.handle = None
        _live_scenes.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sp_PathTraceTile(ctx, tile, pixels, metrics=None):
    """sp_PathTraceTile (src/simd_path_tracer.cpp:178-345): overwrites the
    tile's pixels of `pixels` (float32, width*height*4)."""
    if pixels.dtype != np.float32 or not pixels.flags.c_contiguous:
        raise ValueError("pixels must be a C-contiguous float32 array")
    m = metrics if metrics is not None else Metrics()
    check(lib().spg_path_trace_tile(C.byref(ctx), tile, C.byref(m), pixels.ravel()))
    return m


def path_trace_tiles(ctx, tiles, pixels, metrics=None):
    """All tiles in one wavefront batch (the work queue of
    src/main.cpp:819-844 in one call)."""
    arr = (Tile * len(tiles))(*tiles)
    m = metrics if metrics is not None else Metrics()
    check(lib().spg_path_trace_tiles(C.byref(ctx), arr, len(tiles), C.byref(m),
                                     pixels.ravel()))
    return m


def path_trace_tiles_device(ctx, tiles, device_image, metrics=None):
    arr = (Tile * len(tiles))(*tiles)
    m = metrics if metrics is not None else Metrics()
    check(lib().spg_path_trace_tiles_device(C.byref(ctx), arr, len(tiles),
                                            C.byref(m), device_image.ptr))
    return m


def env_table_build(cube_rgba, size):
    """The next-event estimation table of a size x size x 6 RGBA32F cube
    map: (pdf per texel, alias probabilities, alias indices)."""
    c = _f32(cube_rgba)
    n = 6 * size * size
    pdf, prob = np.zeros(n, np.float32), np.zeros(n, np.float32)
    alias = np.zeros(n, np.uint32)
    check(lib().spg_env_table_build(c, size, size, pdf, prob, alias))
    return pdf, prob, alias


def bvh_build_wide_host(tris9, max_leaf=4, bins=16, threads=0):
    """The same build collapsed to four-wide quantized nodes; returns
    (nodes [n,64] uint8, leaf primitive order, stack bound)."""
    t = _f32(tris9)
    n = t.size // 9
    cap = n + 2
    nodes = np.zeros(64 * cap, np.uint8)
    order = np.zeros(n, np.uint32)
    cnt, pend = C.c_uint32(), C.c_uint32()
    opts = BuildOpts(max_leaf, bins, threads)
    check(lib().spg_bvh_build_wide_host(t, n, C.byref(opts), nodes, cap, order,
                                        C.byref(cnt), C.byref(pend)))
    return nodes[: 64 * cnt.value].reshape(-1, 64), order, pend.value


def bvh_build_host(tris9, max_leaf=4, bins=16, threads=0):
    """SAH BVH2 of a triangle list on the host (no device); returns
    (nodes [n,16] float32, leaf primitive order, max depth)."""
    t = _f32(tris9)
    n = t.size // 9
    cap = 2 * n + 2
    nodes = np.zeros(16 * cap, np.float32)
    order = np.zeros(n, np.uint32)
    cnt, depth = C.c_uint32(), C.c_uint32()
    opts = BuildOpts(max_leaf, bins, threads)
    check(lib().spg_bvh_build_host(t, n, C.byref(opts), nodes, cap, order,
                                   C.byref(cnt), C.byref(depth)))
    return nodes[: 16 * cnt.value].reshape(-1, 16), order, depth.value


# ---- generators ----------------------------------------------------------
def gen_icosphere(level):
    faces = 20 * 4 ** level
    verts = 12 + sum(3 * 20 * 4 ** l for l in range(level))
    v = np.zeros(8 * verts, np.float32)
    i = np.zeros(3 * faces, np.uint32)
    nv, ni = C.c_uint32(), C.c_uint32()
    check(lib().spg_gen_icosphere(level, v, verts, i, 3 * faces, C.byref(nv),
                                  C.byref(ni)))
    return v[: 8 * nv.value].reshape(-1, 8), i[: ni.value]


def gen_plane():
    v = np.zeros(32, np.float32)
    i = np.zeros(6, np.uint32)
    check(lib().spg_gen_plane(v, i))
    return v.reshape(4, 8), i


def gen_triangle_soup(n, seed=1, mu=-5.0, sigma=0.5):
    t = np.zeros(9 * n, np.float32)
    check(lib().spg_gen_triangle_soup(seed, n, mu, sigma, t))
    return t.reshape(n, 9)


def gen_sky_cubemap(size=64, sun=(0.4, 0.75, -0.5), sun_radius=0.0523599,
                    sun_radiance=50.0):
    out = np.zeros(6 * size * size * 4, np.float32)
    check(lib().spg_gen_sky_cubemap(size, _f32(sun, 3), sun_radius, sun_radiance,
                                    out))
    return out


def parse_obj(text):
    """LoadMesh (src/mesh.cpp:5-62) for OBJ text: (verts [n][8] VertexPNT,
    indices [3m]) as spg_obj_parse triangulates and joins them."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    nv, ni = _u32(), _u32()
    check(lib().spg_obj_parse(data, len(data), None, 0, None, 0, C.byref(nv), C.byref(ni)))
    v = np.zeros(nv.value * 8, np.float32)
    i = np.zeros(ni.value, np.uint32)
    check(lib().spg_obj_parse(data, len(data), v.ctypes.data, nv.value, i.ctypes.data,
                              ni.value, C.byref(nv), C.byref(ni)))
    return v.reshape(-1, 8), i


def write_pfm(path, width, height, rgba):
    """PFM file of a host RGBA32F image (row 0 at the top)."""
    check(lib().spg_write_pfm(str(path).encode(), width, height, _f32(rgba)))


def load_obj(path):
    with open(path, "rb") as f:
        return parse_obj(f.read())


def gen_sky_equirect(width=256, height=128, sun=(0.4, 0.75, -0.5),
                     sun_radius=0.0523599, sun_radiance=50.0):
    out = np.zeros(width * height * 4, np.float32)
    check(lib().spg_gen_sky_equirect(width, height, _f32(sun, 3), sun_radius,
                                     sun_radiance, out))
    return out


def pixel_seed(base, index):
    return lib().spg_pixel_seed(base, index)
