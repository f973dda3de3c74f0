"""The scene and camera of the compute-tracer harness tests (the parked
Vulkan compute shader's materials by id: 0 red, 2 checkerboard, 5 white
light, 6 black)."""
import numpy as np

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes


def compute_scene(light=True, sphere=True, ground=True):
    """An icosphere (red) on a 20x20 checkerboard ground under a 6x6 white
    light quad facing down at height 4, objects with scales >= 1."""
    iv, ii = V.gen_icosphere(2)
    pv, pi = V.gen_plane()
    meshes = [(iv, ii, False), (pv, pi, False)]
    objects = []
    if sphere:
        objects.append((0, 0, (0.0, 1.0, 0.0), (0, 0, 0, 1), (1.0, 1.0, 1.0)))
    if ground:
        objects.append((1, 2, (0.0, 0.0, 0.0),
                        tuple(scenes.quat_axis_angle((1, 0, 0), -np.pi / 2)),
                        (20.0, 20.0, 20.0)))
    if light:
        objects.append((1, 5, (0.0, 4.0, 0.0),
                        tuple(scenes.quat_axis_angle((1, 0, 0), np.pi / 2)),
                        (6.0, 6.0, 6.0)))
    return {"meshes": meshes, "objects": objects, "materials": [], "background": 6,
            "camera": (256, 192, (0.0, 1.5, 5.0),
                       tuple(scenes.quat_axis_angle((1, 0, 0), -0.15)), 0.8)}


def camera_transform(desc):
    """The shader's uniform: a column-major mat4 with columns right, up,
    forward, position (from the reference camera basis, sp_ConfigureCamera)."""
    w, h, pos, q, fd = desc["camera"]
    cam = V.sp_ConfigureCamera(w, h, pos, q, fd)
    m = np.zeros(16, np.float32)
    m[0:3] = cam.right
    m[4:7] = cam.up
    m[8:11] = cam.forward
    m[12:15] = cam.position
    m[15] = 1.0
    return m
