"""A mesh ingested from OBJ text (spg_obj_parse) rendered on the GPU equals
the oracle's render of the same triangles bit for bit (config 1 with the
icosphere replaced by the parsed mesh: a displaced icosphere written as
text, plus a pentagon face)."""
import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes
from _libs import SPO_RNG_PCG
from test_gpu_path_trace import gpu_render, oracle_render

pytestmark = pytest.mark.gpu


def _obj_text():
    v, idx = scenes.displaced_icosphere(2, seed=5)[:2]
    v = np.asarray(v, np.float32).reshape(-1, 8)
    lines = ["# generated", "o mesh"]
    lines += ["v %.9g %.9g %.9g" % tuple(p[:3]) for p in v]
    lines += ["vn %.9g %.9g %.9g" % tuple(p[3:6]) for p in v]
    lines += ["f %d//%d %d//%d %d//%d" % (a + 1, a + 1, b + 1, b + 1, c + 1, c + 1)
              for a, b, c in np.asarray(idx).reshape(-1, 3)]
    # a ground pentagon below the sphere (fan-triangulated by the parser)
    n0 = len(v)
    box = [(-1.5, -0.2, -1.5), (1.5, -0.2, -1.5), (1.5, -0.2, 1.5), (-1.5, -0.2, 1.5),
           (0.0, -0.2, 1.5)]
    lines += ["v %.9g %.9g %.9g" % p for p in box]
    lines += ["vn 0 1 0"]
    up = n0 + 1
    lines.append("f %d//%d %d//%d %d//%d %d//%d %d//%d" % (
        n0 + 1, up, n0 + 4, up, n0 + 5, up, n0 + 3, up, n0 + 2, up))
    return "\n".join(lines) + "\n"


def test_obj_mesh_renders_like_the_oracle(device):
    verts, idx = V.parse_obj(_obj_text())
    desc = scenes.config1(64, 48)
    desc["meshes"][0] = (verts.ravel(), idx, True)
    for est in (V.EST_PARITY, V.EST_UNBIASED):
        clamp = 10.0 if est == V.EST_PARITY else 0.0
        img, m = gpu_render(device, desc, 2, 3, est, V.RNG_PCG, clamp=clamp)
        px, met = oracle_render(desc, 2, 3, est, SPO_RNG_PCG, clamp=clamp)
        assert np.array_equal(img.view(np.uint32), px.view(np.uint32))
        assert [m.values[k] for k in (1, 2, 3, 4)] == [int(met[k]) for k in (1, 2, 3, 4)]
