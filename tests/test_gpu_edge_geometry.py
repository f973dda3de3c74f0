"""Edge cases of the geometry path: axis-parallel rays whose origin lies exactly on a BVH box plane, and non-finite input geometry."""
import ctypes as C

import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes
from _libs import oracle, wide_boxes

pytestmark = pytest.mark.gpu


def test_axis_parallel_rays_on_box_planes_match_traversal_twin(device):
    """A ray parallel to a slab (direction component exactly 0, inverse
    infinite) whose origin lies exactly on the slab's plane has the slab
    distance 0 * inf = NaN, where the reference's compare-and-select Min/Max
    and IEEE min/max differ.  Triangle, t bits and the box and triangle test
    counts equal the oracle's restatement of the traversal."""
    rng = np.random.default_rng(5)
    tris = V.gen_triangle_soup(20_000, seed=21, mu=-2.5, sigma=0.5)
    s = V.Scene(device)
    v, i = scenes.soup_mesh(np.asarray(tris, np.float32).reshape(-1, 9))
    m = s.create_mesh(v, i, smooth=False)
    s.build()
    nodes, order, _ = s.wide_export(m)
    # plane coordinates of the decoded child boxes the traversal tests
    lo, hi, present = wide_boxes(nodes)
    planes = np.concatenate([lo, hi], axis=1)[np.concatenate([present, present], 1)]
    n = 8192
    k = rng.integers(0, 3, n)              # the axis the ray is parallel to
    o = rng.uniform(-0.2, 1.2, (n, 3)).astype(np.float32)
    o[np.arange(n), k] = planes[rng.integers(0, len(planes), n), k]
    d = rng.normal(size=(n, 3))
    d[np.arange(n), k] = 0.0
    d[: n // 8] = 0.0                      # and some along a coordinate axis
    d[np.arange(n // 8), (k[: n // 8] + 1) % 3] = rng.choice([-1.0, 1.0], n // 8)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    assert (d[np.arange(n), k] == 0).all()
    bufs = [device.upload(np.ascontiguousarray(x)) for x in
            (o[:, 0], o[:, 1], o[:, 2], d[:, 0], d[:, 1], d[:, 2])]
    outp, outt = device.alloc(4 * n), device.alloc(4 * n)
    st = V.TraversalStats()
    V.check(V.lib().spg_intersect_closest_device(
        s.handle, m, *[b.ptr for b in bufs], n, outp.ptr, outt.ptr, None, None,
        C.byref(st)))
    prim = outp.download(np.uint32, n)
    t = outt.download(np.float32, n)
    O = oracle()
    rays = np.ascontiguousarray(np.concatenate([o, d], 1).ravel())
    tp = np.zeros(n, np.uint32)
    tt = np.zeros(n, np.float32)
    box, tri = C.c_uint64(), C.c_uint64()
    O.spo_wide_closest(np.ascontiguousarray(nodes.ravel()), order,
                       np.ascontiguousarray(tris.ravel()), rays, n, 16, tp, tt,
                       C.byref(box), C.byref(tri))
    assert np.array_equal(prim, tp)
    assert np.array_equal(t.view(np.uint32), tt.view(np.uint32))
    assert st.boxTests == box.value and st.triTests == tri.value
    assert (prim != V.SPG_MISS).sum() > 50


def test_non_finite_geometry_is_rejected(device):
    """Vertex positions and object transforms must be finite (the traversal's
    min/max slab test relies on finite boxes)."""
    s = V.Scene(device)
    v = np.zeros((3, 8), np.float32)
    v[1, 0] = 1.0
    v[2, 1] = 1.0
    idx = np.array([0, 1, 2], np.uint32)
    m = s.create_mesh(v, idx, smooth=False)
    bad = v.copy()
    bad[2, 2] = np.nan
    with pytest.raises(V.SpgError):
        s.create_mesh(bad, idx, smooth=False)
    bad[2, 2] = np.inf
    with pytest.raises(V.SpgError):
        s.create_mesh(bad, idx, smooth=False)
    s.register_material(7, albedo=(0.5, 0.5, 0.5))
    with pytest.raises(V.SpgError):
        s.add_object(m, 7, (0, np.inf, 0))
    with pytest.raises(V.SpgError):
        s.add_object(m, 7, (0, 0, 0), scale=(1, np.nan, 1))
    s.add_object(m, 7, (0, 0, 0))
    s.build()
    s.close()


def test_fast_reciprocal_equals_ieee_division_everywhere(device):
    """The tracer's reciprocal (hardware estimate + one FMA Newton step,
    spg_math.h rcp_cr) against the IEEE division for every float bit
    pattern of its fast range (exponent field 2..252): no difference, so
    normalize, the inverse direction and Moller-Trumbore's determinant keep
    their bit-exact results."""
    import ctypes as C

    counts = (C.c_uint64 * 2)()
    V.check(V.lib().spg_check_arithmetic(device.handle, counts))
    assert counts[0] == 2 * 251 * (1 << 23)
    assert counts[1] == 0
