"""The four-wide quantized hierarchy the GPU walks: its structure, the
exact decode of its 8-bit child boxes, and the CPU restatement of its walk
against brute force (CPU only; the GPU's walk is compared with this
restatement in the GPU closest-hit tests)."""
import ctypes as C

import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes
from _libs import oracle, wide_boxes


def _check(nodes, order, tris):
    pend = C.c_uint32()
    rc = oracle().spo_wide_check(np.ascontiguousarray(nodes.ravel()), len(nodes),
                                 order, np.ascontiguousarray(tris.ravel()), len(tris),
                                 C.byref(pend))
    return rc, pend.value


@pytest.mark.parametrize("leaf", [1, 4, 8])
def test_wide_invariants_and_stack_bound(leaf):
    """Every primitive in one leaf, references in range, present children
    first, every triangle inside the decoded box of every ancestor child;
    the builder's stack bound equals the checker's."""
    tris = np.ascontiguousarray(V.gen_triangle_soup(30000, seed=3, mu=-3.0, sigma=0.7))
    nodes, order, pend = V.bvh_build_wide_host(tris, max_leaf=leaf)
    rc, p2 = _check(nodes, order, tris)
    assert rc == 0 and pend == p2
    _, _, present = wide_boxes(nodes)
    assert present.sum(1).mean() > 3.0           # collapsing fills the nodes
    n2 = len(V.bvh_build_host(tris, max_leaf=leaf)[0])
    assert len(nodes) < 0.6 * n2


def test_checker_catches_a_shrunk_box():
    """A planted mistake: one child's quantized lo raised by one step on an
    axis where it was tight; the containment check must fail."""
    tris = np.ascontiguousarray(V.gen_triangle_soup(2000, seed=4, mu=-2.5, sigma=0.5))
    nodes, order, _ = V.bvh_build_wide_host(tris, max_leaf=4)
    assert _check(nodes, order, tris)[0] == 0
    bad = nodes.copy()
    lo, hi, present = wide_boxes(bad)
    # the root's first child: raise its lo.x by the smallest step that
    # leaves a vertex outside (q = 0 is tight at the node origin)
    q = bad[0, 32 + 0 * 4 + 0]
    bad[0, 32 + 0 * 4 + 0] = min(255, int(q) + 8)
    assert _check(bad, order, tris)[0] == 5


def test_decode_is_exact():
    """q * s is exact, so p + q * s rounds once: the float32 decode equals
    the float64 sum rounded to float32 for every child coordinate."""
    tris = np.ascontiguousarray(V.gen_triangle_soup(20000, seed=5, mu=-3.0, sigma=0.7))
    nodes, _, _ = V.bvh_build_wide_host(tris)
    lo, hi, present = wide_boxes(nodes)
    n = len(nodes)
    p = nodes[:, 0:12].copy().view(np.float32).reshape(n, 3).astype(np.float64)
    s = (nodes[:, 12:15].astype(np.uint32) << 23).view(np.float32).astype(np.float64)
    q = nodes[:, 32:56].reshape(n, 6, 4).transpose(0, 2, 1).astype(np.float64)
    lo64 = (p[:, None, :] + q[:, :, 0:3] * s[:, None, :]).astype(np.float32)
    hi64 = (p[:, None, :] + q[:, :, 3:6] * s[:, None, :]).astype(np.float32)
    assert np.array_equal(lo[present], lo64[present])
    assert np.array_equal(hi[present], hi64[present])
    assert (lo[present] <= hi[present]).all()


def _brute_and_wide(tris, o, d):
    O = oracle()
    rays = np.ascontiguousarray(np.concatenate([o, d], 1).ravel().astype(np.float32))
    n = len(o)
    nodes, order, _ = V.bvh_build_wide_host(tris, max_leaf=4)
    p1 = np.zeros(n, np.uint32)
    t1 = np.zeros(n, np.float32)
    box, tri = C.c_uint64(), C.c_uint64()
    O.spo_wide_closest(np.ascontiguousarray(nodes.ravel()), order, tris.ravel(), rays,
                       n, 8, p1, t1, C.byref(box), C.byref(tri))
    p2 = np.zeros(n, np.uint32)
    t2 = np.zeros(n, np.float32)
    O.spo_closest_hit_brute(tris.ravel(), len(tris), rays, n, 8, p2, t2)
    return (p1, t1), (p2, t2), box.value, tri.value, nodes


def test_wide_twin_equals_brute_force():
    """The GPU's wide walk restated on the CPU returns the brute-force
    closest hit (smallest t > 0, ties to the lower triangle index)."""
    _, o, d = scenes.soup(16, 4096, seed=9)
    tris = np.ascontiguousarray(V.gen_triangle_soup(20000, seed=9, mu=-2.5, sigma=0.5))
    (p1, t1), (p2, t2), box, tri, _ = _brute_and_wide(tris, o, d)
    assert np.array_equal(p1, p2)
    assert np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
    assert (p2 != 0xFFFFFFFF).sum() > 50
    assert box > len(o) and tri > 0


def test_wide_twin_axis_parallel_rays_on_planes():
    """Rays parallel to an axis whose origin lies exactly on a decoded box
    plane (slab distance 0 * inf = NaN) and rays along a coordinate axis:
    the wide walk still returns the brute-force closest hit."""
    rng = np.random.default_rng(6)
    tris = np.ascontiguousarray(V.gen_triangle_soup(8000, seed=22, mu=-2.5, sigma=0.5))
    nodes, _, _ = V.bvh_build_wide_host(tris)
    lo, hi, present = wide_boxes(nodes)
    planes = np.concatenate([lo, hi], axis=1)[np.concatenate([present, present], 1)]
    n = 4096
    k = rng.integers(0, 3, n)
    o = rng.uniform(-0.2, 1.2, (n, 3)).astype(np.float32)
    o[np.arange(n), k] = planes[rng.integers(0, len(planes), n), k]
    d = rng.normal(size=(n, 3))
    d[np.arange(n), k] = 0.0
    d[: n // 8] = 0.0
    d[np.arange(n // 8), rng.integers(0, 3, n // 8)] = rng.choice([-1.0, 1.0], n // 8)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    (p1, t1), (p2, t2), _, _, _ = _brute_and_wide(tris, o, d.astype(np.float32))
    assert np.array_equal(p1, p2)
    assert np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
    assert (p2 != 0xFFFFFFFF).sum() > 20


def test_wide_and_binary_walks_agree_on_a_closed_mesh():
    """Rays from outside a displaced icosphere and from its surface: the
    wide and the binary walks return the same triangle and t bits, and the
    wide walk takes fewer node visits (fewer box-test rounds)."""
    v, idx = scenes.displaced_icosphere(4, seed=3)[:2]
    t = np.ascontiguousarray(v[:, :3][idx.reshape(-1, 3)].reshape(-1, 9), np.float32)
    rng = np.random.default_rng(1)
    n = 8192
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * rng.uniform(1.2, 4, (n, 1))
    d = rng.normal(size=(n, 3)) * 0.6 - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o, d = o.astype(np.float32), d.astype(np.float32)
    O = oracle()
    rays = np.ascontiguousarray(np.concatenate([o, d], 1).ravel())
    nodes2, order2, _ = V.bvh_build_host(t, max_leaf=4)
    pa = np.zeros(n, np.uint32)
    ta = np.zeros(n, np.float32)
    b2, r2 = C.c_uint64(), C.c_uint64()
    O.spo_bvh2_closest(np.ascontiguousarray(nodes2.ravel()), order2, t.ravel(), rays, n,
                       8, pa, ta, C.byref(b2), C.byref(r2))
    (pw, tw), _, bw, rw, _ = _brute_and_wide(t, o, d)
    assert np.array_equal(pa, pw)
    assert np.array_equal(ta.view(np.uint32), tw.view(np.uint32))
    assert (pw != 0xFFFFFFFF).mean() > 0.5
    assert r2.value == rw or abs(r2.value - rw) < 0.1 * r2.value
