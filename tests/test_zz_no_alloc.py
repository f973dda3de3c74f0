"""spg_intersect_closest_device_noalloc: the closest-hit entry point that
allocates nothing (its work counter and stack spill are a static device
array of the library).  Against spg_intersect_closest_device on the same
device buffers, bit for bit, on two calls in a row, the second with fewer
rays (the scratch keeps the first call's counter, which the entry point must
clear); a mesh whose walk exceeds the scratch's stack is refused."""
import ctypes as C

import numpy as np
import pytest

import vk_cinematic_amd as V
from vk_cinematic_amd import scenes

pytestmark = pytest.mark.gpu


def _closest(fn, s, m, bufs, n, outp, outt, outu, outv, *extra):
    V.check(fn(s.handle, m, *[b.ptr for b in bufs], n, outp.ptr, outt.ptr, outu.ptr,
               outv.ptr, *extra))
    return (outp.download(np.uint32, n), outt.download(np.float32, n),
            outu.download(np.float32, n), outv.download(np.float32, n))


def test_noalloc_equals_the_allocating_entry_point(device):
    tris, o, d = scenes.soup(200_000, 1 << 20, seed=3)
    s = V.Scene(device)
    v, i = scenes.soup_mesh(tris)
    m = s.create_mesh(v, i, smooth=False)
    s.build()
    n = len(o)
    bufs = [device.upload(np.ascontiguousarray(x, np.float32)) for x in
            (o[:, 0], o[:, 1], o[:, 2], d[:, 0], d[:, 1], d[:, 2])]
    outs = [device.alloc(4 * n) for _ in range(4)]
    L = V.lib()
    for count in (n, 300_001):
        want = _closest(L.spg_intersect_closest_device, s, m, bufs, count, *outs, None)
        for b in outs:  # poison the outputs between the calls
            b.upload(np.full(n, 0x7FC00001, np.uint32))
        got = _closest(L.spg_intersect_closest_device_noalloc, s, m, bufs, count, *outs)
        print(f"n={count}: hits {int((want[0] != 0xFFFFFFFF).sum())}")
        for a, b in zip(want, got):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert (want[0] != 0xFFFFFFFF).sum() > count // 50
    for b in bufs + outs:
        b.free()
    s.close()


def test_noalloc_limits(device):
    """n = 0 does nothing; NULL arrays are refused; the header's stack limit
    is the one enforced (a mesh within it is accepted)."""
    v, i = V.gen_icosphere(3)
    s = V.Scene(device)
    m = s.create_mesh(v, i)
    s.build()
    L = V.lib()
    assert L.spg_intersect_closest_device_noalloc(s.handle, m, *([None] * 6), 0, None,
                                                  None, None, None) == 0
    assert L.spg_intersect_closest_device_noalloc(s.handle, m, *([None] * 6), 5, None,
                                                  None, None, None) == -1
    _, _, pending = V.bvh_build_wide_host(
        np.ascontiguousarray(v[i.reshape(-1, 3), :3].reshape(-1, 9)))
    assert pending <= 56
    s.close()
