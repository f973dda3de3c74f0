"""OBJ ingestion (spg_obj_parse, LoadMesh src/mesh.cpp:5-62): fans, joined
(position, normal) corners, the first object only, texture coordinates
dropped; the reference's own assets when the reference tree is present.
Host-only calls into the library (no GPU)."""
import os

import numpy as np
import pytest

import vk_cinematic_amd as V

ASSETS = "/root/reference/assets"


def test_polygons_become_fans_and_corners_are_joined():
    text = """# comment
o first
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 2 0.5 0
vn 0 0 1
vn 0 0 -1
vt 0.5 0.5

f 1/1/1 2/1/1 3/1/1 4/1/1
f 2//1 5//1 3//1
f -5//-1 -4//-1 -3//-1
o second
v 9 9 9
f 1//1 2//1 6//1
"""
    v, i = V.parse_obj(text)
    # vertices in order of first use: 1/n1 2/n1 3/n1 4/n1 5/n1 1/n2 2/n2 3/n2
    want_pos = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [2, .5, 0],
                [0, 0, 0], [1, 0, 0], [1, 1, 0]]
    assert np.array_equal(v[:, :3], np.array(want_pos, np.float32))
    assert np.array_equal(v[:5, 3:6], np.tile([0, 0, 1], (5, 1)).astype(np.float32))
    assert np.array_equal(v[5:, 3:6], np.tile([0, 0, -1], (3, 1)).astype(np.float32))
    assert np.all(v[:, 6:8] == 0)
    assert i.tolist() == [0, 1, 2, 0, 2, 3, 1, 4, 2, 5, 6, 7]


@pytest.mark.parametrize("text", [
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n",             # no normals
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 4//1\n",  # index past the end
    "v 0 0 0\nv 1 0 0\nvn 0 0 1\nf 1//1 2//1\n",          # two corners
    "v 0 0 0\n",                                          # no faces
])
def test_rejects_malformed_files(text):
    with pytest.raises(V.SpgError):
        V.parse_obj(text)


def _independent_parse(path):
    """Corner (position, normal) triples of every face of the first object,
    read with plain Python."""
    pos, nrm, faces, objects = [], [], [], 0
    for line in open(path):
        t = line.split()
        if not t:
            continue
        if t[0] == "o":
            objects += 1
            if objects > 1:
                break
        elif t[0] == "v":
            pos.append([float(x) for x in t[1:4]])
        elif t[0] == "vn":
            nrm.append([float(x) for x in t[1:4]])
        elif t[0] == "f":
            c = []
            for tok in t[1:]:
                parts = tok.split("/")
                c.append((int(parts[0]) - 1, int(parts[2]) - 1))
            for k in range(1, len(c) - 1):
                faces.append((c[0], c[k], c[k + 1]))
    return np.array(pos, np.float32), np.array(nrm, np.float32), faces


@pytest.mark.parametrize("name,tris", [("bunny.obj", 4968), ("monkey.obj", 15744)])
def test_reference_assets(name, tris):
    path = os.path.join(ASSETS, name)
    if not os.path.exists(path):
        pytest.skip("reference assets not present")
    v, i = V.load_obj(path)
    pos, nrm, faces = _independent_parse(path)
    assert len(i) == 3 * tris == 3 * len(faces)
    distinct = {c for f in faces for c in f}
    assert len(v) == len(distinct)
    corners = np.array([c for f in faces for c in f])
    assert np.array_equal(v[i, 0:3], pos[corners[:, 0]])
    assert np.array_equal(v[i, 3:6], nrm[corners[:, 1]])


def test_text_without_terminator_is_read_only_to_its_length():
    """The text is (pointer, length): a buffer whose last line ends in a
    digit with more digits right after it in memory must be read only up to
    the length (the bytes after it would make index 37 and a vertex 0.55)."""
    import ctypes as C

    body = b"vn 0 0 1\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1//1 2//1 3//1"
    buf = C.create_string_buffer(body + b"7 v 0.5 5", len(body) + 10)
    nv, ni = C.c_uint32(), C.c_uint32()
    V.check(V.lib().spg_obj_parse(buf, len(body), None, 0, None, 0, C.byref(nv),
                                  C.byref(ni)))
    assert (nv.value, ni.value) == (3, 3)
    # a vertex line cut at the length: "v 0 0 0.5" is read as 0.5, not 0.55
    body2 = b"vn 0 0 1\nv 0 0 0\nv 1 0 0\nv 0 1 0.5"
    tail = b"5\nf 1//1 2//1 3//1\n"
    buf2 = C.create_string_buffer(body2 + tail, len(body2) + len(tail))
    # the cut text has no faces: rejected, but without reading the tail's face
    assert V.lib().spg_obj_parse(buf2, len(body2), None, 0, None, 0, C.byref(nv),
                                 C.byref(ni)) == -1  # SPG_E_INVALID
    full = body2 + b"\nf 1//1 2//1 3//1"
    buf3 = C.create_string_buffer(full + b"5", len(full) + 1)
    v = np.zeros(24, np.float32)
    i = np.zeros(3, np.uint32)
    V.check(V.lib().spg_obj_parse(buf3, len(full), v.ctypes.data, 3, i.ctypes.data, 3,
                                  C.byref(nv), C.byref(ni)))
    assert v.reshape(3, 8)[2, 2] == np.float32(0.5)
