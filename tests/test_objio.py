"""Wavefront OBJ in and out (host code of the C-ABI; no GPU)."""
import numpy as np
import pytest


def test_sphere_roundtrip_is_exact(rtx, tmp_path):
    """The sphere writer's file reads back into the identical mesh (the
    sphere's vertices are in first-use order, so no renumbering shows)."""
    for n in (0, 2, 4):
        v, i = rtx.Sphere(1.0, n).mesh()
        p = str(tmp_path / ("sphere_%d.obj" % n))
        rtx.write_obj(p, [(v, i)], "sphere_%d" % n)
        (v2, i2), = rtx.read_obj(p)
        assert np.array_equal(v2.view(np.uint32), v.view(np.uint32)) and np.array_equal(i2, i)
        txt = open(p).read()
        assert txt.startswith("o sphere_%d\n" % n) and ("# %d vertices" % len(v)) in txt
        assert ("# %d triangles" % len(i)) in txt


def test_renumbering_formats_and_shapes(rtx, tmp_path):
    p = tmp_path / "mixed.obj"
    p.write_text("""# comment
mtllib x.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 5 5 5 1.0
vt 0 0
vn 0 0 1
o quad
usemtl m
f 4/1/1 3/1/1 2//1 1
g tri
s off
f -1 -2 -3
""")
    meshes = rtx.read_obj(str(p))
    assert len(meshes) == 2
    (v0, i0), (v1, i1) = meshes
    # quad 4 3 2 1 -> fan (4,3,2), (4,2,1); vertices in first use order
    assert np.allclose(v0, [[0, 1, 0], [1, 1, 0], [1, 0, 0], [0, 0, 0]])
    assert i0.tolist() == [[0, 1, 2], [0, 2, 3]]
    # negative indices relative to the 5 vertices: -1 -> 5, -2 -> 4, -3 -> 3
    assert np.allclose(v1, [[5, 5, 5], [0, 1, 0], [1, 1, 0]]) and i1.tolist() == [[0, 1, 2]]


def test_objects_feed_scene_meshes(rtx, tmp_path):
    p = tmp_path / "two.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\no a\nf 1 2 3\no b\nf 1 2 4\nf 2 3 4\n")
    obj = rtx.Object(str(p))
    assert len(obj) == 2 and obj[1][1].shape == (2, 3)


@pytest.mark.parametrize("text,what", [("v 0 0 0\nf 1 2 3\n", "out of bounds"), ("v 0 0\n", "three coordinates"),
                                       ("v 0 0 0\nv 1 0 0\nf 1 2\n", "fewer than 3"),
                                       ("v 0 0 0\nf 0 1 1\n", "out of bounds"), ("v 0 0 0\nf a 1 1\n", "malformed")])
def test_bad_files_are_errors(rtx, tmp_path, text, what):
    p = tmp_path / "bad.obj"
    p.write_text(text)
    with pytest.raises(rtx.RtxError, match=what):
        rtx.read_obj(str(p))


def test_missing_file_is_an_error(rtx, tmp_path):
    with pytest.raises(rtx.RtxError, match="cannot open"):
        rtx.read_obj(str(tmp_path / "nope.obj"))
