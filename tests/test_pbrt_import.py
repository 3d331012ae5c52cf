"""PBRT import (skw_pbrt_load / skw_pbrt_descs) against closed forms of the reference importer's
semantics (PbrtSceneImporter_hooks.cpp, Factories/*.h, Math/Transform.cpp):

- transforms post-multiply the current transform (Translate then Scale maps v to s*v + t, not
  s*(v + t)); Transform / ConcatTransform take the 16 numbers column by column; Rotate is about the
  normalized axis by degrees; shapes are placed in world space, normals by the inverse transpose;
- LookAt and the perspective camera: the camera frame has columns right, -up', dir (the frame
  skwarka_amd.camera builds), the field of view belongs to the shorter image axis, a camera
  transform that inverts orientation is mirrored in x first;
- material and light defaults (matte Kd 0.5, plastic Kd 0.25 Ks 0.25 roughness 0.1, L 1), named
  materials, L times scale, ReverseOrientation, object instances under the transform at
  ObjectInstance, Include, attribute and transform blocks;
- a scene written from the test room reads back as the same arrays.
The reference has no tests or fixtures for its importer; these properties pin the restatement."""
import math
import os

import numpy as np
import pytest

import skwarka_amd as sk
from skwarka_amd import workloads as wl
from tests.pbrt_files import room_text

TRI = '"integer indices" [0 1 2] "point P" [0 0 0 1 0 0 0 1 0]'


def load(tmp_path, text, name="s.pbrt"):
    p = tmp_path / name
    p.write_text(text)
    return sk.PbrtScene(str(p))


def world(body, camera=""):
    return camera + "WorldBegin\n" + body + "\nWorldEnd\n"


def test_transform_order_and_normals(tmp_path):
    sc = load(tmp_path, world('Translate 1 2 3\nScale 2 4 8\nShape "trianglemesh" %s '
                              '"normal N" [1 1 1 0 0 1 0 0 1]' % TRI))
    a = sc.arrays()
    m = a["meshes"][0]
    assert np.array_equal(m["verts"], np.array([[1, 2, 3], [3, 2, 3], [1, 6, 3]], np.float32))
    n0 = np.array([0.5, 0.25, 0.125]) / np.linalg.norm([0.5, 0.25, 0.125])
    assert np.allclose(m["normals"][0], n0, rtol=0, atol=1e-7)
    assert np.allclose(m["normals"][1:], [[0, 0, 1], [0, 0, 1]], atol=0)
    assert np.array_equal(a["prim_xform"][0], np.concatenate([np.eye(4).ravel(), np.eye(4).ravel()]))


def test_transform_commands(tmp_path):
    body = "\n".join([
        "TransformBegin",
        "Transform [1 0 0 0  0 1 0 0  0 0 1 0  5 6 7 1]",  # column-major: a translation by (5, 6, 7)
        'Shape "trianglemesh" %s' % TRI,
        "TransformEnd",
        "TransformBegin",
        "Rotate 90 0 0 2",  # about +z (axis normalized)
        'Shape "trianglemesh" %s' % TRI,
        'CoordinateSystem "rot"',
        "TransformEnd",
        "Translate 10 0 0",
        "ConcatTransform [2 0 0 0  0 2 0 0  0 0 2 0  0 0 0 1]",
        'Shape "trianglemesh" %s' % TRI,
        'CoordSysTransform "rot"',
        'Shape "trianglemesh" %s' % TRI,
        "Identity",
        'Shape "trianglemesh" %s' % TRI,
    ])
    a = load(tmp_path, world(body)).arrays()
    v = [m["verts"] for m in a["meshes"]]
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    assert np.array_equal(v[0], base + np.float32([5, 6, 7]))
    assert np.allclose(v[1], [[0, 0, 0], [0, 1, 0], [-1, 0, 0]], atol=1e-15)
    assert np.array_equal(v[2], base * 2 + np.float32([10, 0, 0]))
    assert np.allclose(v[3], v[1], atol=0)
    assert np.array_equal(v[4], base)


def test_perspective_camera(tmp_path):
    for (w, h) in ((320, 240), (240, 320), (100, 100)):
        cam_txt = ('LookAt 1 2 3  1 2 -4  0 1 0\nCamera "perspective" "float fov" [50]\n'
                   'Film "image" "integer xresolution" [%d] "integer yresolution" [%d]\n' % (w, h))
        sc = load(tmp_path, world('Shape "trianglemesh" %s' % TRI, cam_txt))
        cam, cw, ch = sc.camera()
        assert (cw, ch) == (w, h)
        # pbrt's fov is the shorter axis' angle; the x tangent grows by the aspect when wider than tall
        x_tan = math.tan(math.radians(25.0)) * max(1.0, w / h)
        want, _ = sk.camera((1, 2, 3), (0, 0, -1), (0, 1, 0), 2 * math.degrees(math.atan(x_tan)), w, h)
        assert np.allclose(list(cam.origin), list(want.origin), atol=0)
        assert np.allclose(list(cam.rot), list(want.rot), atol=1e-15)
        assert abs(cam.x_tan - x_tan) < 1e-7 * x_tan
        assert abs(cam.y_tan - x_tan * h / w) < 1e-7 * x_tan


def test_camera_defaults_mirroring_and_halffov(tmp_path):
    plain = load(tmp_path, world("", 'LookAt 0 0 5  0 0 0  0 1 0\nCamera "perspective"\n')).camera()
    assert (plain[1], plain[2]) == (640, 480)
    # the frame aspect is a float (640 / 480.f) as in CameraFactory
    assert abs(plain[0].x_tan - math.tan(math.radians(45.0)) * 640 / 480) < 1e-7
    # Scale -1 1 1 makes the camera transform invert orientation; the importer mirrors it back
    mirrored = load(tmp_path, world("", 'Scale -1 1 1\nLookAt 0 0 5  0 0 0  0 1 0\nCamera "perspective"\n')).camera()
    assert list(mirrored[0].rot) == list(plain[0].rot) and list(mirrored[0].origin) == list(plain[0].origin)
    half = load(tmp_path, world("", 'Camera "perspective" "float halffov" [20] "float fov" [90]\n'
                                'Film "image" "integer xresolution" [10] "integer yresolution" [10]\n')).camera()
    assert abs(half[0].x_tan - math.tan(math.radians(20.0))) < 1e-7
    with pytest.raises(sk.SkwError):
        load(tmp_path, world('Shape "trianglemesh" %s' % TRI)).camera()


def test_materials_lights_and_blocks(tmp_path):
    body = "\n".join([
        'Shape "trianglemesh" %s' % TRI,  # default material: matte Kd 0.5
        "AttributeBegin",
        'Material "plastic"',
        'Shape "trianglemesh" %s' % TRI,
        'Material "plastic" "rgb Kd" [0.1 0.2 0.3] "color Ks" [0.4 0.5 0.6] "float roughness" [0.05]',
        'Shape "trianglemesh" %s' % TRI,
        'MakeNamedMaterial "red" "string type" ["matte"] "rgb Kd" [0.9 0.1 0.1]',
        "AttributeEnd",
        'NamedMaterial "red"',
        'Shape "trianglemesh" %s' % TRI,
        "AttributeBegin",
        'AreaLightSource "diffuse"',
        "ReverseOrientation",
        'Shape "trianglemesh" %s' % TRI,
        "AttributeEnd",
        "AttributeBegin",
        'AreaLightSource "area" "rgb L" [4 4 4] "rgb scale" [1 0.5 2]',
        'Shape "trianglemesh" %s' % TRI,
        "AttributeEnd",
        'Shape "trianglemesh" %s' % TRI,  # after the blocks: no light, not reversed
    ])
    a = load(tmp_path, world(body)).arrays()
    f = np.float32
    assert list(a["mat"]) == [0, 1, 1, 0, 0, 0, 0]
    assert np.array_equal(a["kd"][:4], [[0.5] * 3, [0.25] * 3, [f(0.1), f(0.2), f(0.3)], [f(0.9), f(0.1), f(0.1)]])
    assert np.array_equal(a["ks"][1:3], [[0.25] * 3, [f(0.4), f(0.5), f(0.6)]])
    assert a["rough"][1] == f(0.1) and a["rough"][2] == f(0.05)
    assert list(a["light_prim"]) == [4, 5]
    assert np.array_equal(a["light_radiance"], [[1, 1, 1], [4, 2, 8]])
    assert [m["invert_normals"] for m in a["meshes"]] == [False, False, False, False, True, False, False]


def test_object_instances(tmp_path):
    body = "\n".join([
        'ObjectBegin "tri"',
        "Translate 0 0 1",
        'Shape "trianglemesh" %s' % TRI,
        "ObjectEnd",
        'Shape "trianglemesh" %s' % TRI,  # the Translate inside the object does not leak
        "Rotate 90 0 0 1",
        'ObjectInstance "tri"',
        "Scale 2 2 2",
        'ObjectInstance "tri"',
        'ObjectInstance "nothing"',
    ])
    sc = load(tmp_path, world(body))
    a = sc.arrays()
    assert len(a["meshes"]) == 2
    assert np.array_equal(a["meshes"][0]["verts"][:, 2], [1, 1, 1])
    assert np.array_equal(a["meshes"][1]["verts"][:, 2], [0, 0, 0])
    assert [p[0] for p in a["prims"]] == [1, 0, 0]
    r = np.array([[0, -1, 0, 0], [1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float64)
    assert np.allclose(a["prims"][1][1], r, atol=1e-15)
    assert np.allclose(a["prims"][2][1], r @ np.diag([2, 2, 2, 1]), atol=1e-15)
    inv = a["prim_xform"][2, 16:].reshape(4, 4)
    assert np.allclose(inv @ a["prims"][2][1], np.eye(4), atol=1e-15)
    assert any("nothing" in w for w in sc.warnings)
    s = sc.scene(host_only=True)
    info = s.info()  # one static triangle, the object's triangle once, two instances of it
    assert info["triangles"] == 2 and info["instances"] == 2


def test_include_warnings_and_errors(tmp_path):
    (tmp_path / "inc").mkdir()
    (tmp_path / "inc" / "part.pbrt").write_text('Translate 0 0 3\nShape "trianglemesh" %s\n' % TRI)
    sc = load(tmp_path, world('Include "inc/part.pbrt"\nShape "sphere" "float radius" 1\n'
                              'Shape "trianglemesh" "integer indices" [0 1 5] "point P" [0 0 0 1 0 0 0 1 0]\n'
                              'Material "glass"\nLightSource "point"\n'))
    a = sc.arrays()
    assert len(a["meshes"]) == 1 and np.array_equal(a["meshes"][0]["verts"][:, 2], [3, 3, 3])
    w = " | ".join(sc.warnings)
    for word in ("sphere", "out-of-bounds", "glass", "LightSource"):
        assert word in w
    for bad in ('Shape "trianglemesh "integer indices" [0 1 2]', 'Translate 1 2', "Translate 1 x 3",
                'Shape "trianglemesh" "integer indices" [0 1 2', 'Include "missing.pbrt"', "[ 1 2 ]"):
        with pytest.raises(sk.SkwError):
            load(tmp_path, world(bad))
    with pytest.raises(sk.SkwError):
        sk.PbrtScene(str(tmp_path / "no_such_file.pbrt"))


def test_room_file_reads_back_as_the_room(tmp_path):
    s = wl.scene_room_small()
    sc = load(tmp_path, room_text(s, 96, 64, 45.0), "room.pbrt")
    assert sc.warnings == []
    a = sc.arrays()
    assert len(a["meshes"]) == len(s["meshes"])
    for m, ref in zip(a["meshes"], s["meshes"]):
        assert np.array_equal(m["verts"], ref["verts"]) and np.array_equal(m["tris"], ref["tris"])
    f = lambda x: np.asarray(x, np.float32).astype(np.float64)
    assert np.array_equal(a["mat"], s["mat"])
    assert np.array_equal(a["kd"], f(s["kd"])) and np.array_equal(a["ks"][s["mat"] == 1], f(s["ks"])[s["mat"] == 1])
    assert np.array_equal(a["rough"][s["mat"] == 1], f(s["rough"])[s["mat"] == 1])
    assert np.array_equal(a["light_prim"], s["light_prim"])
    assert np.array_equal(a["light_radiance"], f(s["light_radiance"]))
    cam, w, h = sc.camera()
    assert (w, h) == (96, 64)
    want, _ = sk.camera(s["camera"][0], s["camera"][1], s["camera"][2], 2 * math.degrees(math.atan(
        math.tan(math.radians(22.5)) * 1.5)), 96, 64)
    assert np.allclose(list(cam.rot), list(want.rot), atol=1e-7)
    assert np.allclose(list(cam.origin), list(want.origin), atol=1e-6)
