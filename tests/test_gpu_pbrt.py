"""Imported pbrt scenes on the traced path: handles built straight from skw_pbrt_descs render and
trace like the oracle given the same imported arrays.

Bar: direct-lighting radiance within 1e-5 relative and ray counts identical (as for skw_dl_render);
closest hits of an instanced import bit-exact (prim, tri, t), hit points and normals within 1e-9."""
import numpy as np
import pytest

import oracle
import skwarka_amd as sk
from skwarka_amd import workloads as wl
from tests.conftest import empty_dev, from_dev
from tests.pbrt_files import mesh_shape, room_text
from tests.test_gpu_dl import assert_close_rel
from tests.test_gpu_traversal import assert_hits_equal, gpu_trace

pytestmark = pytest.mark.gpu


def test_imported_room_renders_like_the_oracle(torch_cuda, tmp_path):
    print("case: pbrt import, small room 96x64, 2 spp, direct lighting", flush=True)
    s = wl.scene_room_small()
    path = tmp_path / "room.pbrt"
    path.write_text(room_text(s, 96, 64, 45.0))
    imp = sk.PbrtScene(str(path))
    a = imp.arrays()
    cam, w, h = imp.camera()
    cam15 = np.concatenate([list(cam.origin), list(cam.rot), [cam.x_tan, cam.y_tan]])
    acc = oracle.Accel(a["meshes"])
    odl = oracle.DLScene(acc, a["mat"], a["kd"], a["ks"], a["rough"], a["light_prim"], a["light_radiance"])
    ref, rcnt = odl.render(cam15, w, h, 2, 4, 4, 5)
    scene = imp.scene()
    gdl = imp.dl_scene(scene)
    d_rgb = empty_dev(torch_cuda, w * h * 12)
    cnt = gdl.render(cam, w, h, 2, d_rgb.data_ptr(), 4, 4, 5)
    torch_cuda.cuda.synchronize()
    got = from_dev(d_rgb, np.float32, w * h * 3).reshape(w * h, 3)
    assert np.array_equal(cnt, rcnt) and ref.max() > 0
    assert_close_rel(got, ref)


def test_imported_instances_trace_like_the_oracle(torch_cuda, tmp_path):
    print("case: pbrt import, instanced icospheres, closest hits", flush=True)
    rng = np.random.default_rng(4)
    ball = wl.icosphere(3, 1.0, (0.0, 0.0, 0.0), noise=0.02, rng=rng)
    lines = ["WorldBegin\n", 'ObjectBegin "ball"\n', mesh_shape(ball), "ObjectEnd\n", mesh_shape(wl.box(30.0))]
    for i in range(40):
        c = rng.uniform(-20, 20, 3)
        ax = rng.normal(size=3)
        lines.append("AttributeBegin\nTranslate %r %r %r\nRotate %r %r %r %r\nScale %r %r %r\n"
                     'ObjectInstance "ball"\nAttributeEnd\n' % (
                         *map(float, c), float(rng.uniform(0, 360)), *map(float, ax),
                         *map(float, rng.uniform(0.5, 3.0, 3))))
    lines.append("WorldEnd\n")
    path = tmp_path / "inst.pbrt"
    path.write_text("".join(lines))
    imp = sk.PbrtScene(str(path))
    a = imp.arrays()
    assert len(a["prims"]) == 41
    acc = oracle.Accel(a["meshes"], a["prims"], a["prim_xform"])
    scene = imp.scene()
    assert scene.info()["instances"] == 40
    o = rng.uniform(-25, 25, (30000, 3))
    d = rng.normal(size=(30000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = wl.make_rays(o, d)
    got, g = gpu_trace(torch_cuda, scene, rays, geom=True)
    ref, rg = acc.closest_hit(rays, geom=True)
    assert_hits_equal(got, ref)
    hit = (ref["prim"] > 0) & (ref["prim"] != 0xFFFFFFFF)
    assert hit.sum() > 1000
    for f in ("p", "ng", "ns"):
        assert np.allclose(g[f][hit], rg[f][hit], rtol=0, atol=1e-9), f
