"""The triangle tree built on the device (skw_scene_create with SKW_SCENE_DEVICE_BUILD) is node for
node the tree of the reference builder (TriangleAccelerator ctor / Node::Node /
_DetermineBestSplit, TriangleAccelerator.cpp:22-147, 505-729): the same boxes (bit patterns), split
axes, child masks and triangle ranges in the same depth-first layout, the same triangle order and
counts. Checked against the oracle's restatement of the builder on the C0 scene, the reference's
tetra-sphere fixture, seeded spheres, the 200-deep staircase, coplanar and duplicated triangles and
signed zeros (whose node boxes keep the sign of the first extreme in array order), at several
small- and wide-node thresholds (which move nodes between the one-wave subtree kernel, the
workgroup-per-node kernel and the chunked multi-workgroup kernels);
and against the product's host builder (itself pinned to the oracle) on the 2.03M-triangle room
and the batched scenes. The traversal of a device-built scene returns the host-built scene's hit
records bit for bit."""
import os

import numpy as np
import pytest

import oracle
import skwarka_amd as sk
from skwarka_amd import workloads as wl
from tests import scenes
from tests.conftest import empty_dev, from_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _device(torch_cuda):
    return torch_cuda


def _same_tree(ref, sc):
    """ref: (bbox, meta, tri_prim, tri_tri) plus info dict; sc: a device-built scene."""
    (b1, m1, tp1, tt1), info1 = ref
    b2, m2, tp2, tt2 = sc.export_tree()
    assert len(b1) == len(b2), (len(b1), len(b2))
    assert np.array_equal(m1, m2)
    assert np.array_equal(b1.view(np.uint32), b2.view(np.uint32))
    assert np.array_equal(tp1, tp2) and np.array_equal(tt1, tt2)
    si = sc.info()
    for key in ("nodes", "triangles", "max_depth", "instances", "leaves", "stack_need"):
        if key in info1:
            assert info1[key] == si[key], key


def _vs_oracle(meshes, prims=None):
    acc = oracle.Accel(meshes, prims)
    dev = sk.Scene(meshes, prims, device_build=True)
    _same_tree((acc.dump(), acc.info()), dev)
    assert np.array_equal(acc.bounds(), dev.bounds())
    return dev


def _vs_host(meshes, prims=None):
    host = sk.Scene(meshes, prims, host_only=True)
    dev = sk.Scene(meshes, prims, device_build=True)
    _same_tree((host.export_tree(), host.info()), dev)
    assert np.array_equal(host.bounds(), dev.bounds())
    return host, dev


@pytest.fixture
def build_env():
    """Sets the builder's size thresholds for one test (restored afterwards)."""
    names = ("SKW_TRIBUILD_SMALL_MAX", "SKW_TRIBUILD_WIDE_MIN", "SKW_TRIBUILD_CHUNK")
    old = {k: os.environ.get(k) for k in names}

    def set_(**kw):
        for k, v in kw.items():
            os.environ["SKW_TRIBUILD_" + k.upper()] = str(v)

    yield set_
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


def coplanar_plane(n=40, seed=1):
    """A grid of triangles in the plane z = 0 plus jittered duplicates: zero extent on one axis
    (NaN bin positions), identical boxes (all-middle partitions down to the three-axis mask)."""
    rng = np.random.default_rng(seed)
    verts, tris = [], []
    for i in range(n):
        for j in range(n):
            b = len(verts)
            verts += [[i, j, 0], [i + 1, j, 0], [i, j + 1, 0]]
            tris.append([b, b + 1, b + 2])
    for _ in range(300):  # 300 copies of one triangle
        b = len(verts)
        verts += [[3.0, 3.0, 0.0], [4.0, 3.0, 0.0], [3.0, 4.0, 0.0]]
        tris.append([b, b + 1, b + 2])
    for _ in range(200):
        b = len(verts)
        c = rng.uniform(0, n, 2)
        verts += [[c[0], c[1], 0], [c[0] + 0.01, c[1], 0], [c[0], c[1] + 0.01, 0]]
        tris.append([b, b + 1, b + 2])
    return [{"verts": np.array(verts, np.float32), "tris": np.array(tris, np.uint32)}]


def signed_zero_soup(n=3000, seed=2):
    """Small triangles on a lattice whose coordinates include -0.0 and +0.0 (and tiny values), so
    that box extremes tie between signed zeros."""
    rng = np.random.default_rng(seed)
    vals = np.array([-0.0, 0.0, -1.0, 1.0, 0.5, -0.5, 2.0, -2.0, 1e-30, -1e-30], np.float32)
    v = vals[rng.integers(0, len(vals), (3 * n, 3))]
    v += (rng.integers(0, 3, (3 * n, 3)) == 0) * rng.uniform(-0.01, 0.01, (3 * n, 3)).astype(np.float32)
    t = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    return [{"verts": v.astype(np.float32), "tris": t}]


@pytest.mark.parametrize("n", [1, 2, 4, 5, 33, 65, 300])
def test_tiny_scenes(n):
    soup = signed_zero_soup(n=n, seed=n)
    _vs_oracle(soup)


def test_c0_scene():
    sc = _vs_oracle(wl.scene_c0())
    assert sc.info()["triangles"] == 81920 + 12


def test_reference_fixture():
    _vs_oracle(scenes.reference_spheres())


@pytest.mark.parametrize("sm", [1, 4, 32, 256])
def test_small_node_threshold_does_not_change_the_tree(build_env, sm):
    build_env(small_max=sm)
    _vs_oracle(scenes.small_spheres(seed=4, n=5, subdiv=3))
    _vs_oracle(signed_zero_soup())


@pytest.mark.parametrize("wide_min,chunk", [(256, 256), (1000, 512), (3000, 4096)])
def test_wide_node_threshold_does_not_change_the_tree(build_env, wide_min, chunk):
    # nodes above wide_min triangles are spread over chunks (one workgroup each)
    build_env(wide_min=wide_min, chunk=chunk)
    _vs_oracle(scenes.small_spheres(seed=4, n=5, subdiv=3))
    _vs_oracle(signed_zero_soup())
    _vs_oracle(coplanar_plane())
    _vs_oracle(scenes.staircase())


def test_staircase_reaches_the_depth_limit():
    sc = _vs_oracle(scenes.staircase())
    assert sc.info()["max_depth"] == 200


@pytest.mark.parametrize("sm", [1, 32])
def test_coplanar_and_duplicated_triangles(build_env, sm):
    build_env(small_max=sm)
    _vs_oracle(coplanar_plane())


def test_transformed_primitives():
    meshes = scenes.small_spheres(seed=6, n=3, subdiv=2)
    prims = [(i, scenes.rot(i % 3, 0.3 * (i + 1)) @ np.diag([1.5, 1.0, 0.7, 1.0])) for i in range(len(meshes))]
    _vs_oracle(meshes, prims)


def test_empty_scene():
    sc = sk.Scene([{"verts": np.zeros((0, 3), np.float32), "tris": np.zeros((0, 3), np.uint32)}],
                  device_build=True)
    host = sk.Scene([{"verts": np.zeros((0, 3), np.float32), "tris": np.zeros((0, 3), np.uint32)}],
                    host_only=True)
    _same_tree((host.export_tree(), host.info()), sc)


def test_instanced_scene_is_unsupported():
    inst_meshes, inst_prims, _ = scenes.instanced_tetrahedra(4)
    with pytest.raises(sk.SkwError, match="instanced"):
        sk.Scene(inst_meshes, inst_prims, device_build=True)


def test_non_finite_boxes_are_unsupported():
    # a NaN coordinate on one vertex leaves the triangle's box to the other two (min / max keep
    # the current value against NaN), so the tree is still the reference's
    m = wl.box(1.0)
    v = m["verts"].copy()
    v[0, 0] = np.nan
    _vs_oracle([{"verts": v, "tris": m["tris"]}])
    # NaN on every vertex leaves (+inf, -inf) on that axis: rejected
    v[:, 0] = np.nan
    with pytest.raises(sk.SkwError, match="finite"):
        sk.Scene([{"verts": v, "tris": m["tris"]}], device_build=True)


def test_room_2m_triangles_and_traversal(torch_cuda):
    torch = torch_cuda
    s = wl.scene_room()
    host, dev = _vs_host(s["meshes"])
    host_dev = sk.Scene(s["meshes"])
    cam = s["camera"]
    rays = wl.camera_rays(512, 512, cam[0], cam[1], cam[2], cam[3])
    d_rays = torch.from_numpy(rays.view(np.uint8).reshape(-1)).cuda()
    outs = []
    for sc in (host_dev, dev):
        d_hits = empty_dev(torch, len(rays) * 32)
        sc.closest_hit(d_rays.data_ptr(), len(rays), d_hits.data_ptr())
        torch.cuda.synchronize()
        outs.append(from_dev(d_hits, np.uint8, len(rays) * 32))
    assert np.array_equal(outs[0], outs[1])
    assert (outs[0].view(np.uint32).reshape(len(rays), 8)[:, 0] != 0xFFFFFFFF).mean() > 0.99  # a closed room


def test_batched_scenes():
    specs = wl.batch_scenes()
    # a spread of sizes: every eighth scene and the largest
    sizes = [sum(len(m["tris"]) for m in sp[0]) for sp in specs]
    pick = sorted(set(list(range(0, len(specs), 8)) + [int(np.argmax(sizes))]))
    for i in pick:
        _vs_host(specs[i][0])


def test_device_built_scene_saves_and_loads_as_the_host_built(tmp_path):
    # the host copy of a device-built tree is made on first use (save, export)
    meshes = scenes.small_spheres(seed=8, n=4, subdiv=3)
    dev = sk.Scene(meshes, device_build=True)
    path = str(tmp_path / "dev.skws")
    dev.save(path)
    loaded = sk.Scene.load(path)
    host = sk.Scene(meshes, host_only=True)
    _same_tree((host.export_tree(), host.info()), loaded)
    _same_tree((host.export_tree(), host.info()), dev)
