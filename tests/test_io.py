"""On-disk formats (skw_scene_save/load, skw_photon_map_save/load) on host-only structures:
round trips give identical trees, and corrupted, truncated, foreign or empty files are rejected.
The device round trips (hits, gathers, final gather after loading) are in test_gpu_io.py."""
import os

import numpy as np
import pytest

import skwarka_amd as sk
from skwarka_amd import workloads as wl
from tests import scenes


def _scene_export(sc):
    return sc.export_tree()


def test_scene_round_trip_host_only(tmp_path):
    meshes, prims, _ = scenes.instanced_tetrahedra(20)
    for ms, pr in ((wl.scene_c0(), None), (meshes, prims)):
        sc = sk.Scene(ms, prims=pr, host_only=True)
        path = str(tmp_path / "scene.skw")
        sc.save(path)
        back = sk.Scene.load(path, host_only=True)
        a, b = _scene_export(sc), _scene_export(back)
        for x, y in zip(a, b):
            assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
        assert sc.info() == back.info() and np.array_equal(sc.bounds(), back.bounds())


def test_photon_map_round_trip_host_only(tmp_path):
    mesh = wl.merge(wl.scene_c0())
    ph = wl.photons_on_mesh(mesh, 50000, seed=3)
    pm = sk.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"], host_only=True)
    path = str(tmp_path / "photons.skw")
    pm.save(path)
    back = sk.PhotonMap.load(path, host_only=True)
    assert back.n == pm.n
    for x, y in zip(pm.export(), back.export()):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_damaged_files_are_rejected(tmp_path):
    sc = sk.Scene(wl.scene_c0(), host_only=True)
    path = str(tmp_path / "scene.skw")
    sc.save(path)
    data = open(path, "rb").read()
    bad = [
        bytearray(data[:-1]),                          # truncated
        bytearray(data[:100]),                         # cut inside a section
        bytearray(b"SKWPHMAP" + data[8:]),             # another kind of file
        bytearray(data[:8] + b"\x02" + data[9:]),      # another format version
        bytearray(b""),                                # empty
    ]
    flipped = bytearray(data)
    flipped[len(data) // 2] ^= 0x10                    # one bit flipped
    bad.append(flipped)
    for i, d in enumerate(bad):
        p = str(tmp_path / ("bad%d.skw" % i))
        open(p, "wb").write(bytes(d))
        with pytest.raises(sk.SkwError):
            sk.Scene.load(p, host_only=True)
    with pytest.raises(sk.SkwError):
        sk.Scene.load(str(tmp_path / "missing.skw"), host_only=True)
    with pytest.raises(sk.SkwError):
        sk.PhotonMap.load(path, host_only=True)        # a scene file is not a photon map


def _fnv1a64(b):
    h = 0xcbf29ce484222325
    for c in b:
        h = ((h ^ c) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def _rehash(body):
    """A file body with a valid checksum: what a writer bug or a hand edit would leave."""
    return bytes(body) + _fnv1a64(body).to_bytes(8, "little")


def test_inconsistent_trees_with_valid_checksums_are_rejected(tmp_path):
    """Indices the traversal and lookup kernels follow without bounds checks, and the depths that
    size their stacks, are validated on load: a file whose checksum is right but whose tree is not
    is rejected before any upload."""
    import struct
    sc = sk.Scene(scenes.small_spheres(subdiv=2), host_only=True)
    path = str(tmp_path / "scene.skw")
    sc.save(path)
    body = bytearray(open(path, "rb").read()[:-8])
    # header: magic, version, byte order (16); root, max_depth, stack_need (u32 each); leaves (u64);
    # bounds (6 floats); six u64 counts; then section 1 (16-byte header) holding 32-byte nodes
    nodes0 = 16 + 12 + 8 + 24 + 48 + 16
    root = struct.unpack_from("<I", body, 16)[0]
    nn = struct.unpack_from("<Q", body, 16 + 12 + 8 + 24)[0]
    assert root < nn
    open(str(tmp_path / "ok.skw"), "wb").write(_rehash(body))
    assert sk.Scene.load(str(tmp_path / "ok.skw"), host_only=True).info() == sc.info()
    leaf = next(i for i in range(nn) if struct.unpack_from("<I", body, nodes0 + 32 * i + 28)[0] & 3 == 0)
    cases = {}
    b = bytearray(body); struct.pack_into("<I", b, 20, struct.unpack_from("<I", body, 20)[0] - 1)
    cases["understated max depth"] = b
    b = bytearray(body); struct.pack_into("<I", b, 24, struct.unpack_from("<I", body, 24)[0] - 1)
    cases["understated stack"] = b
    b = bytearray(body); struct.pack_into("<I", b, nodes0 + 32 * root + 24, nn)
    cases["child block past the nodes"] = b
    b = bytearray(body); struct.pack_into("<I", b, nodes0 + 32 * root + 24, root)
    cases["cycle to the root"] = b
    word = struct.unpack_from("<I", body, nodes0 + 32 * leaf + 28)[0]
    b = bytearray(body); struct.pack_into("<I", b, nodes0 + 32 * leaf + 28, (word & 7) | (0x1FFFFFF << 3))
    cases["leaf run past the triangles"] = b
    for name, data in cases.items():
        p = str(tmp_path / "bad.skw")
        open(p, "wb").write(_rehash(data))
        with pytest.raises(sk.SkwError, match="inconsistent"):
            sk.Scene.load(p, host_only=True)
    # photon map: magic, version, byte order (16); n (u64); max_depth (u32); section 1 of 16-byte nodes
    mesh = wl.merge(wl.scene_c0())
    ph = wl.photons_on_mesh(mesh, 2000, seed=3)
    pm = sk.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"], host_only=True)
    ppath = str(tmp_path / "photons.skw")
    pm.save(ppath)
    pbody = bytearray(open(ppath, "rb").read()[:-8])
    kd0 = 16 + 8 + 4 + 16
    b = bytearray(pbody); struct.pack_into("<I", b, 24, struct.unpack_from("<I", pbody, 24)[0] - 1)
    bad = {"understated depth": b}
    w0 = struct.unpack_from("<I", pbody, kd0 + 12)[0]
    b = bytearray(pbody); struct.pack_into("<I", b, kd0 + 12, (w0 & 7) | (2000 << 3))
    bad["right child past the nodes"] = b
    w1 = struct.unpack_from("<I", pbody, kd0 + 16 + 12)[0]
    b = bytearray(pbody); struct.pack_into("<I", b, kd0 + 16 + 12, (w1 & 3) | 4 | (1 << 3))  # right = itself
    bad["right child not after the node"] = b
    for name, data in bad.items():
        p = str(tmp_path / "badp.skw")
        open(p, "wb").write(_rehash(data))
        with pytest.raises(sk.SkwError, match="inconsistent"):
            sk.PhotonMap.load(p, host_only=True)
