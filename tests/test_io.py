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
