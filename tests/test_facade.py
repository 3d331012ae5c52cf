"""The C++ classes of include/skwarka.hpp (Scene / TriangleAccelerator / KDTree<Photon> with the
reference's names), driven one call at a time by tests/cpp/facade_check.cpp as the reference's
callers do: every result equals the oracle bit for bit (hit ids, t, b1, b2, occlusion, k-NN ids
and d^2 in heap order, filtered nearest photon), and the bounds equal GetWorldBounds'."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
import skwarka_amd as sk
from skwarka_amd import workloads as wl

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "facade_check")


def _write_input(path, meshes, prims, rays, ph, q, qn, k, r):
    with open(path, "wb") as f:
        f.write(np.uint32(len(meshes)).tobytes())
        for m in meshes:
            v = np.ascontiguousarray(m["verts"], np.float32)
            t = np.ascontiguousarray(m["tris"], np.uint32)
            f.write(np.uint32(len(v)).tobytes() + v.tobytes() + np.uint32(len(t)).tobytes() + t.tobytes())
            f.write(np.uint8(1 if m.get("invert", False) else 0).tobytes())
        f.write(np.uint32(len(prims)).tobytes())
        for mi, xf in prims:
            f.write(np.uint32(mi).tobytes())
            if xf is None:
                f.write(np.uint8(1).tobytes())
            else:
                m = np.asarray(xf, np.float64)
                f.write(np.uint8(0).tobytes() + m.tobytes() + np.linalg.inv(m).tobytes())
        f.write(np.uint32(len(rays)).tobytes() + np.ascontiguousarray(rays).tobytes())
        f.write(np.uint32(len(ph["pos"])).tobytes())
        for key, dt in (("pos", np.float32), ("weight", np.float32), ("dir", np.uint16), ("normal", np.uint16)):
            f.write(np.ascontiguousarray(ph[key], dt).tobytes())
        f.write(np.uint32(len(q)).tobytes() + np.ascontiguousarray(q, np.float64).tobytes())
        f.write(np.ascontiguousarray(qn, np.float64).tobytes())
        f.write(np.uint32(k).tobytes() + np.float64(r).tobytes())


def _read_output(path, n_rays, nq, k):
    buf = open(path, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(buf, dtype, count, off)
        off += a.nbytes
        return a

    bounds = take(np.float64, 6)
    ray_dt = np.dtype([("prim", "<u4"), ("tri", "<u4"), ("t", "<f8"), ("b1", "<f8"), ("b2", "<f8"), ("occ", "u1")])
    hits = take(ray_dt, n_rays)
    n_pts = int(take(np.uint64, 1)[0])
    q_dt = np.dtype([("count", "<u4"), ("set", [("id", "<u4"), ("d2", "<f8")], (k,))])
    out = np.frombuffer(buf, np.dtype([("plain", q_dt), ("filtered", q_dt), ("nearest", "<u4")]), nq, off)
    off += out.nbytes
    assert off == len(buf)
    return bounds, hits, n_pts, out


def _case():
    sphere = wl.icosphere(3, 1.0, (0.0, 0.0, 0.0), noise=0.02, rng=np.random.default_rng(5))
    box = wl.box(3.0)
    meshes = [sphere, box]
    shift = np.eye(4)
    shift[:3, :3] *= 0.4
    shift[:3, 3] = (1.6, 0.5, -0.7)
    prims = [(0, None), (1, None), (0, shift)]
    rng = np.random.default_rng(6)
    dirs = rng.normal(size=(200, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    rays = np.concatenate([wl.camera_rays(24, 24, (0.0, 0.0, -2.9), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0), 60.0),
                           wl.make_rays(rng.uniform(-2.5, 2.5, (200, 3)), dirs)])
    mesh = wl.merge(meshes)
    ph = wl.photons_on_mesh(mesh, 80000, seed=3)
    acc = oracle.Accel(meshes, prims)
    hits, geom = acc.closest_hit(rays, geom=True)
    keep = np.flatnonzero(hits["prim"] != sk.MISS)[:150]
    q = geom["p"][keep].copy()
    qn = geom["ng"][keep].copy()
    return meshes, prims, rays, ph, q, qn, acc


def test_facade_without_a_device_reports_an_error(tmp_path):
    if sk.lib().skw_device_count() > 0:
        pytest.skip("a HIP device is visible; the error path needs none")
    meshes, prims, rays, ph, q, qn, _ = _case()
    inp = str(tmp_path / "in.bin")
    _write_input(inp, meshes, prims, rays[:4], ph, q[:2], qn[:2], 4, 0.1)
    res = subprocess.run([BIN, inp, str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=60)
    assert res.returncode == 1 and "skw_scene_create" in res.stderr


@pytest.mark.gpu
def test_facade_scalar_calls_match_oracle(torch_cuda):
    print("case: C++ facade, scalar Intersect / IntersectTest / GetNearestPoints / GetNearestPoint", flush=True)
    meshes, prims, rays, ph, q, qn, acc = _case()
    k, r = 16, 0.1
    with tempfile.TemporaryDirectory() as tmp:
        inp, outp = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        _write_input(inp, meshes, prims, rays, ph, q, qn, k, r)
        res = subprocess.run([BIN, inp, outp], capture_output=True, text=True, timeout=110)
        assert res.returncode == 0, res.stderr
        bounds, got, n_pts, gathers = _read_output(outp, len(rays), len(q), k)
    assert np.array_equal(bounds, acc.bounds())
    ref = acc.closest_hit(rays)
    for f in ("prim", "tri"):
        assert np.array_equal(got[f], ref[f])
    for f in ("t", "b1", "b2"):
        assert np.array_equal(got[f].view(np.uint64), ref[f].view(np.uint64)), f
    assert np.array_equal(got["occ"], acc.any_hit(rays))
    assert n_pts == len(ph["pos"])
    tree = oracle.KDTree(ph["pos"])
    rids, rd2, rc = tree.knn(q, r, k)
    om = oracle.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"])
    _, _, fids, fd2, fc = om.irradiance(q, qn, k, r)
    for name, (ids, d2, cnt) in (("plain", (rids, rd2, rc)), ("filtered", (fids, fd2, fc))):
        g = gathers[name]
        assert np.array_equal(g["count"], cnt), name
        mask = np.arange(k)[None, :] < cnt[:, None]
        assert np.array_equal(np.where(mask, g["set"]["id"], 0), np.where(mask, ids, 0)), name
        assert np.array_equal(np.where(mask, g["set"]["d2"], 0).view(np.uint64),
                              np.where(mask, d2, 0).view(np.uint64)), name
    assert (gathers["filtered"]["count"] > 0).mean() > 0.5
    _, _, ids1, _, c1 = om.irradiance(q, qn, 1, r)
    assert np.array_equal(gathers["nearest"], np.where(c1 > 0, ids1[:, 0], sk.MISS))
