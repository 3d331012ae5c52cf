"""The device kd build (skw_photon_map_build, skw_photon_set_map) against the host build
(skw_photon_map_create, node for node the oracle's libstdc++ KDTree::_Build).

Bar: permutation, split coordinates and node words bit-exact for every size class (single thread
below 33 items, one wave up to 2048, a 1024-thread workgroup above), with ties, identical points,
points on a plane, and 1M photons; gathers over the two maps bit-exact (payload and parent-split
tables)."""
import numpy as np
import pytest

import skwarka_amd as sk
from skwarka_amd import workloads as wl
from tests.conftest import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu


def _clouds():
    rng = np.random.default_rng(5)
    out = []
    for n in (1, 2, 3, 4, 5, 17, 32, 33, 34, 100, 2047, 2048, 2049, 2050, 5000, 70001):
        out.append(("random %d" % n, rng.uniform(-1, 1, (n, 3)).astype(np.float32)))
    out.append(("grid ties 20000", (rng.integers(0, 8, (20000, 3))).astype(np.float32)))
    out.append(("identical 6000", np.full((6000, 3), 0.25, np.float32)))
    plane = rng.uniform(-1, 1, (30000, 3)).astype(np.float32)
    plane[:, 1] = 0.5
    out.append(("plane 30000", plane))
    line = np.zeros((9000, 3), np.float32)
    line[:, 2] = np.repeat(np.arange(3000), 3).astype(np.float32)
    out.append(("repeated line 9000", line))
    return out


def _device_map(torch, pos, w=None, d=None, nrm=None):
    n = len(pos)
    dp = to_dev(torch, np.ascontiguousarray(pos, np.float32))
    dw = to_dev(torch, np.ascontiguousarray(w, np.float32)) if w is not None else None
    dd = to_dev(torch, np.ascontiguousarray(d, np.uint16)) if d is not None else None
    dn = to_dev(torch, np.ascontiguousarray(nrm, np.uint16)) if nrm is not None else None
    pm = sk.PhotonMap.build_on_device(n, dp.data_ptr(), dw.data_ptr() if dw is not None else None,
                                      dd.data_ptr() if dd is not None else None,
                                      dn.data_ptr() if dn is not None else None)
    torch.cuda.synchronize()
    return pm


def _same_tree(a, b):
    pa, sa, wa = a.export()
    pb, sb, wb = b.export()
    assert np.array_equal(pa, pb)
    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
    assert np.array_equal(wa, wb)


def test_device_build_equals_host_build(torch_cuda):
    for name, pos in _clouds():
        print("case: kd build %s" % name, flush=True)
        n = len(pos)
        host = sk.PhotonMap(pos, np.zeros((n, 3), np.float32), np.zeros(n, np.uint16), np.zeros(n, np.uint16))
        dev = _device_map(torch_cuda, pos)
        _same_tree(dev, host)


def test_device_build_million_photons_gathers_bit_exact(torch_cuda):
    print("case: kd build 1M photons (c1) and gathers over both maps", flush=True)
    mesh = wl.merge(wl.scene_c0())
    ph = wl.photons_on_mesh(mesh, 1_000_000, seed=1)
    host = sk.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"])
    dev = _device_map(torch_cuda, ph["pos"], ph["weight"], ph["dir"], ph["normal"])
    _same_tree(dev, host)
    rng = np.random.default_rng(2)
    idx = rng.integers(0, len(ph["pos"]), 20000)
    q = ph["pos"][idx].astype(np.float64)
    qn = wl.cdir_table()[ph["normal"][idx]]
    nq, k = len(q), 50
    dq, dn = to_dev(torch_cuda, q), to_dev(torch_cuda, qn)
    res = []
    for pm in (host, dev):
        ee, ei = empty_dev(torch_cuda, nq * 12), empty_dev(torch_cuda, nq * 12)
        ids, d2, cnt = empty_dev(torch_cuda, nq * k * 4), empty_dev(torch_cuda, nq * k * 8), empty_dev(torch_cuda, nq * 4)
        pm.irradiance(dq.data_ptr(), dn.data_ptr(), nq, k, 0.05, 1, 0.5, ee.data_ptr(), ei.data_ptr(), None,
                      ids.data_ptr(), d2.data_ptr(), cnt.data_ptr())
        torch_cuda.cuda.synchronize()
        res.append([from_dev(ee, np.uint32, nq * 3), from_dev(ei, np.uint32, nq * 3), from_dev(cnt, np.uint32, nq),
                    from_dev(ids, np.uint32, nq * k), from_dev(d2, np.uint64, nq * k)])
    cnt = res[0][2]
    mask = (np.arange(k)[None, :] < cnt[:, None]).ravel()
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    assert np.array_equal(res[0][2], res[1][2])
    assert np.array_equal(res[0][3][mask], res[1][3][mask]) and np.array_equal(res[0][4][mask], res[1][4][mask])


def test_shot_photon_set_map_is_the_host_tree(torch_cuda):
    print("case: kd build of a shot photon set", flush=True)
    s = wl.scene_room_small()
    scene = sk.Scene(s["meshes"])
    gdl = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    ps = gdl.shoot_photons(100000, (10 ** 12,) * 3, 1 << 16, seed=4)
    for j in (1, 2):
        m = ps.export(j)
        host = sk.PhotonMap(m["pos"], m["weight"], m["dir"], m["normal"])
        _same_tree(ps.photon_map(j), host)
