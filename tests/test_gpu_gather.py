"""GPU parity of the photon-map gather (through the C ABI) against the oracle.

Bar: k-NN index sets bit-exact, in the same final heap order as a libstdc++ build of the
reference (so the estimator sums in the same order); d^2 bit-exact; E_ext / E_int bit-exact;
radiance = rho/pi * E_ext within 1e-6 relative (one float rounding).
"""
import numpy as np
import pytest

import oracle
import skwarka_amd as sk
from skwarka_amd import workloads as wl
from tests.conftest import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu
MISS = 0xFFFFFFFF


@pytest.fixture(scope="module")
def c1_data():
    mesh = wl.merge(wl.scene_c0())
    ph = wl.photons_on_mesh(mesh, 1_000_000, seed=1)
    om = oracle.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"])
    # gather points: C0 primary hits (every 8th pixel) with their geometric normals
    acc = oracle.Accel(wl.scene_c0())
    rays = wl.rays_c0()[::8]
    h, g = acc.closest_hit(rays, geom=True)
    keep = h["prim"] != MISS
    return ph, om, g["p"][keep].copy(), g["ng"][keep].copy()


def gpu_irradiance(torch, pm, p, n, k, r, paths, rho):
    nq = len(p)
    dp, dn = to_dev(torch, p), to_dev(torch, n)
    ee, ei, rad = empty_dev(torch, nq * 12), empty_dev(torch, nq * 12), empty_dev(torch, nq * 12)
    ids, d2, cnt = empty_dev(torch, nq * k * 4), empty_dev(torch, nq * k * 8), empty_dev(torch, nq * 4)
    pm.irradiance(dp.data_ptr(), dn.data_ptr(), nq, k, r, paths, rho, ee.data_ptr(), ei.data_ptr(), rad.data_ptr(),
                  ids.data_ptr(), d2.data_ptr(), cnt.data_ptr())
    torch.cuda.synchronize()
    return (from_dev(ee, np.float32, nq * 3).reshape(nq, 3), from_dev(ei, np.float32, nq * 3).reshape(nq, 3),
            from_dev(rad, np.float32, nq * 3).reshape(nq, 3), from_dev(ids, np.uint32, nq * k).reshape(nq, k),
            from_dev(d2, np.float64, nq * k).reshape(nq, k), from_dev(cnt, np.uint32, nq))


def check_sets(cnt, ids, d2, rc, rids, rd2):
    assert np.array_equal(cnt, rc)
    mask = np.arange(ids.shape[1])[None, :] < cnt[:, None]
    assert np.array_equal(np.where(mask, ids, 0), np.where(mask, rids, 0))
    assert np.array_equal(np.where(mask, d2, 0).view(np.uint64), np.where(mask, rd2, 0).view(np.uint64))


@pytest.mark.parametrize("k,r", [(100, 0.05), (8, 0.05), (32, 0.2)])
def test_irradiance_c1_photons(torch_cuda, c1_data, k, r):
    print("case: irradiance 1M photons k=%d r=%g" % (k, r), flush=True)
    ph, om, p, n = c1_data
    pm = sk.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"])
    ee, ei, rad, ids, d2, cnt = gpu_irradiance(torch_cuda, pm, p, n, k, r, 1, 0.5)
    ree, rei, rids, rd2, rc = om.irradiance(p, n, k, r, paths=1)
    check_sets(cnt, ids, d2, rc, rids, rd2)
    assert np.array_equal(ee.view(np.uint32), ree.view(np.uint32))
    assert np.array_equal(ei.view(np.uint32), rei.view(np.uint32))
    expect = (0.5 * 0.31830988618379067154 * ree.astype(np.float64)).astype(np.float32)
    assert np.allclose(rad, expect, rtol=1e-6, atol=0)
    if k == 8:
        assert (cnt == k).mean() > 0.5  # the bounded heap really was exercised
    assert cnt.max() > 0


def test_knn_and_nearest_without_filter(torch_cuda, c1_data):
    print("case: knn / nearest, filter off and on", flush=True)
    ph, om, p, n = c1_data
    pm = sk.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"])
    tree = oracle.KDTree(ph["pos"])  # KDTree<Point3D_f>: same tree as the photon map
    nq = len(p)
    dp, dn = to_dev(torch_cuda, p), to_dev(torch_cuda, n)
    k = 16
    ids, d2, cnt = empty_dev(torch_cuda, nq * k * 4), empty_dev(torch_cuda, nq * k * 8), empty_dev(torch_cuda, nq * 4)
    pm.knn(dp.data_ptr(), 0, nq, k, 0.1, 0, ids.data_ptr(), d2.data_ptr(), cnt.data_ptr())
    torch_cuda.cuda.synchronize()
    rids, rd2, rc = tree.knn(p, 0.1, k)
    check_sets(from_dev(cnt, np.uint32, nq), from_dev(ids, np.uint32, nq * k).reshape(nq, k),
               from_dev(d2, np.float64, nq * k).reshape(nq, k), rc, rids, rd2)
    near = empty_dev(torch_cuda, nq * 4)
    pm.nearest(dp.data_ptr(), 0, nq, 0.02, 0, near.data_ptr())
    torch_cuda.cuda.synchronize()
    assert np.array_equal(from_dev(near, np.uint32, nq), tree.nearest(p, 0.02))
    # filtered nearest: the first k=1 set of the filtered lookup equals the filtered nearest
    pm.nearest(dp.data_ptr(), dn.data_ptr(), nq, 0.05, 1, near.data_ptr())
    torch_cuda.cuda.synchronize()
    got = from_dev(near, np.uint32, nq)
    _, _, ids1, _, c1 = om.irradiance(p, n, 1, 0.05)
    assert np.array_equal(got, np.where(c1 > 0, ids1[:, 0], MISS))


def test_duplicates_small_and_empty_maps(torch_cuda):
    print("case: duplicate photons, 1- and 0-photon maps", flush=True)
    rng = np.random.default_rng(2)
    pos = rng.uniform(-1, 1, (20000, 3)).astype(np.float32)
    pos[:, 2] = 0.0
    pos = np.concatenate([pos, pos[:2000], pos[:500]])
    nrm = np.full(len(pos), wl.cdir_encode(np.array([[0.0, 0.0, 1.0]]))[0], np.uint16)
    dirs = wl.cdir_encode(rng.normal(size=(len(pos), 3)) / 1.0)
    dirs = wl.cdir_encode(np.asarray([v / np.linalg.norm(v) for v in rng.normal(size=(len(pos), 3))]))
    w = rng.uniform(0, 1, (len(pos), 3)).astype(np.float32)
    q = np.c_[rng.uniform(-1, 1, (4000, 2)), np.zeros(4000)]
    qn = np.tile([0.0, 0.0, 1.0], (4000, 1))
    om = oracle.PhotonMap(pos, w, dirs, nrm)
    pm = sk.PhotonMap(pos, w, dirs, nrm)
    for k, r in ((50, 0.1), (7, 0.1)):
        ee, ei, rad, ids, d2, cnt = gpu_irradiance(torch_cuda, pm, q, qn, k, r, 2, 0.5)
        ree, rei, rids, rd2, rc = om.irradiance(q, qn, k, r, paths=2)
        check_sets(cnt, ids, d2, rc, rids, rd2)
        assert np.array_equal(ee.view(np.uint32), ree.view(np.uint32))
        assert np.array_equal(ei.view(np.uint32), rei.view(np.uint32))
    for m in (1, 0):
        om1 = oracle.PhotonMap(pos[:m], w[:m], dirs[:m], nrm[:m])
        pm1 = sk.PhotonMap(pos[:m], w[:m], dirs[:m], nrm[:m])
        ee, ei, rad, ids, d2, cnt = gpu_irradiance(torch_cuda, pm1, q[:100], qn[:100], 4, 2.0, 1, 0.5)
        ree, rei, rids, rd2, rc = om1.irradiance(q[:100], qn[:100], 4, 2.0)
        check_sets(cnt, ids, d2, rc, rids, rd2)
        assert np.array_equal(ee, ree)


def test_cone_and_caustic_estimates_match_oracle(torch_cuda, c1_data):
    print("case: cone-filtered irradiance and caustic radiance, 1M photons", flush=True)
    ph, om, p, n = c1_data
    pm = sk.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"])
    nq = len(p)
    dp, dn = to_dev(torch_cuda, p), to_dev(torch_cuda, n)
    e = empty_dev(torch_cuda, nq * 12)
    for k, r, kc in ((100, 0.05, 1.0), (16, 0.2, 1.5)):
        pm.cone_irradiance(dp.data_ptr(), dn.data_ptr(), nq, k, r, 2, kc, e.data_ptr())
        torch_cuda.cuda.synchronize()
        got = from_dev(e, np.float32, nq * 3).reshape(nq, 3)
        ref = om.cone(p, n, k, r, 2, kc)
        # float outputs of double sums over the same (bit-exact) sets; the weights take a sqrt,
        # so the bar is the floating-point one: 1e-6 relative
        assert np.all(np.abs(got - ref) <= 1e-6 * np.abs(ref)), np.abs(got - ref).max()
        assert ref.max() > 0
    # caustic radiance toward a direction on the normal's side for half the points, the other side
    # for the rest
    rng = np.random.default_rng(4)
    wo = n + rng.normal(0, 0.3, n.shape)
    wo /= np.linalg.norm(wo, axis=1)[:, None]
    wo[1::2] *= -1.0
    dwo = to_dev(torch_cuda, wo)
    rad = empty_dev(torch_cuda, nq * 24)
    for k, r in ((100, 0.05), (5, 0.1)):
        pm.caustic_radiance(dp.data_ptr(), dn.data_ptr(), dwo.data_ptr(), nq, k, r, 3, 0.5, rad.data_ptr())
        torch_cuda.cuda.synchronize()
        got = from_dev(rad, np.float64, nq * 3).reshape(nq, 3)
        ref = om.caustic(p, n, wo, k, r, 3, 0.5)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), np.abs(got - ref).max()
        assert ref.max() > 0


@pytest.mark.parametrize("k,r", [(100, 0.05), (8, 0.05), (32, 0.2)])
def test_irradiance_without_sets_and_lookup_counts(torch_cuda, c1_data, k, r):
    """The disc estimate without returned sets (slots written only for sets that fill, the lookup
    repeated for those) equals the oracle bit for bit; the counting variant reports the oracle's
    visits, photons inside the radius and set sizes."""
    print("case: irradiance without sets + lookup counts, 1M photons k=%d r=%g" % (k, r), flush=True)
    torch = torch_cuda
    ph, om, p, n = c1_data
    pm = sk.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"])
    nq = len(p)
    dp, dn = to_dev(torch, p), to_dev(torch, n)
    ee, ei, rad = empty_dev(torch, nq * 12), empty_dev(torch, nq * 12), empty_dev(torch, nq * 12)
    cnt = empty_dev(torch, nq * 4)
    pm.irradiance(dp.data_ptr(), dn.data_ptr(), nq, k, r, 1, 0.5, ee.data_ptr(), ei.data_ptr(), rad.data_ptr(),
                  None, None, cnt.data_ptr())
    torch.cuda.synchronize()
    ree, rei, _, _, rc = om.irradiance(p, n, k, r, paths=1, with_sets=False)
    assert np.array_equal(from_dev(ee, np.uint32, nq * 3), ree.view(np.uint32).reshape(-1))
    assert np.array_equal(from_dev(ei, np.uint32, nq * 3), rei.view(np.uint32).reshape(-1))
    assert np.array_equal(from_dev(cnt, np.uint32, nq), rc)
    if k == 8:
        assert (rc == k).mean() > 0.5  # the repeated lookups of filled sets were exercised
    if k == 100:
        assert (rc < k).all()  # none fills: no slot written at all
    stats = empty_dev(torch, nq * 12)
    ee2 = empty_dev(torch, nq * 12)
    pm.irradiance_stats(dp.data_ptr(), dn.data_ptr(), nq, k, r, 1, 0.5, ee2.data_ptr(), ei.data_ptr(), None,
                        stats.data_ptr())
    torch.cuda.synchronize()
    got = from_dev(stats, np.uint32, nq * 3).reshape(nq, 3)
    assert np.array_equal(got, om.lookup_counts(p, n, k, r))
    assert np.array_equal(from_dev(ee2, np.uint32, nq * 3), ree.view(np.uint32).reshape(-1))
