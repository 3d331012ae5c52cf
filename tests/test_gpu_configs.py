"""GPU parity at the benchmark configurations' full sizes, through the C ABI, against the oracle.

* C2: the seeded 2,026,260-triangle room at 1024x1024, 16 spp, 4 light + 4 BSDF samples: every
  pixel's radiance within 1e-5 relative of the oracle's full frame, the three ray counts equal, and
  the primary hits of all 1024x1024 camera rays (1 spp) bit-exact.
* C3: all 64 batched scenes (1.3k-874k triangles, 128/256/512 square frames, 4 spp) in one
  persistent launch: every scene's hit records bit-exact against the oracle's traversal on a
  strided sample of its rays (every 61st), after a captured-frame replay.
* C4: the room at 512x512, 8 spp, L2 loss against a perturbed-parameter target: the adjoint (tape
  and recomputed) equals the oracle's FP64 forward-mode gradient on a band of 64 rows (the adjoint
  input is zero elsewhere) within 1e-6 of the largest entry, and central finite differences of the
  GPU estimator on all 132 parameters (step 1e-4 relative) within 1e-3 relative.
"""
import concurrent.futures
import ctypes

import numpy as np
import pytest

import oracle
import skwarka_amd as sk
from skwarka_amd import workloads as wl
from tests.conftest import empty_dev, from_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def room():
    s = wl.scene_room()
    acc = oracle.Accel(s["meshes"])
    scene = sk.Scene(s["meshes"])
    return s, acc, scene


def assert_close_rel(got, ref, rel=1e-5):
    g = got.astype(np.float64)
    r = ref.astype(np.float64)
    err = np.abs(g - r)
    bad = err > rel * np.maximum(np.abs(g), np.abs(r)) + 1e-30
    assert not bad.any(), "%d values off; worst rel %g" % (bad.sum(), (err / np.maximum(np.abs(r), 1e-30)).max())


def test_c2_full_frame_matches_oracle(torch_cuda, room):
    print("case: C2 room 1024x1024 x 16 spp, 4+4 samples, whole frame vs oracle", flush=True)
    torch = torch_cuda
    s, acc, scene = room
    w = h = 1024
    spp, seed = 16, 1234
    cam, cam15 = sk.camera(*s["camera"], w, h)
    gdl = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    d_rgb = empty_dev(torch, w * h * 12)
    cnt = gdl.render(cam, w, h, spp, d_rgb.data_ptr(), 4, 4, seed)
    torch.cuda.synchronize()
    got = from_dev(d_rgb, np.float32, w * h * 3).reshape(w * h, 3)
    odl = oracle.DLScene(acc, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    ref, rcnt = odl.render(cam15, w, h, spp, 4, 4, seed)
    assert np.array_equal(np.asarray(cnt, np.int64), rcnt.astype(np.int64)), (cnt, rcnt)
    assert rcnt[1] > 10 ** 7 and rcnt[2] > 10 ** 7
    assert_close_rel(got, ref)
    assert ref.max() > 0
    # primary hits of every pixel at 1 spp: the camera rays of this view as the batch form makes them
    batch = sk.Batch([scene], [(cam, w, h, 1, seed)])
    n = w * h
    d_rays, d_hits = empty_dev(torch, n * 64), empty_dev(torch, n * 32)
    batch.rays(0, d_rays.data_ptr())
    scene.closest_hit(d_rays.data_ptr(), n, d_hits.data_ptr())
    torch.cuda.synchronize()
    rays = from_dev(d_rays, wl.RAY_DTYPE(), n)
    hits = from_dev(d_hits, sk.HIT_DTYPE, n)
    ref_hits = acc.closest_hit(rays)
    assert np.array_equal(hits.view(np.uint8), ref_hits.view(np.uint8))
    assert (hits["prim"] != sk.MISS).mean() > 0.99
    batch.close()


def _batch_hits(torch, batch, i):
    ptr, n = batch.hits(i)
    hip = ctypes.CDLL("libamdhip64.so")
    dst = empty_dev(torch, n * 32)
    torch.cuda.synchronize()
    assert hip.hipMemcpy(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(ptr), ctypes.c_size_t(n * 32), 3) == 0
    return from_dev(dst, sk.HIT_DTYPE, n)


def test_c3_all_64_scenes_match_oracle(torch_cuda):
    print("case: C3 64 batched scenes, one persistent launch, strided rays vs oracle", flush=True)
    torch = torch_cuda
    specs = wl.batch_scenes()
    assert len(specs) == 64
    scenes, views = [], []
    for meshes, camspec, size, spp, seed in specs:
        scenes.append(sk.Scene(meshes))
        cam, _ = sk.camera(*camspec, size, size)
        views.append((cam, size, size, spp, seed))
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        accs = list(ex.map(lambda sp: oracle.Accel(sp[0]), specs))
    batch = sk.Batch(scenes, views)
    for _ in range(2):  # the captured frame, replayed
        batch.frame()
        batch.check()
    torch.cuda.synchronize()
    tris = 0
    for i, (sc, acc) in enumerate(zip(scenes, accs)):
        got = _batch_hits(torch, batch, i)
        n = len(got)
        assert n == views[i][1] * views[i][2] * views[i][3]
        d_rays = empty_dev(torch, n * 64)
        batch.rays(i, d_rays.data_ptr())
        torch.cuda.synchronize()
        rays = from_dev(d_rays, wl.RAY_DTYPE(), n)
        sel = np.arange(i % 61, n, 61)
        ref = acc.closest_hit(np.ascontiguousarray(rays[sel]))
        assert np.array_equal(got[sel].view(np.uint8), ref.view(np.uint8)), "scene %d" % i
        tris += sc.info()["triangles"]
    assert tris > 7_000_000
    lat = batch.latency_ns()
    assert np.all(lat > 0)
    batch.close()


def test_c4_gradient_at_config_size(torch_cuda, room):
    print("case: C4 room 512x512 x 8 spp, 132-parameter gradient vs oracle band and central FD", flush=True)
    torch = torch_cuda
    s, acc, scene = room
    w = h = 512
    spp, seed = 8, 1234
    cam, cam15 = sk.camera(*s["camera"], w, h)
    gdl = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"],
                     prim_group=s["prim_group"], n_groups=s["n_groups"])
    assert gdl.n_params == 132
    th0 = s["theta0"]
    tt = wl.perturbed_theta(th0)
    target = torch.empty((w * h, 3), dtype=torch.float64, device="cuda")
    gdl.render_replay(tt, tt, cam, w, h, spp, target.data_ptr(), 4, 4, 4321)
    img = torch.empty_like(target)

    def replay(te):
        gdl.render_replay(te, th0, cam, w, h, spp, img.data_ptr(), 4, 4, seed)
        torch.cuda.synchronize()
        return img

    adj_full = (2.0 * (replay(th0) - target)).contiguous()
    # tape and recomputed adjoints of the full loss
    g_rc = torch.empty(132, dtype=torch.float64, device="cuda")
    gdl.adjoint(th0, cam, w, h, spp, adj_full.data_ptr(), g_rc.data_ptr(), 4, 4, seed)
    img2 = torch.empty_like(target)
    tape = gdl.render_tape(th0, cam, w, h, spp, img2.data_ptr(), 1 << 36, 4, 4, seed)
    g_tape = torch.empty_like(g_rc)
    tape.adjoint(adj_full.data_ptr(), g_tape.data_ptr())
    torch.cuda.synchronize()
    g = g_rc.cpu().numpy()
    assert torch.equal(img2, replay(th0))
    assert np.all(np.abs(g_tape.cpu().numpy() - g) <= 1e-10 * np.abs(g).max())
    # FP64 oracle gradient on rows 224-287; the GPU adjoint of the same band (zero input elsewhere)
    r0, r1 = 224, 288
    p0, p1 = r0 * w, r1 * w
    adj_band = torch.zeros_like(adj_full)
    adj_band[p0:p1] = adj_full[p0:p1]
    g_band = torch.empty_like(g_rc)
    tape.adjoint(adj_band.data_ptr(), g_band.data_ptr())
    torch.cuda.synchronize()
    tape.close()
    odl = oracle.DLScene(acc, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"],
                         prim_group=s["prim_group"], n_groups=s["n_groups"])
    g_ref, img_ref = odl.gradient(th0, adj_band[p0:p1].cpu().numpy(), cam15, w, h, spp, 4, 4, seed, p0, p1)
    band = g_band.cpu().numpy()
    scale = np.abs(g_ref).max()
    assert scale > 0 and np.all(np.abs(band - g_ref) <= 1e-6 * scale), np.abs(band - g_ref).max() / scale
    assert np.all(np.abs(img_ref - replay(th0).cpu().numpy()[p0:p1]) <= 1e-5 * np.abs(img_ref) + 1e-12)
    # central differences of the GPU estimator on every parameter, sampling held at theta0

    def loss_at(te):
        return float(((replay(te) - target) ** 2).sum().item())

    fd = np.zeros(132)
    for j in range(132):
        step = 1e-4 * max(abs(th0[j]), 1e-3)
        tp, tm = th0.copy(), th0.copy()
        tp[j] += step
        tm[j] -= step
        fd[j] = (loss_at(tp) - loss_at(tm)) / (tp[j] - tm[j])
    fscale = np.abs(fd).max()
    assert np.all(np.abs(g - fd) <= 1e-3 * np.abs(fd) + 1e-6 * fscale), np.max(np.abs(g - fd) / np.abs(fd))
    assert np.count_nonzero(g[:64]) > 30 and np.count_nonzero(g[64:128]) > 10 and np.all(g[128:] != 0)


def test_c3_batch_with_photon_gathers_matches_oracle(torch_cuda):
    """The 64 scenes and 8 photon-gather jobs (250k photons each, 65,536 queries, k = 100) in one
    persistent launch, replayed: every job's disc estimates bit-exact against the oracle on every
    7th query, and each scene's hits against the oracle on every 127th ray."""
    print("case: C3 64 scenes + 8 photon-gather jobs in one launch vs oracle", flush=True)
    torch = torch_cuda
    specs = wl.batch_scenes()
    scenes, views = [], []
    for meshes, camspec, size, spp, seed in specs:
        scenes.append(sk.Scene(meshes))
        cam, _ = sk.camera(*camspec, size, size)
        views.append((cam, size, size, spp, seed))
    jobs = wl.batch_gather_jobs(specs)
    gathers, keep = [], []
    for jb in jobs:
        pm = sk.PhotonMap(jb["ph"]["pos"], jb["ph"]["weight"], jb["ph"]["dir"], jb["ph"]["normal"])
        nq = len(jb["p"])
        dp = torch.from_numpy(jb["p"].view(np.uint8).reshape(-1).copy()).cuda()
        dn = torch.from_numpy(jb["n"].view(np.uint8).reshape(-1).copy()).cuda()
        outs = [empty_dev(torch, nq * 12) for _ in range(3)]
        gathers.append((pm, dp.data_ptr(), dn.data_ptr(), nq, jb["k"], jb["r"], jb["paths"], jb["rho"],
                        outs[0].data_ptr(), outs[1].data_ptr(), outs[2].data_ptr()))
        keep.append((pm, dp, dn, outs))
    batch = sk.Batch(scenes, views, gathers)
    for shape in ((1, 64, 256, 16), (2, 64, 1024, 8)):
        batch.configure(*shape)
        for _ in range(2):
            batch.frame()
            batch.check()
    torch.cuda.synchronize()
    for jb, (pm, dp, dn, outs) in zip(jobs, keep):
        om = oracle.PhotonMap(jb["ph"]["pos"], jb["ph"]["weight"], jb["ph"]["dir"], jb["ph"]["normal"])
        nq = len(jb["p"])
        sel = np.arange(0, nq, 7)
        ree, rei, _, _, rc = om.irradiance(jb["p"][sel], jb["n"][sel], jb["k"], jb["r"], paths=1, with_sets=False)
        ee = from_dev(outs[0], np.float32, nq * 3).reshape(nq, 3)[sel]
        ei = from_dev(outs[1], np.float32, nq * 3).reshape(nq, 3)[sel]
        assert np.array_equal(ee.view(np.uint32), ree.view(np.uint32))
        assert np.array_equal(ei.view(np.uint32), rei.view(np.uint32))
        assert rc.mean() > 20
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        accs = list(ex.map(lambda sp: oracle.Accel(sp[0]), specs))
    for i, acc in enumerate(accs):
        got = _batch_hits(torch, batch, i)
        n = len(got)
        d_rays = empty_dev(torch, n * 64)
        batch.rays(i, d_rays.data_ptr())
        torch.cuda.synchronize()
        rays = from_dev(d_rays, wl.RAY_DTYPE(), n)
        sel = np.arange(i % 127, n, 127)
        ref = acc.closest_hit(np.ascontiguousarray(rays[sel]))
        assert np.array_equal(got[sel].view(np.uint8), ref.view(np.uint8)), "scene %d" % i
    lat = batch.latency_ns()
    assert len(lat) == 72 and np.all(lat > 0)
    batch.close()
