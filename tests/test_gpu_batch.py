"""GPU checks of the batched multi-scene form: every scene's hits from the persistent queue kernel
equal skw_closest_hit on the same camera rays (bit-exact); the captured frame replays with the
same result, with eager work between replays; the autotuner returns a valid shape."""
import ctypes

import numpy as np
import pytest

import skwarka_amd as sk
from skwarka_amd import workloads as wl
from tests.conftest import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu


def _small_batch(n=7):
    specs = wl.batch_scenes(n=n, seed=17, tri_lo=500, tri_hi=60000, sizes=(16, 48, 64), spp=2)
    scenes, views = [], []
    for meshes, camspec, size, spp, seed in specs:
        scenes.append(sk.Scene(meshes))
        cam, _ = sk.camera(*camspec, size, size)
        views.append((cam, size, size, spp, seed))
    return scenes, views


def _batch_hits(torch, batch, i):
    """A scene's hit records, copied from the batch's device buffer (hipMemcpy device to device)."""
    ptr, n = batch.hits(i)
    hip = ctypes.CDLL("libamdhip64.so")
    dst = empty_dev(torch, n * 32)
    torch.cuda.synchronize()
    rc = hip.hipMemcpy(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(ptr), ctypes.c_size_t(n * 32), 3)
    assert rc == 0
    return from_dev(dst, sk.HIT_DTYPE, n)


def test_batch_hits_equal_single_scene_traversal(torch_cuda):
    print("case: batch of 7 scenes vs per-scene closest_hit", flush=True)
    scenes, views = _small_batch()
    batch = sk.Batch(scenes, views)
    batch.frame()
    batch.check()
    torch_cuda.cuda.synchronize()
    for i, sc in enumerate(scenes):
        got = _batch_hits(torch_cuda, batch, i)
        n = len(got)
        d_rays = empty_dev(torch_cuda, n * 64)
        batch.rays(i, d_rays.data_ptr())
        d_hits = empty_dev(torch_cuda, n * 32)
        sc.closest_hit(d_rays.data_ptr(), n, d_hits.data_ptr())
        torch_cuda.cuda.synchronize()
        ref = from_dev(d_hits, sk.HIT_DTYPE, n)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), "scene %d" % i
        assert (ref["prim"] != sk.MISS).mean() > 0.5


def test_batch_graph_replays_and_configs(torch_cuda):
    print("case: batch graph replays, explicit configs, autotune", flush=True)
    scenes, views = _small_batch(5)
    batch = sk.Batch(scenes, views)
    batch.frame()
    torch_cuda.cuda.synchronize()
    first = [_batch_hits(torch_cuda, batch, i) for i in range(len(scenes))]
    for shape in ((2, 32, 64, 8), (4, 64, 1024, 16), (1, 32, 4096, 32)):
        batch.configure(*shape)
        for _ in range(3):
            batch.frame()
            batch.check()
            x = torch_cuda.ones(1 << 16, device="cuda")  # eager work between replays
            assert x.sum().item() == float(1 << 16)
            torch_cuda.cuda.synchronize()
            for i in range(len(scenes)):
                assert np.array_equal(_batch_hits(torch_cuda, batch, i).view(np.uint8), first[i].view(np.uint8))
    cfg, ms = batch.autotune(frames=1)
    assert ms > 0 and cfg["chunk_rays"] % 64 == 0
    batch.frame()
    lat = batch.latency_ns()
    assert np.all(lat > 0)
    assert batch.ray_count() == sum(v[1] * v[2] * v[3] for v in views)
    for i in range(len(scenes)):
        assert np.array_equal(_batch_hits(torch_cuda, batch, i).view(np.uint8), first[i].view(np.uint8))


def test_batch_rejects_bad_shapes(torch_cuda):
    scenes, views = _small_batch(2)
    batch = sk.Batch(scenes, views)
    for shape in ((3, 32, 256, 16), (1, 16, 256, 16), (1, 32, 100, 16)):
        with pytest.raises(sk.SkwError):
            batch.configure(*shape)


def test_batch_with_photon_gathers(torch_cuda):
    print("case: batch of 3 scenes + 2 photon-gather jobs vs skw_photon_irradiance", flush=True)
    torch = torch_cuda
    scenes, views = _small_batch(3)
    jobs, refs = [], []
    for seed, (nph, nq, k, r) in enumerate(((200000, 20000, 100, 0.05), (30000, 3000, 8, 0.2))):
        mesh = wl.merge(wl.scene_c0())
        ph = wl.photons_on_mesh(mesh, nph, seed=seed + 1)
        pm = sk.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"])
        rng = np.random.default_rng(seed)
        sel = rng.choice(len(ph["pos"]), nq, replace=False)
        p = ph["pos"][sel].astype(np.float64)
        n = wl.cdir_table()[ph["normal"][sel]]
        dp, dn = to_dev(torch, p), to_dev(torch, n)
        outs = [empty_dev(torch, nq * 12) for _ in range(3)]
        jobs.append((pm, dp.data_ptr(), dn.data_ptr(), nq, k, r, 2, 0.5, outs[0].data_ptr(), outs[1].data_ptr(),
                     outs[2].data_ptr()))
        ref = [empty_dev(torch, nq * 12) for _ in range(3)]
        pm.irradiance(dp.data_ptr(), dn.data_ptr(), nq, k, r, 2, 0.5, ref[0].data_ptr(), ref[1].data_ptr(),
                      ref[2].data_ptr())
        refs.append((outs, ref, (dp, dn), nq))
    batch = sk.Batch(scenes, views, jobs)
    torch.cuda.synchronize()
    for shape in ((1, 32, 256, 16), (4, 64, 4096, 8)):
        batch.configure(*shape)
        for _ in range(2):
            batch.frame()
            batch.check()
            torch.cuda.synchronize()
            for outs, ref, _, nq in refs:
                for o, r in zip(outs, ref):
                    assert np.array_equal(from_dev(o, np.uint32, nq * 3), from_dev(r, np.uint32, nq * 3))
            for i, sc in enumerate(scenes):
                got = _batch_hits(torch, batch, i)
                d_rays = empty_dev(torch, len(got) * 64)
                batch.rays(i, d_rays.data_ptr())
                d_hits = empty_dev(torch, len(got) * 32)
                sc.closest_hit(d_rays.data_ptr(), len(got), d_hits.data_ptr())
                torch.cuda.synchronize()
                assert np.array_equal(got.view(np.uint8), from_dev(d_hits, sk.HIT_DTYPE, len(got)).view(np.uint8))
    lat = batch.latency_ns()
    assert len(lat) == 5 and np.all(lat > 0)
    assert from_dev(refs[0][1][0], np.float32, 100).max() > 0
