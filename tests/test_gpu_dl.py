"""GPU parity of the direct-lighting render (skw_dl_render, through the C ABI) against the oracle's
restatement of DirectLightingIntegrator on the same sample stream.

Bar: per-pixel radiance within 1e-5 relative (the GPU's pow / sin / cos / atan may differ from
glibc's in the last bits); ray counts identical (same emission decisions); primary-ray hit ids
bit-exact (checked through the traversal API on the same camera rays)."""
import numpy as np
import pytest

import oracle
import skwarka_amd as sk
from skwarka_amd import workloads as wl
from tests.conftest import empty_dev, from_dev

pytestmark = pytest.mark.gpu


def render_both(torch, s, w, h, spp, nls=4, nbs=4, seed=7):
    acc = oracle.Accel(s["meshes"])
    odl = oracle.DLScene(acc, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    cam, cam15 = sk.camera(*s["camera"], w, h)
    ref, rcnt = odl.render(cam15, w, h, spp, nls, nbs, seed)
    scene = sk.Scene(s["meshes"])
    gdl = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    d_rgb = empty_dev(torch, w * h * 12)
    cnt = gdl.render(cam, w, h, spp, d_rgb.data_ptr(), nls, nbs, seed)
    torch.cuda.synchronize()
    got = from_dev(d_rgb, np.float32, w * h * 3).reshape(w * h, 3)
    return got, cnt, ref, rcnt


def assert_close_rel(got, ref, rel=1e-5):
    g = got.astype(np.float64)
    r = ref.astype(np.float64)
    err = np.abs(g - r)
    tol = rel * np.maximum(np.abs(g), np.abs(r)) + 1e-30
    bad = err > tol
    assert not bad.any(), "%d pixels off; worst rel %g" % (bad.sum(), (err / np.maximum(np.abs(r), 1e-30)).max())


def test_dl_small_room_matches_oracle(torch_cuda):
    print("case: direct lighting, small room 96x64, 3 spp", flush=True)
    s = wl.scene_room_small()
    got, cnt, ref, rcnt = render_both(torch_cuda, s, 96, 64, 3)
    assert np.array_equal(cnt, rcnt)
    assert rcnt[1] > 1000 and rcnt[2] > 1000
    assert ref.max() > 0
    assert_close_rel(got, ref)


def test_dl_sample_counts_and_seeds(torch_cuda):
    print("case: direct lighting, 1/2 light+bsdf samples, other seed", flush=True)
    s = wl.scene_room_small(seed=11)
    for nls, nbs in ((1, 2), (3, 1)):
        got, cnt, ref, rcnt = render_both(torch_cuda, s, 40, 30, 2, nls, nbs, seed=99)
        assert np.array_equal(cnt, rcnt)
        assert_close_rel(got, ref)


def test_dl_rejects_bad_parameters(torch_cuda):
    s = wl.scene_room_small()
    scene = sk.Scene(s["meshes"])
    gdl = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    cam, _ = sk.camera(*s["camera"], 8, 8)
    d_rgb = empty_dev(torch_cuda, 8 * 8 * 12)
    with pytest.raises(sk.SkwError):
        gdl.render(cam, 8, 8, 1, d_rgb.data_ptr(), 0, 4)
    with pytest.raises(sk.SkwError):
        gdl.render(cam, 0, 8, 1, d_rgb.data_ptr())


def test_profiled_frame_is_the_render_and_counts_the_walk(torch_cuda):
    print("case: profiled direct-lighting frame, counting walk", flush=True)
    torch = torch_cuda
    s = wl.scene_room_small()
    scene = sk.Scene(s["meshes"])
    gdl = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    w, h, spp, seed = 40, 32, 2, 21
    cam, _ = sk.camera(*s["camera"], w, h)
    a, b, c = (empty_dev(torch, w * h * 12) for _ in range(3))
    rays = gdl.render(cam, w, h, spp, a.data_ptr(), 3, 2, seed)
    p1 = gdl.render_profiled(cam, w, h, spp, b.data_ptr(), 3, 2, seed, False)
    p2 = gdl.render_profiled(cam, w, h, spp, c.data_ptr(), 3, 2, seed, True)
    torch.cuda.synchronize()
    img = from_dev(a, np.uint32, w * h * 3)
    assert np.array_equal(img, from_dev(b, np.uint32, w * h * 3))
    assert np.array_equal(img, from_dev(c, np.uint32, w * h * 3))
    assert p1["rays"] == p2["rays"] == [int(x) for x in rays]
    assert all(v > 0 for v in p1["stage_ms"].values())
    assert p1["node_tests"] == [0, 0, 0] and all(v > 0 for v in p2["node_tests"])
    # the primary rays are the camera rays the batch form generates for the same view: the
    # counting walk's totals equal the per-ray counts of skw_closest_hit_stats over them
    batch = sk.Batch([scene], [(cam, w, h, spp, seed)])
    n = w * h * spp
    d_rays, d_hits = empty_dev(torch, n * 64), empty_dev(torch, n * 32)
    d_pops, d_tests = empty_dev(torch, n * 4), empty_dev(torch, n * 4)
    batch.rays(0, d_rays.data_ptr())
    scene.closest_hit_stats(d_rays.data_ptr(), n, d_hits.data_ptr(), d_pops.data_ptr(), d_tests.data_ptr())
    torch.cuda.synchronize()
    assert p2["node_tests"][0] == int(from_dev(d_pops, np.uint32, n).astype(np.int64).sum())
    assert p2["tri_tests"][0] == int(from_dev(d_tests, np.uint32, n).astype(np.int64).sum())


def test_passes_per_launch_do_not_change_the_frame(torch_cuda, monkeypatch):
    """A frame's launches cover P sample passes of every pixel (P a power of two dividing spp, the
    largest such not above SKW_DL_PASSES_PER_LAUNCH): each sample keeps its own values and the film
    adds a pixel's samples in pass order, so the image, the ray counts, the replayed image and the
    tape are those of one pass per launch, bit for bit; adjoints (whose FP64 atomics add in any
    order) agree to rounding. The default P renders as the oracle does."""
    print("case: direct lighting, 1 / 2 / 4 / 8 passes per launch, small room 48x40, 8 and 6 spp", flush=True)
    torch = torch_cuda
    s = wl.scene_room_small()
    w, h = 48, 40
    scene = sk.Scene(s["meshes"])
    gdl = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"],
                     prim_group=s["prim_group"], n_groups=s["n_groups"])
    cam, cam15 = sk.camera(*s["camera"], w, h)
    th0 = np.asarray(s["theta0"], np.float64)
    rng = np.random.default_rng(3)
    adj = torch.from_numpy(rng.normal(size=(w * h, 3))).cuda()
    runs = {}
    for spp in (8, 6):
        for P in (1, 2, 4, 8):
            monkeypatch.setenv("SKW_DL_PASSES_PER_LAUNCH", str(P))
            d_rgb = empty_dev(torch, w * h * 12)
            cnt = gdl.render(cam, w, h, spp, d_rgb.data_ptr(), 2, 3, 21)
            img = torch.empty((w * h, 3), dtype=torch.float64, device="cuda")
            rcnt = gdl.render_replay(th0, th0, cam, w, h, spp, img.data_ptr(), 2, 3, 21)
            timg = torch.empty_like(img)
            tape = gdl.render_tape(th0, cam, w, h, spp, timg.data_ptr(), 1 << 30, 2, 3, 21)
            g_tape = torch.empty(gdl.n_params, dtype=torch.float64, device="cuda")
            tape.adjoint(adj.data_ptr(), g_tape.data_ptr())
            g_adj = torch.empty_like(g_tape)
            gdl.adjoint(th0, cam, w, h, spp, adj.data_ptr(), g_adj.data_ptr(), 2, 3, 21)
            torch.cuda.synchronize()
            tape.close()
            runs[(spp, P)] = (from_dev(d_rgb, np.uint8, w * h * 12), cnt, rcnt, img.cpu().numpy(), timg.cpu().numpy(),
                              g_tape.cpu().numpy(), g_adj.cpu().numpy())
        base = runs[(spp, 1)]
        for P in (2, 4, 8):
            r = runs[(spp, P)]
            for k in range(5):
                assert np.array_equal(base[k], r[k]), (spp, P, k)
            for k in (5, 6):
                assert np.all(np.abs(base[k] - r[k]) <= 1e-12 * np.abs(base[k]).max()), (spp, P, k)
    monkeypatch.delenv("SKW_DL_PASSES_PER_LAUNCH")
    got, cnt, ref, rcnt = render_both(torch, s, w, h, 4, 2, 3, 21)
    assert np.array_equal(cnt, rcnt)
    assert_close_rel(got, ref)
