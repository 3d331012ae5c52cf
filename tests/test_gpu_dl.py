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
