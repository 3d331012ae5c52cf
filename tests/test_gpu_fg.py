"""GPU parity of the photon final gather (skw_irradiance_map_create, skw_dl_final_gather, through
the C ABI) against the oracle's restatement of PhotonLTEIntegrator::_ConstructIrradiancePhotonMap
and _FinalGather on the same photons and sample stream.

Bar: irradiance photons (positions, normals, E_ext, E_int) and the lookup distance bit-exact (the
same k = 100 sets summed in the same order); gather-ray counts identical (same sampling decisions);
per-sample radiance within 1e-5 relative (the GPU's pow / sin / cos may differ from glibc's in the
last bits); the closed Lambertian box equal to kd * kd' * E / pi within 1e-12 in every pixel."""
import numpy as np
import pytest

import oracle
import skwarka_amd as sk
from skwarka_amd import workloads as wl
from tests.conftest import empty_dev, from_dev
from tests.test_gpu_dl import assert_close_rel

pytestmark = pytest.mark.gpu


def _maps(s, n_direct, n_indirect, n_caustic, seed):
    mesh = wl.merge(s["meshes"])
    sets = wl.photon_maps(mesh, n_direct, n_indirect, n_caustic, seed)
    gpu = [sk.PhotonMap(p["pos"], p["weight"], p["dir"], p["normal"]) if p is not None else None for p in sets]
    ref = [oracle.PhotonMap(p["pos"], p["weight"], p["dir"], p["normal"]) if p is not None else None for p in sets]
    return gpu, ref


@pytest.fixture(scope="module")
def room():
    s = wl.scene_room_small()
    gpu, ref = _maps(s, 20000, 60000, 8000, seed=21)
    area = wl.scene_area(s["meshes"])
    paths = (40000, 30000, 50000)  # caustic, direct, indirect
    irr = sk.IrradianceMap(direct=gpu[0], indirect=gpu[1], caustic=gpu[2], caustic_paths=paths[0],
                           direct_paths=paths[1], indirect_paths=paths[2], scene_area=area)
    oirr = oracle.IrradianceMap(direct=ref[0], indirect=ref[1], caustic=ref[2], paths=paths, scene_area=area)
    return s, gpu, ref, area, irr, oirr


def test_irradiance_map_matches_oracle(torch_cuda, room):
    print("case: irradiance photon map, small room, 20k/60k/8k photons", flush=True)
    s, gpu, ref, area, irr, oirr = room
    assert irr.n == oirr.n == 6000
    got, want = irr.export(), oirr.export()
    for g, w in zip(got, want):
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8))
    assert irr.max_lookup_dist == oirr.max_dist
    assert np.isfinite(irr.max_lookup_dist) and irr.max_lookup_dist > 0
    assert (got[2] > 0).any() and (got[3] >= 0).all()


def test_irradiance_map_without_caustics(torch_cuda, room):
    print("case: irradiance photon map, no caustic map", flush=True)
    s, gpu, ref, area, _, _ = room
    irr = sk.IrradianceMap(direct=gpu[0], indirect=gpu[1], direct_paths=7, indirect_paths=9, scene_area=area)
    oirr = oracle.IrradianceMap(direct=ref[0], indirect=ref[1], paths=(1, 7, 9), scene_area=area)
    for g, w in zip(irr.export(), oirr.export()):
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8))
    assert irr.max_lookup_dist == oirr.max_dist


def fg_both(torch, s, irr, oirr, ind_gpu, ind_ref, area, w, h, spp, gs, seed):
    cam, cam15 = sk.camera(*s["camera"], w, h)
    acc = oracle.Accel(s["meshes"])
    odl = oracle.DLScene(acc, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    ref, rrays = oracle.fg_render(acc, odl, oirr, ind_ref, area, cam15, w, h, spp, gs, seed)
    scene = sk.Scene(s["meshes"])
    gdl = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    d_rad = empty_dev(torch, w * h * spp * 24)
    d_rgb = empty_dev(torch, w * h * 12)
    rays = gdl.final_gather(irr, ind_gpu, area, cam, w, h, spp, gs, d_rgb.data_ptr(), d_rad.data_ptr(), seed)
    torch.cuda.synchronize()
    got = from_dev(d_rad, np.float64, w * h * spp * 3).reshape(w * h * spp, 3)
    rgb = from_dev(d_rgb, np.float32, w * h * 3).reshape(w * h, 3)
    return got, rgb, rays, ref, rrays


def test_final_gather_small_room_matches_oracle(torch_cuda, room):
    print("case: final gather, small room 48x32, 2 spp, 9 gather samples", flush=True)
    s, gpu, ref, area, irr, oirr = room
    w, h, spp = 48, 32, 2
    got, rgb, rays, want, rrays = fg_both(torch_cuda, s, irr, oirr, gpu[1], ref[1], area, w, h, spp, 8, seed=3)
    assert rays[0] == w * h * spp
    assert rays[1] == rrays and rrays > w * h * spp * 9
    assert want.max() > 0
    assert_close_rel(got, want)
    mean = want.reshape(w * h, spp, 3).mean(axis=1)
    assert_close_rel(rgb, mean.astype(np.float32), rel=2e-6)


def test_final_gather_odd_size_25_samples_matches_oracle(torch_cuda, room):
    # 37 x 23 pixels at 3 spp (2553 slots: the last block of 64 slots is partly empty) with 25
    # gather samples: each wave of the sample kernel takes 12 or 13 of the 50 BSDF and cone samples.
    print("case: final gather, small room 37x23, 3 spp, 25 gather samples", flush=True)
    s, gpu, ref, area, irr, oirr = room
    w, h, spp = 37, 23, 3
    got, rgb, rays, want, rrays = fg_both(torch_cuda, s, irr, oirr, gpu[1], ref[1], area, w, h, spp, 25, seed=11)
    assert rays[0] == w * h * spp
    assert rays[1] == rrays and rrays > w * h * spp * 25
    assert want.max() > 0
    assert_close_rel(got, want)
    mean = want.reshape(w * h, spp, 3).mean(axis=1)
    assert_close_rel(rgb, mean.astype(np.float32), rel=2e-6)


def test_final_gather_launch_shape_changes_no_value(torch_cuda, room, monkeypatch):
    # The same frame as pixel batches of 28 pixels with one pass per launch (2^10 gather rays) and
    # as every pixel with all 4 passes in one launch: per-sample values and the image bit for bit.
    print("case: final gather, small room 48x32, 4 spp, 2^10 gather rays per launch vs 4 passes per launch",
          flush=True)
    s, gpu, ref, area, irr, oirr = room
    w, h, spp = 48, 32, 4
    scene = sk.Scene(s["meshes"])
    gdl = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    cam, _ = sk.camera(*s["camera"], w, h)
    out = []
    for log2 in ("10", "27"):
        monkeypatch.setenv("SKW_FG_GATHER_RAYS_LOG2", log2)
        d_rad = empty_dev(torch_cuda, w * h * spp * 24)
        d_rgb = empty_dev(torch_cuda, w * h * 12)
        rays = gdl.final_gather(irr, gpu[1], area, cam, w, h, spp, 9, d_rgb.data_ptr(), d_rad.data_ptr(), seed=7)
        torch_cuda.cuda.synchronize()
        out.append((from_dev(d_rad, np.uint8, w * h * spp * 24), from_dev(d_rgb, np.uint8, w * h * 12), list(rays)))
    assert out[0][2] == out[1][2] and out[0][2][1] > w * h * spp * 9
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_final_gather_without_indirect_photons(torch_cuda, room):
    print("case: final gather, no indirect map (BSDF samples only), 16 gather samples", flush=True)
    s, gpu, ref, area, irr, oirr = room
    got, _, rays, want, rrays = fg_both(torch_cuda, s, irr, oirr, None, None, area, 32, 24, 1, 16, seed=5)
    assert rays[1] == rrays
    assert_close_rel(got, want)


def _box(inward):
    return [wl.box(1.0, inward=inward)], np.zeros(1, np.uint8), np.full((1, 3), 0.6), np.zeros((1, 3)), np.full(1, 0.2)


def test_final_gather_closed_box_over_pixel_batches(torch_cuda, monkeypatch):
    # At most 2^23 gather rays per launch: 676 gather samples hold 6204 pixels per batch, so an
    # 80 x 80 image takes two batches.
    monkeypatch.setenv("SKW_FG_GATHER_RAYS_LOG2", "23")
    print("case: final gather, closed box 80x80, 676 gather samples (two pixel batches)", flush=True)
    e_ext, e_int = np.array([1.0, 0.5, 0.25], np.float32), np.array([0.3, 0.2, 0.1], np.float32)
    for inward, e in ((True, e_ext), (False, e_int)):
        meshes, mat, kd, ks, rough = _box(inward)
        rng = np.random.default_rng(1)
        pos, nrm = [], []
        for axis in range(3):
            for sg in (-1.0, 1.0):
                p = rng.uniform(-0.99, 0.99, (300, 3))
                p[:, axis] = sg
                n = np.zeros((300, 3))
                n[:, axis] = -sg if inward else sg
                pos.append(p)
                nrm.append(n)
        pos = np.concatenate(pos).astype(np.float32)
        nid = wl.cdir_encode(np.concatenate(nrm))
        irr = sk.IrradianceMap(explicit=(pos, nid, np.tile(e_ext, (len(pos), 1)), np.tile(e_int, (len(pos), 1)),
                                         np.inf))
        scene = sk.Scene(meshes)
        gdl = sk.DLScene(scene, mat, kd, ks, rough, np.zeros(0, np.uint32), np.zeros((0, 3)))
        w = h = 80
        cam, _ = sk.camera((0.0, 0.2, 0.0), (0.1, -1.0, 0.05), (0.0, 0.0, 1.0), 60.0, w, h)
        d_rad = empty_dev(torch_cuda, w * h * 24)
        rays = gdl.final_gather(irr, None, wl.scene_area(meshes), cam, w, h, 1, 676, None, d_rad.data_ptr(), seed=2)
        torch_cuda.cuda.synchronize()
        got = from_dev(d_rad, np.float64, w * h * 3).reshape(w * h, 3)
        expect = 0.6 * 0.6 * e.astype(np.float64) / np.pi
        assert rays[0] == w * h and rays[1] == w * h * 676
        assert np.allclose(got, expect[None, :], rtol=1e-12, atol=0)


def test_final_gather_rejects_bad_parameters(torch_cuda, room):
    print("case: final gather argument checks", flush=True)
    s, gpu, ref, area, irr, oirr = room
    scene = sk.Scene(s["meshes"])
    gdl = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    cam, _ = sk.camera(*s["camera"], 8, 8)
    for gs in (0, 677):
        with pytest.raises(sk.SkwError):
            gdl.final_gather(irr, gpu[1], area, cam, 8, 8, 1, gs)
    with pytest.raises(sk.SkwError):
        gdl.final_gather(irr, gpu[1], -1.0, cam, 8, 8, 1, 4)
    with pytest.raises(sk.SkwError):
        gdl.final_gather(irr, gpu[1], area, cam, 0, 8, 1, 4)
