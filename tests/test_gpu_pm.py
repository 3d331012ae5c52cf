"""GPU parity of the composed photon-mapping render (skw_pm_render, through the C ABI) against the
oracle's orc_pm_render, over photon maps shot on the device (skw_photon_shoot), built into kd-trees
on the device (skw_photon_set_map) and turned into the irradiance map (skw_irradiance_map_create);
the oracle builds its maps from the same photons.

Bar: per-pixel radiance within 1e-5 relative (as direct lighting and the final gather alone);
camera, shadow, BSDF and gather ray counts identical; without an irradiance map the composition is
skw_dl_render's image bit for bit; a caustic map holding photons is refused."""
import numpy as np
import pytest

import oracle
import skwarka_amd as sk
from skwarka_amd import workloads as wl
from tests.conftest import empty_dev, from_dev
from tests.test_gpu_dl import assert_close_rel

pytestmark = pytest.mark.gpu


def shot_maps(s, scene, gdl, n_paths, seed):
    """Photons shot on the device, their device-built maps and irradiance map, and the oracle's
    maps and irradiance map over the same photons."""
    ps = gdl.shoot_photons(n_paths, (40_000_000,) * 3, 1 << 16, seed)
    area = wl.scene_area(s["meshes"])
    gmaps = [ps.photon_map(j) if ps.counts[j] > 0 else None for j in range(3)]
    paths = [int(x) for x in ps.paths]
    irr = sk.IrradianceMap(direct=gmaps[1], indirect=gmaps[2], caustic=gmaps[0], caustic_paths=paths[0],
                           direct_paths=paths[1], indirect_paths=paths[2], scene_area=area)
    omaps = []
    for j in range(3):
        m = ps.export(j)
        omaps.append(oracle.PhotonMap(m["pos"], m["weight"], m["dir"], m["normal"]) if ps.counts[j] > 0 else None)
    oirr = oracle.IrradianceMap(direct=omaps[1], indirect=omaps[2], caustic=omaps[0], paths=tuple(paths),
                                scene_area=area)
    return ps, gmaps, irr, omaps, oirr, area


def test_pm_render_small_room_matches_oracle(torch_cuda):
    print("case: photon-mapping render, small room 64x48 x 3 spp, shot maps", flush=True)
    torch = torch_cuda
    s = wl.scene_room_small()
    scene = sk.Scene(s["meshes"])
    gdl = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    ps, gmaps, irr, omaps, oirr, area = shot_maps(s, scene, gdl, 200_000, 11)
    assert ps.counts[0] == 0 and ps.counts[1] > 0 and ps.counts[2] > 0
    for g, w in zip(irr.export(), oirr.export()):
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8))
    w, h, spp, gs = 64, 48, 3, 8
    cam, cam15 = sk.camera(*s["camera"], w, h)
    d_rgb = empty_dev(torch, w * h * 12)
    rays = gdl.pm_render(irr, gmaps[0], gmaps[2], area, cam, w, h, spp, d_rgb.data_ptr(), 4, 4, gs, 29)
    torch.cuda.synchronize()
    got = from_dev(d_rgb, np.float32, w * h * 3).reshape(w * h, 3)
    acc = oracle.Accel(s["meshes"])
    odl = oracle.DLScene(acc, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    ref, rcnt = oracle.pm_render(acc, odl, oirr, omaps[2], area, cam15, w, h, spp, 4, 4, gs, 29)
    assert np.array_equal(np.asarray(rays, np.int64), rcnt.astype(np.int64)), (rays, rcnt)
    assert rcnt[3] > 10000
    assert_close_rel(got, ref)
    # the gather term matters: the composition differs from direct lighting alone
    d_dl = empty_dev(torch, w * h * 12)
    gdl.render(cam, w, h, spp, d_dl.data_ptr(), 4, 4, 29)
    torch.cuda.synchronize()
    dl = from_dev(d_dl, np.float32, w * h * 3).reshape(w * h, 3)
    assert (got > dl * (1 + 1e-3)).mean() > 0.5
    # without an irradiance map: skw_dl_render's image bit for bit
    rays0 = gdl.pm_render(None, None, None, area, cam, w, h, spp, d_rgb.data_ptr(), 4, 4, gs, 29)
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(d_rgb, np.uint32, w * h * 3), dl.view(np.uint32).reshape(-1))
    assert rays0[3] == 0
    # a caustic map with photons cannot arise from these materials and is refused
    with pytest.raises(sk.SkwError, match="caustic"):
        gdl.pm_render(irr, gmaps[2], gmaps[2], area, cam, w, h, spp, d_rgb.data_ptr(), 4, 4, gs, 29)


def test_pm_render_c2_room_band_matches_oracle(torch_cuda):
    """The 2,026,260-triangle room at 1024x1024, 16 spp, 4 + 4 samples per hit, 8 (9) gather samples,
    over maps of 2^20 shot paths: rows 496-527 against the oracle."""
    print("case: photon-mapping render, C2 room 1024x1024 x 16 spp, shot maps, rows 496-527", flush=True)
    torch = torch_cuda
    s = wl.scene_room()
    scene = sk.Scene(s["meshes"])
    gdl = sk.DLScene(scene, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    ps, gmaps, irr, omaps, oirr, area = shot_maps(s, scene, gdl, 1 << 20, 77)
    for g, w in zip(irr.export(), oirr.export()):
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8))
    w = h = 1024
    spp, gs, seed = 16, 8, 1234
    cam, cam15 = sk.camera(*s["camera"], w, h)
    d_rgb = empty_dev(torch, w * h * 12)
    rays = gdl.pm_render(irr, gmaps[0], gmaps[2], area, cam, w, h, spp, d_rgb.data_ptr(), 4, 4, gs, seed)
    torch.cuda.synchronize()
    got = from_dev(d_rgb, np.float32, w * h * 3).reshape(w * h, 3)
    acc = oracle.Accel(s["meshes"])
    odl = oracle.DLScene(acc, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    p0, p1 = 496 * w, 528 * w
    ref, rcnt = oracle.pm_render(acc, odl, oirr, omaps[2], area, cam15, w, h, spp, 4, 4, gs, seed, p0, p1)
    assert_close_rel(got[p0:p1], ref)
    assert rays[0] == w * h * spp and rays[3] > 10 ** 8
    assert ref.max() > 0
