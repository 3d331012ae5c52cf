"""Pins the oracle's composed photon-mapping render (orc_pm_render, PhotonLTEIntegrator::
_SurfaceRadiance, PhotonLTEIntegrator.cpp:320-385) to its parts:

* without an irradiance map (no final gather) it is the direct-lighting render bit for bit, with
  the same ray counts (the caustic lookup of a photon-less caustic map adds a zero spectrum);
* with one, every pixel is the direct-lighting pixel plus the mean of the final gather's
  per-sample values (no sample of these scenes is rejected by the renderer's rule), so dropping
  or doubling either term fails; the gather-ray count is the final gather's own.
"""
import numpy as np

import oracle
import skwarka_amd as sk
from skwarka_amd import workloads as wl


def _room():
    """The small room with photon maps shot from its own lights (so the gather term has the
    direct term's scale) and the irradiance map built from them."""
    s = wl.scene_room_small()
    acc = oracle.Accel(s["meshes"])
    odl = oracle.DLScene(acc, s["mat"], s["kd"], s["ks"], s["rough"], s["light_prim"], s["light_radiance"])
    shot = oracle.photon_shoot(acc, odl, 60000, (10 ** 12,) * 3, 1 << 14, seed=3)
    ref = [oracle.PhotonMap(m["pos"], m["weight"], m["dir"], m["normal"]) if len(m["pos"]) else None
           for m in shot["maps"]]
    area = wl.scene_area(s["meshes"])
    oirr = oracle.IrradianceMap(direct=ref[1], indirect=ref[2], caustic=ref[0], paths=tuple(shot["paths"]),
                                scene_area=area)
    return s, acc, odl, ref, area, oirr


def test_pm_render_without_final_gather_is_direct_lighting():
    s, acc, odl, ref, area, oirr = _room()
    w, h, spp = 40, 30, 3
    _, cam15 = sk.camera(*s["camera"], w, h)
    rgb, cnt = oracle.pm_render(acc, odl, None, None, area, cam15, w, h, spp, 3, 2, 4, 17)
    dl, dcnt = odl.render(cam15, w, h, spp, 3, 2, 17)
    assert np.array_equal(rgb.view(np.uint32), dl.view(np.uint32))
    assert np.array_equal(cnt[:3], dcnt) and cnt[3] == 0
    assert dl.max() > 0


def test_pm_render_is_direct_lighting_plus_final_gather():
    s, acc, odl, ref, area, oirr = _room()
    w, h, spp, gs = 32, 24, 2, 4
    _, cam15 = sk.camera(*s["camera"], w, h)
    rgb, cnt = oracle.pm_render(acc, odl, oirr, ref[2], area, cam15, w, h, spp, 4, 4, gs, 5)
    dl, dcnt = odl.render(cam15, w, h, spp, 4, 4, 5)
    fg, grays = oracle.fg_render(acc, odl, oirr, ref[2], area, cam15, w, h, spp, gs, 5)
    expect = dl.astype(np.float64) + fg.reshape(w * h, spp, 3).mean(axis=1)
    assert np.allclose(rgb, expect, rtol=2e-6, atol=1e-12)
    assert np.array_equal(cnt[:3], dcnt) and cnt[3] == grays
    gmean = fg.reshape(w * h, spp, 3).mean(axis=1)
    assert (gmean > 1e-2 * np.maximum(dl, 1e-30)).mean() > 0.5  # the gather term is not negligible
