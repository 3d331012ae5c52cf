"""Pins the photon-shooting oracle (orc_photon_shoot) to closed forms.

In a closed Lambertian box of reflectance rho lit by one gray quad light of area A and radiance Le
(PhotonLTEIntegrator::ShootPhotons, PhotonShootingPipeline.cpp:110-310):
* every path's first intersection deposits one direct photon of weight
  Le cos / (cos / pi / A) = pi A Le (the light's power: the pdfs of SamplePhoton cancel);
* Lambertian cosine sampling gives f cos / pdf = rho, so the roulette continues with probability
  rho and keeps the weight at pi A Le: the indirect photons per path are geometric with mean
  rho / (1 - rho);
* the share of direct photons landing on the floor is the light-to-floor form factor.
The lights' power CDF gives each light's photons the weight Phi_l Y(Phi) / Y(Phi_l) and a share
Y(Phi_l) / Y(Phi) of the paths. Rounds take photons into a map until it is full.
"""
import numpy as np

import oracle
from skwarka_amd import workloads as wl


def test_radical_inverse_known_values():
    cases = [(1, 2, 0.5), (2, 2, 0.25), (3, 2, 0.75), (4, 2, 0.125), (6, 2, 0.375), (1, 3, 1 / 3),
             (2, 3, 2 / 3), (3, 3, 1 / 9), (5, 3, 7 / 9), (7, 5, 0.44), (23, 23, 1 / 23 ** 2)]
    for n, b, v in cases:
        assert abs(oracle.radical_inverse(n, b) - v) < 1e-15
    # base 2: the first 2^k values are the 2^k strata
    v = sorted(oracle.radical_inverse(n, 2) for n in range(0, 1024))
    assert np.array_equal(v, np.arange(0, 1024) / 1024.0)


def _lit_box(rho=0.5, le=(2.0, 2.0, 2.0), lights=None):
    """[-1, 1]^3 with inward normals and quad lights of side 0.5 facing down just under the ceiling."""
    meshes = [wl.box(1.0)]
    lights = lights or [((0.0, 0.0), le)]
    y = 0.99
    for (x, z), _ in lights:
        meshes.append(wl.quad((x - 0.25, y, z - 0.25), (x + 0.25, y, z - 0.25), (x + 0.25, y, z + 0.25),
                              (x - 0.25, y, z + 0.25)))
    n = len(meshes)
    mat = np.zeros(n, np.uint8)
    kd = np.full((n, 3), rho)
    ks = np.zeros((n, 3))
    rough = np.full(n, 0.2)
    acc = oracle.Accel(meshes)
    dls = oracle.DLScene(acc, mat, kd, ks, rough, np.arange(1, n, dtype=np.uint32),
                         np.array([l[1] for l in lights], np.float64))
    return acc, dls


BIG = 10 ** 12


def test_direct_photons_carry_the_light_power():
    acc, dls = _lit_box(rho=0.5)
    n = 20000
    out = oracle.photon_shoot(acc, dls, n, (BIG, BIG, BIG), 4096, seed=1)
    d = out["maps"][1]
    assert len(d["pos"]) == n and out["paths"] == [n, n, n] and out["total_paths"] == n
    assert len(out["maps"][0]["pos"]) == 0  # no specular surfaces: no caustic photons
    assert np.allclose(d["weight"], np.pi * 0.25 * 2.0, rtol=1e-6, atol=0)
    # photons lie on the box or the light, arrive from the side their surface normal faces
    p = d["pos"].astype(np.float64)
    on_box = np.isclose(np.abs(p), 1.0, atol=1e-5).any(axis=1)
    assert np.all(on_box)  # the light faces down: no direct photon on it
    table = wl.cdir_table()
    cos = np.sum(table[d["dir"]] * table[d["normal"]], axis=1)
    assert np.all(cos > -0.02)


def test_indirect_photons_per_path_are_geometric():
    rho = 0.5
    acc, dls = _lit_box(rho=rho)
    n = 100000
    out = oracle.photon_shoot(acc, dls, n, (BIG, BIG, BIG), 16384, seed=2)
    ind = out["maps"][2]
    mean = len(ind["pos"]) / n
    assert abs(mean - rho / (1 - rho)) < 0.03
    assert np.allclose(ind["weight"], np.pi * 0.25 * 2.0, rtol=1e-6, atol=0)


def _form_factor_light_to_floor(h=1.99, half=0.25, m_l=20, m_f=200):
    xl = (np.arange(m_l) + 0.5) / m_l * 2 * half - half
    xf = (np.arange(m_f) + 0.5) / m_f * 2 - 1.0
    lx, lz = np.meshgrid(xl, xl)
    fx, fz = np.meshgrid(xf, xf)
    lx, lz, fx, fz = lx.ravel(), lz.ravel(), fx.ravel(), fz.ravel()
    total = 0.0
    for i in range(len(lx)):
        r2 = (fx - lx[i]) ** 2 + (fz - lz[i]) ** 2 + h * h
        total += np.sum(h * h / (np.pi * r2 * r2))
    return total * (4.0 / (m_f * m_f)) / (m_l * m_l)


def test_direct_photons_follow_the_form_factor():
    acc, dls = _lit_box(rho=0.5)
    n = 200000
    out = oracle.photon_shoot(acc, dls, n, (0, BIG, 0), 1 << 17, seed=3)
    p = out["maps"][1]["pos"]
    share = np.mean(np.abs(p[:, 1] + 1.0) < 1e-5)
    ff = _form_factor_light_to_floor()
    assert 0.2 < ff < 0.5
    assert abs(share - ff) / ff < 5e-3


def test_power_cdf_splits_paths_and_weights():
    le1, le2 = np.array([1.0, 2.0, 3.0]), np.array([4.0, 1.0, 0.5])
    acc, dls = _lit_box(lights=[((-0.5, 0.0), tuple(le1)), ((0.5, 0.0), tuple(le2))])
    n = 40000
    out = oracle.photon_shoot(acc, dls, n, (0, BIG, 0), 8192, seed=4)
    w = out["maps"][1]["weight"].astype(np.float64)
    lum = np.array([float.fromhex('0x1.b38293da589a6p-5'), float.fromhex('0x1.2b993cc221566p-1'),
                    float.fromhex('0x1.725d3400723f6p-2')])
    phi1, phi2 = np.pi * 0.25 * le1, np.pi * 0.25 * le2
    y1, y2 = lum @ phi1, lum @ phi2
    w1, w2 = phi1 * (y1 + y2) / y1, phi2 * (y1 + y2) / y2
    is1 = np.all(np.isclose(w, w1, rtol=1e-6), axis=1)
    is2 = np.all(np.isclose(w, w2, rtol=1e-6), axis=1)
    assert np.all(is1 ^ is2)
    assert abs(is1.mean() - y1 / (y1 + y2)) < 2e-3


def test_rounds_stop_full_maps():
    acc, dls = _lit_box(rho=0.6)
    out = oracle.photon_shoot(acc, dls, 100000, (0, 50000, BIG), 16384, seed=5)
    counts = [len(m["pos"]) for m in out["maps"]]
    # every path deposits one direct photon: full after the 4th round of 16384
    assert counts[0] == 0 and out["paths"][0] == 0
    assert counts[1] == 4 * 16384 and out["paths"][1] == 4 * 16384
    assert out["paths"][2] == 100000 and out["total_paths"] == 100000
    # the maps are in path order: the same call with one round of 100000 paths holds the same photons
    one = oracle.photon_shoot(acc, dls, 100000, (0, BIG, BIG), 100000, seed=5)
    assert np.array_equal(one["maps"][2]["pos"], out["maps"][2]["pos"])
    assert np.array_equal(one["maps"][1]["pos"][:counts[1]], out["maps"][1]["pos"])
