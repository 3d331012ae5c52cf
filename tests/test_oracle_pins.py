"""Pins of the oracle's BSDF, Fresnel and light-sampling pieces: the values the reference's own unit
tests assert, and closed forms a wrong constant, sign or operand would fail.

* BlinnDistribution: D(e=0) = 1/pi (BlinnDistribution.test.h:28-34); the normalisation
  integral of D(h) cos(theta_h) over the hemisphere is 1 for every exponent (closed form); the
  pdf of Microfacet<., Blinn(20)> summed over the upper hemisphere at incident (0.5,0.5,0.5)/|.| with
  400x400 stratified uniform-hemisphere samples is 0.9083 +- 1e-3 (Microfacet.test.h:107-125).
* Microfacet reciprocity (Microfacet.test.h:53-61); sampled pdf equals the BSDF pdf
  (Microfacet.test.h:81-100, BlinnDistribution.test.h:78-98).
* FresnelDielectric(1.5, 1): range [0, 1] and total internal reflection (Fresnel.test.h:28-52);
  0.04 at normal incidence; the textbook angle form sin^2(i - t) / sin^2(i + t), tan^2 likewise, and
  Brewster's angle (p-polarised reflectance 0).
* Lambertian: f = R/pi, pdf = |cos|/pi (Lambertian.cpp:25-36, BxDF.cpp:20-41).
* DiffuseAreaLightSource on the reference tests' tetrahedron: Power = pi * area with area = 8/sqrt(3)
  (DiffuseAreaLightSource.test.h:44-50); sampled points on the surface, normalised directions,
  radiance only from the lit side, LightingPDF equal to the sampled pdf (:52-130); LightingPDF
  integrates to 1 +- 0.02 over the sphere of directions (:131-158).
* IrradianceLightsSamplingStrategy on five tetrahedra at x = 10 i: CDF in [0, 1], non-decreasing,
  ending at 1, area-light probabilities decreasing with distance, and every light at zero weight
  (opposite hemisphere) giving the equal-probability CDF of _Weights_To_CDF
  (IrradianceLightsSamplingStrategy.test.h:58-130, IrradianceLightsSamplingStrategy.cpp:110-129).
"""
import numpy as np
import pytest

import oracle
from tests import scenes

INV_PI = 1.0 / np.pi


def stratified(n, rng):
    """StratifiedSampling2D with jitter: n x n cells, one uniform point per cell."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    u = (i.reshape(-1) + rng.uniform(0, 1, n * n)) / n
    v = (j.reshape(-1) + rng.uniform(0, 1, n * n)) / n
    return np.minimum(u, np.nextafter(1.0, 0)), np.minimum(v, np.nextafter(1.0, 0))


def uniform_hemisphere(u, v):
    z = u
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * v
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


def uniform_sphere(u, v):
    z = 1.0 - 2.0 * u
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * v
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


# ------------------------------------------------------------------ Blinn / microfacet


def test_blinn_zero_exponent_is_inv_pi():
    assert oracle.blinn_d(0.0, unit([0.5, -0.5, 0.5])) == INV_PI


@pytest.mark.parametrize("e", [0.0, 1.0, 3.3, 10.0, 100.0, 1000.0])
def test_blinn_normalisation_integral(e):
    # integral over the hemisphere of D(h) cos(theta_h) d omega = 2 pi int_0^1 D(z) z dz = 1
    x, w = np.polynomial.legendre.leggauss(400)
    z = 0.5 * (x + 1.0)
    d = np.array([oracle.blinn_d(e, [np.sqrt(max(0.0, 1 - zz * zz)), 0.0, zz]) for zz in z])
    assert abs(2.0 * np.pi * 0.5 * np.sum(w * d * z) - 1.0) < 1e-9


def test_microfacet_pdf_sum_reference_value():
    rng = np.random.default_rng(7)
    u, v = stratified(400, rng)
    ex = uniform_hemisphere(u, v)
    inc = unit([0.5, 0.5, 0.5])
    total = sum(oracle.microfacet_pdf(20.0, inc, e) for e in ex) * 2.0 * np.pi / len(ex)
    assert abs(total - 0.9083) < 1e-3, total


def test_microfacet_reciprocity():
    v1, v2 = unit([0.2, 0.5, 0.8]), unit([-0.1, -0.3, 0.9])
    for e in (20.0, 3.0, 500.0):
        a, b = oracle.microfacet_eval(e, v1, v2), oracle.microfacet_eval(e, v2, v1)
        assert abs(a[0] - b[0]) <= 1e-14 * a[0] and abs(a[1] - b[1]) <= 1e-15
    kd, ks = [0.3, 0.5, 0.7], [0.04, 0.04, 0.04]
    f1, _ = oracle.bsdf_local(1, kd, ks, 20.0, v1, v2)
    f2, _ = oracle.bsdf_local(1, kd, ks, 20.0, v2, v1)
    assert np.allclose(f1, f2, rtol=1e-14, atol=0)


@pytest.mark.parametrize("plastic,e,inc", [(1, 10.0, [0.5, -0.5, -0.5]), (1, 20.0, [0.5, 0.5, 0.5]),
                                           (1, 300.0, [0.1, 0.2, 0.9]), (0, 1.0, [0.5, 0.5, 0.5])])
def test_sampled_pdf_equals_bsdf_pdf(plastic, e, inc):
    rng = np.random.default_rng(11)
    kd, ks = [0.4, 0.5, 0.6], [0.04, 0.04, 0.04]
    wo = unit(inc)
    checked = 0
    for _ in range(1000):
        u0, u1, uc = rng.uniform(0, 1, 3)
        wi, pdf, f = oracle.bsdf_local_sample(plastic, kd, ks, e, wo, u0, u1, uc)
        assert pdf >= 0.0 and np.all(f >= 0.0)
        assert abs(np.linalg.norm(wi) - 1.0) < 1e-12
        if pdf > 0.0:
            f2, pdf2 = oracle.bsdf_local(plastic, kd, ks, e, wo, wi)
            assert abs(pdf - pdf2) <= 1e-5 * pdf
            assert np.allclose(f, f2, rtol=1e-12, atol=0)
            checked += 1
    assert checked > 500


def test_lambertian_value_and_pdf():
    kd = np.array([0.2, 0.5, 0.9])
    wo, wi = unit([0.3, -0.2, 0.9]), unit([-0.4, 0.1, 0.6])
    f, pdf = oracle.bsdf_local(0, kd, [0, 0, 0], 1.0, wo, wi)
    assert np.array_equal(f, kd * INV_PI) and abs(pdf - wi[2] * INV_PI) < 1e-16
    f, pdf = oracle.bsdf_local(0, kd, [0, 0, 0], 1.0, wo, wi * [1, 1, -1])
    assert np.all(f == 0.0) and pdf == 0.0  # transmission side: no reflection, no pdf
    # plastic = Lambertian + microfacet lobe, pdf averaged over the two components
    f, pdf = oracle.bsdf_local(1, kd, [0.04] * 3, 20.0, wo, wi)
    sc, fr = oracle.microfacet_eval(20.0, wo, wi)
    assert np.allclose(f, kd * INV_PI + 0.04 * fr * sc, rtol=1e-15, atol=0)
    assert abs(pdf - 0.5 * (wi[2] * INV_PI + oracle.microfacet_pdf(20.0, wo, wi))) < 1e-15


# ------------------------------------------------------------------ Fresnel


def test_fresnel_range_and_total_internal_reflection():
    th = np.arange(0.0, np.pi + 1e-12, 0.01)
    r = np.array([oracle.fresnel_dielectric(np.cos(t)) for t in th])
    assert np.all((r >= 0.0) & (r <= 1.0))
    assert oracle.fresnel_dielectric(np.cos(np.pi / 2 + 0.1)) == 1.0


def test_fresnel_closed_forms():
    assert abs(oracle.fresnel_dielectric(1.0) - 0.04) < 1e-15
    assert abs(oracle.fresnel_dielectric(-1.0) - 0.04) < 1e-15
    n1, n2 = 1.0, 1.5  # light arriving from outside (cos > 0) enters the 1.5 medium
    for ti in np.linspace(0.05, 1.5, 12):
        tt = np.arcsin(n1 / n2 * np.sin(ti))
        rs = np.sin(ti - tt) ** 2 / np.sin(ti + tt) ** 2
        rp = np.tan(ti - tt) ** 2 / np.tan(ti + tt) ** 2
        assert abs(oracle.fresnel_dielectric(np.cos(ti)) - 0.5 * (rs + rp)) < 1e-12
        # from inside (cos < 0): n1 and n2 swap; below the critical angle the same formula holds
        if np.sin(ti) * n2 / n1 < 1.0:
            tt2 = np.arcsin(n2 / n1 * np.sin(ti))
            rs2 = np.sin(ti - tt2) ** 2 / np.sin(ti + tt2) ** 2
            rp2 = np.tan(ti - tt2) ** 2 / np.tan(ti + tt2) ** 2
            assert abs(oracle.fresnel_dielectric(-np.cos(ti)) - 0.5 * (rs2 + rp2)) < 1e-12
        else:
            assert oracle.fresnel_dielectric(-np.cos(ti)) == 1.0
    tb = np.arctan(n2 / n1)  # Brewster: p-polarised reflectance vanishes
    tt = np.pi / 2 - tb
    assert abs(oracle.fresnel_dielectric(np.cos(tb)) - 0.5 * np.sin(tb - tt) ** 2) < 1e-12


# ------------------------------------------------------------------ area lights


def tetra_light_scene(origins):
    meshes = []
    for o in origins:
        m = scenes.tetra_mesh()
        meshes.append({"verts": (m["verts"] + np.asarray(o, np.float32)).astype(np.float32), "tris": m["tris"]})
    acc = oracle.Accel(meshes)
    n = len(meshes)
    dls = oracle.DLScene(acc, np.zeros(n, np.uint8), np.full((n, 3), 0.5), np.zeros((n, 3)), np.full(n, 0.1),
                         np.arange(n, dtype=np.uint32), np.ones((n, 3)))
    return acc, dls, meshes


def on_triangle(p, v):
    a, b, c = v
    area = 0.5 * np.linalg.norm(np.cross(a - b, a - c))
    s = (0.5 * np.linalg.norm(np.cross(p - a, p - b)) + 0.5 * np.linalg.norm(np.cross(p - b, p - c)) +
         0.5 * np.linalg.norm(np.cross(p - c, p - a)))
    return abs(s - area) < 1e-9


def test_area_light_power_and_sampling():
    acc, dls, meshes = tetra_light_scene([(0, 0, 0)])
    area, _ = dls.light_info(0)
    assert abs(np.pi * area - np.pi * 4 * 2.0 / np.sqrt(3.0)) < 1e-5
    v = meshes[0]["verts"].astype(np.float64)
    tris = meshes[0]["tris"]
    e1 = v[tris[:, 1]] - v[tris[:, 0]]
    e2 = v[tris[:, 2]] - v[tris[:, 0]]
    normals = np.cross(e1, e2)
    normals /= np.linalg.norm(normals, axis=1)[:, None]
    rng = np.random.default_rng(3)
    p = np.array([3.0, 4.0, 5.0])
    for _ in range(1000):
        tu, u0, u1 = rng.uniform(0, 1, 3)
        t, wi, dist, pdf, lit = dls.light_sample(0, p, tu, u0, u1)
        assert abs(np.linalg.norm(wi) - 1.0) < 1e-12 and pdf >= 0.0
        q = p + wi * dist
        assert on_triangle(q, v[tris[t]])
        assert lit == (normals[t] @ wi < 0.0)
        assert abs(dls.light_pdf(0, t, wi, dist) - pdf) <= 1e-10


def test_area_light_pdf_integrates_to_one():
    acc, dls, meshes = tetra_light_scene([(0, 0, 0)])
    v = meshes[0]["verts"].astype(np.float64)
    tris = meshes[0]["tris"]
    rng = np.random.default_rng(5)
    u, w = stratified(300, rng)
    dirs = uniform_sphere(u, w)
    p = np.array([1.5, 2.0, 1.2])
    total = 0.0
    for t, (a, b, c) in enumerate(v[tris]):
        # Moller-Trumbore of every direction against triangle t (the reference test's helper)
        e1, e2 = b - a, c - a
        s1 = np.cross(dirs, e2)
        div = s1 @ e1
        ok = div != 0.0
        inv = np.where(ok, 1.0 / np.where(ok, div, 1.0), 0.0)
        d = p - a
        b1 = (s1 @ d) * inv
        s2 = np.cross(d, e1)
        b2 = (dirs @ s2) * inv
        tt = (s2 @ e2) * inv
        hit = ok & (b1 >= 0) & (b2 >= 0) & (b1 + b2 <= 1) & (tt > 0)
        for k in np.nonzero(hit)[0]:
            q = a + b1[k] * e1 + b2[k] * e2
            dist = np.linalg.norm(q - p)
            total += dls.light_pdf(0, t, (q - p) / dist, dist) * 4.0 * np.pi / len(dirs)
    assert abs(total - 1.0) < 0.02, total


def test_lights_cdf_reference_cases():
    acc, dls, _ = tetra_light_scene([(10.0 * i, 0, 0) for i in range(5)])
    for n in ([0.0, 0.0, 1.0], [1.0, 0.0, 0.0]):
        cdf = dls.lights_cdf([0, 0, 0], n, 5)
        assert np.all((cdf >= 0) & (cdf <= 1)) and np.all(np.diff(cdf) >= 0)
        assert abs(cdf[-1] - 1.0) <= 1e-15
        pdf = np.diff(np.r_[0.0, cdf])
        assert np.all(np.diff(pdf[1:]) < 0)  # farther lights get lower probability
    # the origin lies inside light 0's box, which keeps its weight pi * intensity whatever the
    # normal, while the lights behind the normal get 0 (the reference's WithNormal2 case)
    cdf = dls.lights_cdf([0, 0, 0], [-1.0, 0.0, 0.0], 5)
    assert abs(cdf[0] - 1.0) <= 1e-15 and np.all(cdf[1:] == cdf[0])  # w * (1 / w) rounds to 1 +- 1 ulp
    # outside every box with every light behind the normal: all weights 0 -> equal probabilities
    cdf = dls.lights_cdf([-5.0, 0, 0], [-1.0, 0.0, 0.0], 5)
    assert np.array_equal(cdf, (np.arange(5) + 1.0) * (1.0 / 5.0))  # (i + 1) * inv, as _Weights_To_CDF
