"""Pins the direct-lighting oracle to closed forms, as the reference's own tests do
(UnitTests/MainTests/Raytracer/Core/DirectLightingIntegrator.test.h:46-201: area light within 1e-2,
occluded light exactly 0)."""
import numpy as np

import oracle
import skwarka_amd as sk
from skwarka_amd import workloads as wl


def _floor_and_light(occluder=False, plastic=False):
    floor = wl.quad((-20, 0, -20), (-20, 0, 20), (20, 0, 20), (20, 0, -20))  # normal +y
    light = wl.quad((-0.5, 2, -0.5), (0.5, 2, -0.5), (0.5, 2, 0.5), (-0.5, 2, 0.5))  # normal -y
    meshes = [floor, light]
    if occluder:
        meshes.insert(1, wl.quad((-3, 1, -3), (-3, 1, 3), (3, 1, 3), (3, 1, -3)))
    n = len(meshes)
    mat = np.zeros(n, np.uint8)
    kd = np.zeros((n, 3))
    kd[0] = 0.5
    ks = np.zeros((n, 3))
    rough = np.full(n, 0.1)
    if plastic:
        mat[0] = 1
        ks[0] = 0.3
        rough[0] = 0.2
    return meshes, mat, kd, ks, rough, np.array([n - 1], np.uint32), np.array([[5.0, 4.0, 3.0]])


def _irradiance_at_origin(le):
    # E = Le * integral over the light of cos(theta) cos(theta') / r^2 dA, midpoint rule
    m = 400
    xs = (np.arange(m) + 0.5) / m - 0.5
    x, z = np.meshgrid(xs, xs)
    r2 = x * x + z * z + 4.0
    cos = 2.0 / np.sqrt(r2)
    return le * np.sum(cos * cos / r2) / (m * m)


def test_area_light_on_lambertian_floor():
    meshes, mat, kd, ks, rough, lp, lr = _floor_and_light()
    acc = oracle.Accel(meshes)
    dls = oracle.DLScene(acc, mat, kd, ks, rough, lp, lr)
    _, cam15 = sk.camera((0, 1.0, 0), (0, -1, 0), (0, 0, 1), 0.01, 1, 1)
    rgb, cnt = dls.render(cam15, 1, 1, 4096, seed=3)
    expect = 0.5 / np.pi * _irradiance_at_origin(lr[0])
    assert np.allclose(rgb[0], expect, rtol=1e-2)
    assert cnt[0] == 4096 and cnt[1] > 0 and cnt[2] > 0


def test_occluded_light_is_exactly_black():
    meshes, mat, kd, ks, rough, lp, lr = _floor_and_light(occluder=True)
    acc = oracle.Accel(meshes)
    dls = oracle.DLScene(acc, mat, kd, ks, rough, lp, lr)
    _, cam15 = sk.camera((0, 0.5, 0), (0, -1, 0), (0, 0, 1), 0.01, 1, 1)
    rgb, cnt = dls.render(cam15, 1, 1, 256, seed=4)
    assert np.all(rgb == 0.0)
    assert cnt[1] > 0  # shadow rays were cast and all were blocked


def test_looking_at_the_light_sees_its_radiance():
    meshes, mat, kd, ks, rough, lp, lr = _floor_and_light()
    kd[:] = 0.0
    acc = oracle.Accel(meshes)
    dls = oracle.DLScene(acc, mat, kd, ks, rough, lp, lr)
    _, cam15 = sk.camera((0, 1.0, 0), (0, 1, 0), (0, 0, 1), 0.01, 1, 1)
    rgb, _ = dls.render(cam15, 1, 1, 16, seed=5)
    assert np.allclose(rgb[0], lr[0].astype(np.float32), rtol=0, atol=0)


def test_plastic_floor_estimators_agree():
    # Light sampling and BSDF sampling estimate the same integral: the MIS-combined render with
    # 4+4 samples and the render with 1+1 samples converge to the same value.
    meshes, mat, kd, ks, rough, lp, lr = _floor_and_light(plastic=True)
    acc = oracle.Accel(meshes)
    dls = oracle.DLScene(acc, mat, kd, ks, rough, lp, lr)
    _, cam15 = sk.camera((0.3, 1.0, -0.8), (-0.2, -1, 0.7), (0, 0, 1), 0.01, 1, 1)
    a, _ = dls.render(cam15, 1, 1, 8192, 4, 4, seed=6)
    b, _ = dls.render(cam15, 1, 1, 8192, 1, 1, seed=7)
    assert np.allclose(a, b, rtol=3e-2)
