"""GPU parity of closest-hit / any-hit / hit geometry (through the C ABI) against the oracle.

Bar: hit ids, t, b1, b2 bit-exact (the plain form keeps the reference's FP64 arithmetic); hit
geometry (point, normals, cross, dot) within 1e-12 absolute / bit-exact where no sqrt is involved.
"""
import numpy as np
import pytest

import oracle
import skwarka_amd as sk
from skwarka_amd import workloads as wl
from tests import scenes
from tests.conftest import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu
MISS = 0xFFFFFFFF


def gpu_trace(torch, scene, rays, any_hit=False, geom=False):
    n = len(rays)
    d_rays = to_dev(torch, rays)
    if any_hit:
        d_occ = empty_dev(torch, n)
        scene.any_hit(d_rays.data_ptr(), n, d_occ.data_ptr())
        torch.cuda.synchronize()
        return from_dev(d_occ, np.uint8, n)
    d_hits = empty_dev(torch, n * 32)
    scene.closest_hit(d_rays.data_ptr(), n, d_hits.data_ptr())
    if geom:
        d_geom = empty_dev(torch, n * 104)
        scene.hit_geometry(d_rays.data_ptr(), d_hits.data_ptr(), n, d_geom.data_ptr())
        torch.cuda.synchronize()
        return from_dev(d_hits, sk.HIT_DTYPE, n), from_dev(d_geom, sk.GEOM_DTYPE, n)
    torch.cuda.synchronize()
    return from_dev(d_hits, sk.HIT_DTYPE, n)


def assert_hits_equal(got, ref):
    assert np.array_equal(got["prim"], ref["prim"])
    assert np.array_equal(got["tri"], ref["tri"])
    assert np.array_equal(got["t"].view(np.uint64), ref["t"].view(np.uint64))
    assert np.array_equal(got["b1"].view(np.uint64), ref["b1"].view(np.uint64))
    assert np.array_equal(got["b2"].view(np.uint64), ref["b2"].view(np.uint64))


def test_closest_hit_c0_full(torch_cuda):
    print("case: closest_hit C0 512x512", flush=True)
    meshes = wl.scene_c0()
    rays = wl.rays_c0()
    acc = oracle.Accel(meshes)
    ref = acc.closest_hit(rays)
    sc = sk.Scene(meshes)
    got = gpu_trace(torch_cuda, sc, rays)
    assert_hits_equal(got, ref)
    occ = gpu_trace(torch_cuda, sc, rays, any_hit=True)
    assert np.array_equal(occ, acc.any_hit(rays))


def test_visit_counts_match_oracle(torch_cuda):
    # Per-ray node-box tests and triangle tests equal the oracle's: the GPU walk visits the same
    # nodes in the same order with the same shrinking tmax (any divergence changes the counts).
    print("case: closest_hit_stats C0 256x256", flush=True)
    meshes = wl.scene_c0()
    rays = wl.rays_c0(256, 256)
    acc = oracle.Accel(meshes)
    ref, rpops, rtests = acc.closest_hit(rays, counters=True)
    sc = sk.Scene(meshes)
    n = len(rays)
    d_rays, d_hits = to_dev(torch_cuda, rays), empty_dev(torch_cuda, n * 32)
    d_pops, d_tests = empty_dev(torch_cuda, n * 4), empty_dev(torch_cuda, n * 4)
    sc.closest_hit_stats(d_rays.data_ptr(), n, d_hits.data_ptr(), d_pops.data_ptr(), d_tests.data_ptr())
    torch_cuda.cuda.synchronize()
    assert_hits_equal(from_dev(d_hits, sk.HIT_DTYPE, n), ref)
    assert np.array_equal(from_dev(d_pops, np.uint32, n), rpops)
    assert np.array_equal(from_dev(d_tests, np.uint32, n), rtests)


def test_reference_fixture_vs_oracle_and_brute_force(torch_cuda):
    print("case: reference spheres 1000 random rays", flush=True)
    meshes = scenes.reference_spheres()
    acc = oracle.Accel(meshes)
    sc = sk.Scene(meshes)
    rays = wl.reference_test_rays(acc.bounds(), 1000)
    got = gpu_trace(torch_cuda, sc, rays)
    assert_hits_equal(got, acc.closest_hit(rays))
    bh = acc.brute_closest(rays)
    assert np.array_equal(got["prim"], bh["prim"]) and np.array_equal(got["tri"], bh["tri"])
    occ = gpu_trace(torch_cuda, sc, rays, any_hit=True)
    assert np.array_equal(occ.astype(bool), bh["prim"] != MISS)


def test_random_rays_many_spheres(torch_cuda):
    print("case: random rays in icospheres+box, tmin/tmax windows", flush=True)
    meshes = scenes.small_spheres(3, n=12, subdiv=4)
    acc = oracle.Accel(meshes)
    sc = sk.Scene(meshes)
    rays = scenes.random_rays(200_000, -2.9, 2.9, seed=4)
    rng = np.random.default_rng(5)
    rays["tmin"][::3] = rng.uniform(-1, 1, len(rays[::3]))
    rays["tmax"][1::3] = rng.uniform(0.1, 3, len(rays[1::3]))
    assert_hits_equal(gpu_trace(torch_cuda, sc, rays), acc.closest_hit(rays))
    assert np.array_equal(gpu_trace(torch_cuda, sc, rays, any_hit=True), acc.any_hit(rays))


def test_axis_aligned_and_degenerate_rays(torch_cuda):
    # Zero direction components give infinite inverses and NaN slab terms (TriangleAccelerator.cpp:
    # 176-192), -0.0 takes the "else" child order (:199), NaN rays must miss without faulting.
    print("case: axis-aligned / signed-zero / NaN rays", flush=True)
    meshes = scenes.small_spheres(6, n=5, subdiv=3)
    acc = oracle.Accel(meshes)
    sc = sk.Scene(meshes)
    rng = np.random.default_rng(6)
    o = rng.uniform(-2.5, 2.5, (6000, 3))
    d = np.zeros((6000, 3))
    axis = rng.integers(0, 3, 6000)
    d[np.arange(6000), axis] = rng.choice([-1.0, 1.0], 6000)
    d[::7] = -0.0
    d[::7, 0] = 1.0
    # origins exactly on node planes: reuse vertex coordinates
    v = np.concatenate([m["verts"] for m in meshes]).astype(np.float64)
    o[::5] = v[rng.integers(0, len(v), len(o[::5]))]
    rays = wl.make_rays(o, d)
    nan_rays = wl.make_rays(np.full((64, 3), np.nan), d[:64])
    rays = np.concatenate([rays, nan_rays])
    assert_hits_equal(gpu_trace(torch_cuda, sc, rays), acc.closest_hit(rays))
    assert np.array_equal(gpu_trace(torch_cuda, sc, rays, any_hit=True), acc.any_hit(rays))


def test_transformed_and_instanced_primitives(torch_cuda):
    print("case: transformed prims and instanced tetrahedra", flush=True)
    inst_meshes, inst_prims, flat = scenes.instanced_tetrahedra(1000, seed=9)
    rng = np.random.default_rng(10)
    o = rng.uniform(0, 100, (20000, 3))
    d = rng.uniform(0, 1, (20000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = wl.make_rays(o, d)
    for meshes, prims in ((inst_meshes, inst_prims),
                          ([scenes.tetra_mesh() for _ in inst_prims], [(i, m) for i, (_, m) in enumerate(inst_prims)])):
        acc = oracle.Accel(meshes, prims)
        sc = sk.Scene(meshes, prims)
        got, g = gpu_trace(torch_cuda, sc, rays, geom=True)
        ref, rg = acc.closest_hit(rays, geom=True)
        assert_hits_equal(got, ref)
        hit = ref["prim"] != MISS
        assert hit.sum() > 100
        for f in ("p", "ng", "ns", "cross"):
            assert np.allclose(g[f][hit], rg[f][hit], rtol=0, atol=1e-9), f
        assert np.array_equal(gpu_trace(torch_cuda, sc, rays, any_hit=True), acc.any_hit(rays))


def test_hit_geometry_matches_oracle(torch_cuda):
    print("case: hit geometry on C0", flush=True)
    meshes = wl.scene_c0()
    rays = wl.rays_c0(128, 128)
    acc = oracle.Accel(meshes)
    sc = sk.Scene(meshes)
    got, g = gpu_trace(torch_cuda, sc, rays, geom=True)
    ref, rg = acc.closest_hit(rays, geom=True)
    assert_hits_equal(got, ref)
    hit = ref["prim"] != MISS
    # point, cross and dot use only + - * /: bit-exact; normals go through sqrt: 1e-15 relative
    assert np.array_equal(g["p"][hit].view(np.uint64), rg["p"][hit].view(np.uint64))
    assert np.array_equal(g["cross"][hit].view(np.uint64), rg["cross"][hit].view(np.uint64))
    assert np.array_equal(g["dot"][hit].view(np.uint64), rg["dot"][hit].view(np.uint64))
    assert np.allclose(g["ng"][hit], rg["ng"][hit], rtol=0, atol=1e-15)
    assert np.allclose(g["ns"][hit], rg["ns"][hit], rtol=0, atol=1e-15)
    assert np.all(g["p"][~hit] == 0)


def test_empty_scene_and_empty_batch(torch_cuda):
    print("case: empty scene / zero rays", flush=True)
    sc = sk.Scene([{"verts": np.zeros((0, 3), np.float32), "tris": np.zeros((0, 3), np.uint32)}])
    rays = scenes.random_rays(100, -1, 1, seed=1)
    got = gpu_trace(torch_cuda, sc, rays)
    assert np.all(got["prim"] == MISS) and np.all(np.isinf(got["t"]))
    assert not gpu_trace(torch_cuda, sc, rays, any_hit=True).any()
    sc2 = sk.Scene(wl.scene_c0())
    sc2.closest_hit(0, 0, 0)  # n = 0 is a no-op even with NULL pointers
    with pytest.raises(sk.SkwError):
        sc2.closest_hit(0, 5, 0)


def test_deepest_tree_uses_global_stack(torch_cuda):
    """A tree at the builder's depth limit (200 levels): closest and any hit through the
    global-memory stack variant, bit-exact against the oracle and brute force."""
    print("case: 200-level tree, global-stack traversal, 20000 rays", flush=True)
    meshes = scenes.staircase()
    scene = sk.Scene(meshes)
    info = scene.info()
    assert info["max_depth"] == 200 and info["stack_need"] > 128
    rays = scenes.staircase_rays(20000, 3)
    acc = oracle.Accel(meshes)
    ref = acc.closest_hit(rays)
    assert_hits_equal(gpu_trace(torch_cuda, scene, rays), ref)
    assert np.array_equal(gpu_trace(torch_cuda, scene, rays, any_hit=True), acc.any_hit(rays))
    brute = acc.brute_closest(rays)
    assert np.array_equal(ref["prim"], brute["prim"]) and np.array_equal(ref["tri"], brute["tri"])
    assert (ref["prim"] != MISS).mean() > 0.5


def test_five_wave_instance_matches_oracle(torch_cuda, monkeypatch):
    # SKW_TRACE_WAVES_PER_CU above 16 runs every launch whose stack fits 32 entries on the walk
    # compiled for five waves per SIMD (the instance large launches take by default): the same
    # closest hits and occlusion, bit for bit, on the cases above.
    print("case: five-wave trace instance (C0, random windows, degenerate, instanced rays)", flush=True)
    monkeypatch.setenv("SKW_TRACE_WAVES_PER_CU", "19")
    test_closest_hit_c0_full(torch_cuda)
    test_random_rays_many_spheres(torch_cuda)
    test_axis_aligned_and_degenerate_rays(torch_cuda)
    test_transformed_and_instanced_primitives(torch_cuda)
