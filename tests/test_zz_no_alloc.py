"""skw_closest_hit_noalloc (through the C ABI) against skw_closest_hit: bit-identical hit records on
two calls in a row, the second with fewer rays, for a scene whose walks use the LDS stacks and for
a tree of the builder's maximum depth (200 levels), whose walks use the fixed global stack area with
a grid limited to it. The entry point shares one device array per GPU between calls, so the second
call also checks that nothing of the first call's state leaks into it."""
import numpy as np
import pytest

import oracle
import skwarka_amd as sk
from skwarka_amd import workloads as wl
from tests.conftest import empty_dev, from_dev, to_dev
from tests.scenes import staircase, staircase_rays

pytestmark = pytest.mark.gpu


def run_pair(torch, scene, rays_a, rays_b):
    for rays in (rays_a, rays_b):
        n = len(rays)
        d_rays = to_dev(torch, rays)
        ref = empty_dev(torch, n * 32)
        got = empty_dev(torch, n * 32)
        scene.closest_hit(d_rays.data_ptr(), n, ref.data_ptr())
        scene.closest_hit_noalloc(d_rays.data_ptr(), n, got.data_ptr())
        torch.cuda.synchronize()
        r = from_dev(ref, sk.HIT_DTYPE, n)
        g = from_dev(got, sk.HIT_DTYPE, n)
        assert np.array_equal(g.view(np.uint8), r.view(np.uint8))
        yield r


def test_noalloc_matches_closest_hit_c0(torch_cuda):
    print("case: noalloc closest hit, c0 scene, 512x512 then 64x64 rays", flush=True)
    meshes = wl.scene_c0()
    scene = sk.Scene(meshes)
    rays_a = wl.rays_c0()
    rays_b = wl.rays_c0(64, 64)
    outs = list(run_pair(torch_cuda, scene, rays_a, rays_b))
    assert (outs[0]["prim"] != sk.MISS).mean() > 0.99


def test_noalloc_matches_closest_hit_deepest_tree(torch_cuda):
    print("case: noalloc closest hit, 200-level tree (global stack), 50000 then 3000 rays", flush=True)
    meshes = staircase()
    scene = sk.Scene(meshes)
    info = scene.info()
    assert info["max_depth"] == 200 and info["stack_need"] > 128
    rays_a, rays_b = staircase_rays(50000, 1), staircase_rays(3000, 2)
    outs = list(run_pair(torch_cuda, scene, rays_a, rays_b))
    acc = oracle.Accel(meshes)
    ref = acc.closest_hit(rays_b)
    assert np.array_equal(outs[1].view(np.uint8), ref.view(np.uint8))
    assert (outs[1]["prim"] != sk.MISS).mean() > 0.5
