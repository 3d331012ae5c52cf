"""Test scenes, including the fixtures of the reference's own TriangleAccelerator tests."""
import numpy as np

from skwarka_amd import workloads as wl


def reference_spheres():
    """UnitTests/MainTests/Raytracer/Core/TriangleAccelerator.test.h:35-43: four tetra-spheres."""
    return [wl.tetra_sphere((0, 0, 0), 10.0, 5), wl.tetra_sphere((-20, 0, 0), 20.0, 5),
            wl.tetra_sphere((0, -10, 0), 0.5, 5), wl.tetra_sphere((1000, 0, 0), 1000.0, 6)]


def small_spheres(seed=0, n=6, subdiv=3):
    rng = np.random.default_rng(seed)
    meshes = []
    for _ in range(n):
        c = rng.uniform(-2, 2, 3)
        r = rng.uniform(0.2, 0.8)
        meshes.append(wl.icosphere(subdiv, r, c, noise=0.02, rng=rng))
    meshes.append(wl.box(3.0))
    return meshes


def tetra_mesh():
    """TriangleMeshTestHelper::ConstructTetrahedron (TriangleMeshTestHelper.h:46-66) at the origin."""
    br = np.sqrt(8.0) / 3.0
    m_pi_3 = 1.04719755119659774615
    v = np.array([[0.0, 0.0, 1.0], [br, 0.0, -1.0 / 3.0],
                  [br * np.cos(2.0 * m_pi_3), br * np.sin(2.0 * m_pi_3), -1.0 / 3.0],
                  [br * np.cos(2.0 * m_pi_3), -br * np.sin(2.0 * m_pi_3), -1.0 / 3.0]], np.float32)
    t = np.array([[1, 3, 2], [1, 0, 3], [1, 2, 0], [0, 2, 3]], np.uint32)
    return {"verts": v, "tris": t}


def rot(axis, a):
    c, s = np.cos(a), np.sin(a)
    m = np.eye(4)
    i, j = {0: (1, 2), 1: (2, 0), 2: (0, 1)}[axis]
    m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
    return m


def instanced_tetrahedra(n, seed=3):
    """TriangleAccelerator.test.h:164-205: n instances of one tetrahedron (translation, scale,
    rotations) and the same tetrahedra pre-transformed into separate meshes."""
    rng = np.random.default_rng(seed)
    tet = tetra_mesh()
    xfs = []
    for _ in range(n):
        m = np.eye(4)
        m[:3, 3] = rng.uniform(0, 100, 3)
        s = np.diag([rng.uniform(0, 3) + 0.01, rng.uniform(0, 3) + 0.01, rng.uniform(0, 3) + 0.01, 1.0])
        m = m @ s @ rot(1, rng.uniform(0, 6)) @ rot(2, rng.uniform(0, 6)) @ rot(0, rng.uniform(0, 6))
        xfs.append(m)
    inst_meshes = [tet]
    inst_prims = [(0, m) for m in xfs]
    flat = []
    for m in xfs:
        v = tet["verts"].astype(np.float64)
        w = (np.c_[v, np.ones(len(v))] @ m.T)[:, :3]
        flat.append({"verts": w.astype(np.float32), "tris": tet["tris"]})
    return inst_meshes, inst_prims, flat


def random_rays(n, lo, hi, seed, tmin=0.0, tmax=np.inf):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return wl.make_rays(o, d, tmin, tmax)


def staircase(n=400, ratio=1.1):
    """Triangles at exponentially shrinking x: the split search peels them off one at a time, so the
    tree reaches the builder's depth limit of 200 (TriangleAccelerator.cpp:632) and the traversal
    needs more than the 128 stack entries the LDS variants hold."""
    xs = ratio ** (-np.arange(n, dtype=np.float64))
    verts, tris = [], []
    for i, x in enumerate(xs):
        s = x * 0.05
        verts += [[x, -1, -1], [x + s, 1, -1], [x, 1, 1]]
        tris.append([3 * i, 3 * i + 1, 3 * i + 2])
    return [{"verts": np.array(verts, np.float32), "tris": np.array(tris, np.uint32)}]


def staircase_rays(n, seed):
    """Rays along +x through the staircase, slightly spread."""
    rng = np.random.default_rng(seed)
    o = np.c_[np.full(n, -0.5), rng.uniform(-0.9, 0.9, (n, 2))]
    d = np.c_[np.ones(n), rng.normal(0, 0.02, (n, 2))]
    d /= np.linalg.norm(d, axis=1)[:, None]
    return wl.make_rays(o, d)
