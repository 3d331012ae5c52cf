"""Seeded synthetic workloads: scenes, camera rays and photons.

Definitions (all generated on the host with numpy, float64 math, then stored as the reference's
types: float32 vertices, float64 rays, float32 photon powers, 16-bit CompressedDirection ids):

* icosphere(s): the 12-vertex icosahedron subdivided s times (20*4^s triangles), midpoints pushed
  to the unit sphere, optionally displaced radially by a relative uniform noise.
* box: the cube [-a, a]^3 as 12 triangles whose geometric normals (v1-v0)x(v2-v0) point inward.
* tetra_sphere(s): the reference's own test sphere (Source/RayLibs/Shapes/Sphere.cpp:42-197):
  a tetrahedron subdivided s times (4*4^s triangles), used by its TriangleAccelerator tests.
* camera rays: pinhole camera, x field of view, pixel centres, as PerspectiveCamera::GenerateRay
  (Source/RayLibs/Raytracer/Cameras/PerspectiveCamera.cpp:36-64) with a look-at frame
  (Source/RayLibs/Math/Transform.cpp:138-165).
* photons: positions uniform by area (triangle by area CDF, UniformTriangleSampling
  b0 = 1 - sqrt(u1), b1 = u2 sqrt(u1)), power U[0, pmax)^3, incident direction uniform over the
  hemisphere of the surface normal, direction and normal stored as CompressedDirection ids
  (Source/RayLibs/Math/CompressedDirection.h:92-124).
"""
import numpy as np

NO_GROUP = 0xFFFFFFFF

# ---------------------------------------------------------------- meshes


def icosphere(subdiv, radius=1.0, center=(0.0, 0.0, 0.0), noise=0.0, rng=None):
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                  (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                  (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], np.int64)
    for _ in range(subdiv):
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        e = np.stack([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)], 1).reshape(-1, 2)
        key = np.minimum(e[:, 0], e[:, 1]) * len(v) + np.maximum(e[:, 0], e[:, 1])
        uk, inv = np.unique(key, return_inverse=True)
        lo, hi = uk // len(v), uk % len(v)
        mid = v[lo] + v[hi]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        m = (inv + base).reshape(-1, 3)
        ab, bc, ca = m[:, 0], m[:, 1], m[:, 2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    normals = v.copy()
    r = np.full(len(v), float(radius))
    if noise:
        r = r * (1.0 + (rng.random(len(v)) * 2.0 - 1.0) * noise)
    verts = v * r[:, None] + np.asarray(center, np.float64)
    return {"verts": verts.astype(np.float32), "tris": f.astype(np.uint32), "normals": normals.astype(np.float32)}


def box(a=3.0, center=(0.0, 0.0, 0.0), inward=True):
    c = np.asarray(center, np.float64)
    p = np.array([[x, y, z] for x in (-a, a) for y in (-a, a) for z in (-a, a)], np.float64) + c
    # corner index = 4*ix + 2*iy + iz; faces listed with outward orientation, flipped for inward.
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for q in quads:
        tris.append((q[0], q[1], q[2]))
        tris.append((q[0], q[2], q[3]))
    tris = np.array(tris, np.uint32)
    # check orientation against the outward direction and flip as requested
    out = []
    for t in tris:
        v0, v1, v2 = p[t[0]], p[t[1]], p[t[2]]
        n = np.cross(v1 - v0, v2 - v0)
        outward = np.dot(n, (v0 + v1 + v2) / 3.0 - c) > 0
        if outward == inward:
            t = (t[0], t[2], t[1])
        out.append(t)
    return {"verts": p.astype(np.float32), "tris": np.array(out, np.uint32)}


def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3)."""
    v = np.array([p0, p1, p2, p3], np.float64)
    return {"verts": v.astype(np.float32), "tris": np.array([(0, 1, 2), (0, 2, 3)], np.uint32)}


def tetra_sphere(center, radius, subdiv):
    """Source/RayLibs/Shapes/Sphere.cpp:42-197 (vertex order and edge reuse as there)."""
    br = np.sqrt(8.0) / 3.0
    m_pi_3 = 1.04719755119659774615
    verts = [np.array([0.0, 0.0, 1.0]), np.array([br, 0.0, -1.0 / 3.0]),
             np.array([br * np.cos(2.0 * m_pi_3), br * np.sin(2.0 * m_pi_3), -1.0 / 3.0]),
             np.array([br * np.cos(2.0 * m_pi_3), -br * np.sin(2.0 * m_pi_3), -1.0 / 3.0])]
    tris = [(1, 3, 2), (1, 0, 3), (1, 2, 0), (0, 2, 3)]
    deleted = [False] * 4
    e2v = {}
    for _ in range(subdiv):
        n0 = len(tris)
        for j in range(n0):
            if deleted[j]:
                continue
            i0, i1, i2 = tris[j]
            deleted[j] = True

            def mid(a, b):
                m = (verts[a] + verts[b]) * 0.5
                return m * (1.0 / np.sqrt(np.dot(m, m)))

            nv = []
            for a, b in ((i0, i1), (i1, i2), (i2, i0)):
                if (a, b) not in e2v:
                    verts.append(mid(a, b))
                    e2v[(a, b)] = e2v[(b, a)] = len(verts) - 1
                nv.append(e2v[(a, b)])
            tris += [(i0, nv[0], nv[2]), (i1, nv[1], nv[0]), (i2, nv[2], nv[1]), (nv[0], nv[1], nv[2])]
            deleted += [False] * 4
    tris = np.array([t for t, d in zip(tris, deleted) if not d], np.uint32)
    v = np.array(verts)
    world = v * radius + np.asarray(center, np.float64)
    return {"verts": world.astype(np.float32), "tris": tris, "normals": v.astype(np.float32)}


def merge(meshes):
    """Concatenates meshes into one."""
    vs, ts, ns = [], [], []
    off = 0
    for m in meshes:
        vs.append(m["verts"])
        ts.append(m["tris"].astype(np.int64) + off)
        ns.append(m.get("normals") if m.get("normals") is not None else np.zeros_like(m["verts"]))
        off += len(m["verts"])
    return {"verts": np.concatenate(vs), "tris": np.concatenate(ts).astype(np.uint32), "normals": np.concatenate(ns)}


# ---------------------------------------------------------------- rays


def camera_rays(width, height, origin, target_dir, up, x_fov_deg, spp=1, rng=None):
    """Primary rays at pixel centres (spp=1) or uniformly jittered inside the pixel."""
    d = np.asarray(target_dir, np.float64)
    d = d * (1.0 / np.sqrt(np.dot(d, d)))
    right = np.cross(d, np.asarray(up, np.float64))
    right = right * (1.0 / np.sqrt(np.dot(right, right)))
    new_up = np.cross(right, d)
    m = np.stack([right, -new_up, d], 1)  # camera-to-world rotation (columns)
    x_tan = np.tan(np.deg2rad(x_fov_deg) / 2.0)
    y_tan = x_tan * float(height) / float(width)
    ys, xs = np.meshgrid(np.arange(height, dtype=np.float64), np.arange(width, dtype=np.float64), indexing="ij")
    xs = np.repeat(xs.reshape(-1), spp)
    ys = np.repeat(ys.reshape(-1), spp)
    if spp == 1 and rng is None:
        px, py = xs + 0.5, ys + 0.5
    else:
        px, py = xs + rng.random(len(xs)), ys + rng.random(len(ys))
    cx = 2.0 * x_tan * (px - width * 0.5) / width
    cy = 2.0 * y_tan * (py - height * 0.5) / height
    cam = np.stack([cx, cy, np.ones_like(cx)], 1)
    w = cam @ m.T
    w = w * (1.0 / np.sqrt((w * w).sum(1)))[:, None]
    rays = np.zeros(len(w), RAY_DTYPE())
    rays["o"] = np.asarray(origin, np.float64)
    rays["d"] = w
    rays["tmin"] = 0.0
    rays["tmax"] = np.inf
    return rays


def RAY_DTYPE():
    from . import RAY_DTYPE as dt
    return dt


def make_rays(o, d, tmin=0.0, tmax=np.inf):
    o = np.asarray(o, np.float64).reshape(-1, 3)
    d = np.asarray(d, np.float64).reshape(-1, 3)
    r = np.zeros(max(len(o), len(d)), RAY_DTYPE())
    r["o"] = o
    r["d"] = d
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def mt19937_raw(seed, n):
    """Raw 32-bit outputs of std::mt19937 / boost::mt19937 seeded with `seed`."""
    bg = np.random.MT19937(0)
    bg._legacy_seeding(int(seed))
    return bg.random_raw(n).astype(np.uint32)


class RefRandom:
    """RandomGenerator<double> (Source/RayLibs/Math/RandomGenerator.h:81-99): value =
    raw * (range * 2^-32) over a default-seeded (5489) mt19937."""

    def __init__(self, seed=5489):
        self.bg = np.random.MT19937(0)
        self.bg._legacy_seeding(int(seed))

    def __call__(self, lo, hi=None, n=None):
        if hi is None:
            lo, hi = 0.0, lo
        raw = self.bg.random_raw(1 if n is None else n).astype(np.float64)
        v = lo + raw * ((hi - lo) * (1.0 / 4294967296.0))
        return v[0] if n is None else v


def reference_test_rays(bounds, n, rg=None):
    """The random rays of the reference's TriangleAccelerator tests (TriangleAccelerator.test.h:65-99):
    origin uniform in the world box, direction from three U[0,1) values normalised, tmin U[-10,10),
    tmax U[100,10000). Arguments are drawn left to right."""
    rg = rg or RefRandom()
    out = np.zeros(n, RAY_DTYPE())
    for i in range(n):
        o = [rg(bounds[0], bounds[3]), rg(bounds[1], bounds[4]), rg(bounds[2], bounds[5])]
        d = np.array([rg(1.0), rg(1.0), rg(1.0)])
        d = d * (1.0 / np.sqrt(np.dot(d, d)))
        out[i]["o"] = o
        out[i]["d"] = d
        out[i]["tmin"] = rg(-10.0, 10.0)
        out[i]["tmax"] = rg(100.0, 10000.0)
    return out


# ---------------------------------------------------------------- photons


def cdir_encode(v):
    """CompressedDirection(Vector3D) for an array of unit vectors (N,3) -> uint16 ids."""
    v = np.asarray(v, np.float64).reshape(-1, 3)
    data = np.zeros(len(v), np.uint32)
    a = v.copy()
    for c, bit in ((0, 13), (1, 14), (2, 15)):
        neg = a[:, c] < 0
        data[neg] |= np.uint32(1 << bit)
        a[neg, c] = -a[neg, c]
    scale = 126.999 / (a[:, 0] + a[:, 1] + a[:, 2])
    x = np.floor(a[:, 0] * scale).astype(np.uint32)
    y = np.floor(a[:, 1] * scale).astype(np.uint32)
    flip = y >= 64
    x = np.where(flip, 127 - x, x)
    y = np.where(flip, 127 - y, y)
    data |= (x << 6) | y
    return data.astype(np.uint16)


def cdir_table():
    """The 65536-entry decode table of CompressedDirection.cpp:20-44 (float64)."""
    i = np.arange(65536, dtype=np.int64)
    x = (i & 0x1FC0) >> 6
    y = i & 0x3F
    flip = (x + y) >= 127
    x = np.where(flip, 127 - x, x)
    y = np.where(flip, 127 - y, y)
    d = np.stack([x + 0.5, y + 0.5, (126 - x - y).astype(np.float64)], 1)
    l2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
    inv = 1.0 / np.sqrt(l2)
    d = d * inv[:, None]
    for c, bit in ((0, 13), (1, 14), (2, 15)):
        s = (i >> bit) & 1
        d[:, c] = np.where(s == 1, -d[:, c], d[:, c])
    return d


def triangle_normals(verts, tris):
    v = verts.astype(np.float64)
    e1 = v[tris[:, 1]] - v[tris[:, 0]]
    e2 = v[tris[:, 2]] - v[tris[:, 0]]
    n = np.cross(e1, e2)
    area = 0.5 * np.linalg.norm(n, axis=1)
    n = n / np.linalg.norm(n, axis=1, keepdims=True)
    return n, area


def photons_on_mesh(mesh, n, seed, power_max=1e-6):
    rng = np.random.default_rng(seed)
    v = mesh["verts"].astype(np.float64)
    t = mesh["tris"]
    nrm, area = triangle_normals(mesh["verts"], t)
    cdf = np.cumsum(area)
    cdf /= cdf[-1]
    tri = np.minimum(np.searchsorted(cdf, rng.random(n), side="right"), len(t) - 1)
    u1, u2 = rng.random(n), rng.random(n)
    su = np.sqrt(u1)
    b0 = 1.0 - su
    b1 = u2 * su
    b2 = 1.0 - b0 - b1
    pos = v[t[tri, 0]] * b0[:, None] + v[t[tri, 1]] * b1[:, None] + v[t[tri, 2]] * b2[:, None]
    weight = (rng.random((n, 3)) * power_max).astype(np.float32)
    # uniform hemisphere around the normal: z = u, phi = 2 pi u'
    z = rng.random(n)
    phi = 2.0 * np.pi * rng.random(n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    nn = nrm[tri]
    helper = np.where(np.abs(nn[:, 0:1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    tu = np.cross(nn, helper)
    tu /= np.linalg.norm(tu, axis=1, keepdims=True)
    tv = np.cross(nn, tu)
    d = tu * (r * np.cos(phi))[:, None] + tv * (r * np.sin(phi))[:, None] + nn * z[:, None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return {
        "pos": pos.astype(np.float32),
        "weight": weight,
        "dir": cdir_encode(d),
        "normal": cdir_encode(nn),
    }


def mesh_area(mesh):
    """TriangleMesh::GetArea (Source/RayLibs/Raytracer/Core/TriangleMesh.cpp:195-205): float
    triangle areas (Triangle3D<float>::GetArea, Math/Triangle3D.h:82-86) summed in a float, in
    triangle order."""
    v = mesh["verts"].astype(np.float32)
    t = mesh["tris"].astype(np.int64)
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    e1, e2 = b - a, c - a
    cx = e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1]
    cy = e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2]
    cz = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
    area = np.float32(0.5) * np.sqrt(cx * cx + cy * cy + cz * cz)
    if len(area) == 0:
        return np.float32(0.0)
    return np.cumsum(area, dtype=np.float32)[-1]  # sequential float accumulation


def scene_area(meshes, prims=None):
    """PhotonLTEIntegrator's scene area (PhotonLTEIntegrator.cpp:51-54): the double sum over
    primitives of their mesh's float area (the untransformed mesh)."""
    areas = [mesh_area(m) for m in meshes]
    total = 0.0
    for p in (range(len(meshes)) if prims is None else [pr[0] for pr in prims]):
        total += float(areas[p])
    return total


def photon_maps(mesh, n_direct, n_indirect, n_caustic, seed, power_max=1e-6):
    """Three seeded photon sets on a mesh (photons_on_mesh with seeds seed, seed+1, seed+2): direct,
    indirect and caustic (None when n_caustic is 0)."""
    d = photons_on_mesh(mesh, n_direct, seed, power_max)
    i = photons_on_mesh(mesh, n_indirect, seed + 1, power_max)
    c = photons_on_mesh(mesh, n_caustic, seed + 2, power_max) if n_caustic > 0 else None
    return d, i, c


# ---------------------------------------------------------------- named workloads


def scene_c0(seed=0):
    """81,920-triangle icosphere (r=1, radial noise 0.02) inside the 12-triangle box [-3,3]^3."""
    rng = np.random.default_rng(seed)
    return [icosphere(6, 1.0, (0, 0, 0), noise=0.02, rng=rng), box(3.0)]


def rays_c0(width=512, height=512):
    return camera_rays(width, height, (0.0, 0.0, -2.9), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0), 60.0)


def scene_room(seed=2, n_spheres=200, subdiv_counts=(70, 111, 19), subdiv_levels=(5, 4, 3), room=4.0,
               light_radiance=8.0):
    """Seeded direct-lighting room: the cube [-room, room]^3 (inward normals, matte 0.5), n_spheres
    displaced icospheres (radii log-uniform in [0.05, 0.5], +-2% radial noise, centres uniform in
    the room), the largest subdiv_counts[0] spheres at subdiv_levels[0] and so on (70/111/19 at
    5/4/3 gives 2,026,240 sphere triangles), and 4 quad area lights of 1 x 1 just under the ceiling
    facing down. Primitive order: room, spheres, lights. Materials: prim p belongs to reflectance /
    roughness group p % 64 (64 groups, gray reflectance U[0.1, 0.9], roughness U[0.01, 0.3]);
    spheres with odd index are plastic (glossy reflectance 0.04), the rest matte; lights are black
    matte. Light radiance is light_radiance times a per-light tint U[0.8, 1.2]^3."""
    rng = np.random.default_rng(seed)
    radii = np.exp(rng.uniform(np.log(0.05), np.log(0.5), n_spheres))
    centers = rng.uniform(-room + 0.6, room - 0.6, (n_spheres, 3))
    order = np.argsort(-radii, kind="stable")
    level = np.empty(n_spheres, np.int64)
    start = 0
    for cnt, lev in zip(subdiv_counts, subdiv_levels):
        level[order[start:start + cnt]] = lev
        start += cnt
    level[order[start:]] = subdiv_levels[-1]
    meshes = [box(room)]
    for i in range(n_spheres):
        meshes.append(icosphere(int(level[i]), float(radii[i]), centers[i], noise=0.02, rng=rng))
    y = room - 0.01
    light_pos = [(-1.5, -1.5), (1.5, -1.5), (-1.5, 1.5), (1.5, 1.5)]
    for (x, z) in light_pos:
        # vertices ordered so that (v1 - v0) x (v2 - v0) points down (-y)
        meshes.append(quad((x - 0.5, y, z - 0.5), (x + 0.5, y, z - 0.5), (x + 0.5, y, z + 0.5), (x - 0.5, y, z + 0.5)))
    n_prims = len(meshes)
    group_refl = rng.uniform(0.1, 0.9, 64)
    group_refl[0] = 0.5  # group 0 holds the room
    group_rough = rng.uniform(0.01, 0.3, 64)
    mat = np.zeros(n_prims, np.uint8)
    kd = np.zeros((n_prims, 3))
    ks = np.zeros((n_prims, 3))
    rough = np.zeros(n_prims)
    prim_group = np.full(n_prims, NO_GROUP, np.uint32)
    for p in range(n_prims - 4):
        g = p % 64
        prim_group[p] = g
        kd[p] = group_refl[g]
        rough[p] = group_rough[g]
        if 1 <= p <= n_spheres and (p - 1) % 2 == 1:
            mat[p] = 1
            ks[p] = 0.04
    rough[n_prims - 4:] = 0.1
    light_prim = np.arange(n_prims - 4, n_prims, dtype=np.uint32)
    light_rad = light_radiance * rng.uniform(0.8, 1.2, (4, 3))
    theta0 = np.concatenate([group_refl, group_rough, np.ones(4)])
    return {"meshes": meshes, "mat": mat, "kd": kd, "ks": ks, "rough": rough, "light_prim": light_prim,
            "light_radiance": light_rad, "group_refl": group_refl, "group_rough": group_rough,
            "prim_group": prim_group, "n_groups": 64, "theta0": theta0,
            "camera": ((0.0, 0.0, -room + 0.1), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0), 60.0)}


def perturbed_theta(theta0, n_groups=64, seed=4):
    """Target parameters of the material-fitting workload: reflectance + N(0, 0.05), roughness +
    N(0, 0.02), emission scales x 1.2."""
    rng = np.random.default_rng(seed)
    t = np.array(theta0, np.float64)
    t[:n_groups] += rng.normal(0.0, 0.05, n_groups)
    t[n_groups:2 * n_groups] += rng.normal(0.0, 0.02, n_groups)
    t[2 * n_groups:] *= 1.2
    return t


def scene_room_small(seed=5):
    """A small room (20 spheres at subdivision 3 and 2) for parity tests."""
    return scene_room(seed=seed, n_spheres=20, subdiv_counts=(8, 12), subdiv_levels=(3, 2))


def batch_scenes(n=64, seed=3, tri_lo=1e3, tri_hi=1e6, sizes=(128, 256, 512), spp=4):
    """The batched workload: n seeded scenes whose triangle counts are log-uniform in
    [tri_lo, tri_hi] (icospheres of subdivision 2-5 filling a unit-ish volume until the count is
    reached, inside a 12-triangle box), image sizes drawn from `sizes` (square), spp samples.
    Returns a list of (meshes, camera tuple, size, spp, seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        target = float(np.exp(rng.uniform(np.log(tri_lo), np.log(tri_hi))))
        meshes = [box(3.0)]
        total = 12
        while total < target:
            left = target - total
            sub = 5 if left > 20480 * 1.5 else (4 if left > 5120 * 1.5 else (3 if left > 1280 * 1.5 else 2))
            r = float(np.exp(rng.uniform(np.log(0.1), np.log(0.8))))
            c = rng.uniform(-2.2, 2.2, 3)
            m = icosphere(sub, r, c, noise=0.02, rng=rng)
            meshes.append(m)
            total += len(m["tris"])
        size = int(sizes[rng.integers(0, len(sizes))])
        out.append((meshes, ((0.0, 0.0, -2.9), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0), 60.0), size, spp, 1000 + i))
    return out


def batch_gather_jobs(specs, n_jobs=8, photons=250_000, queries=65_536, k=100, seed=5):
    """Photon-gather jobs for the batched workload: job j gathers on scene j * len(specs) / n_jobs,
    whose surfaces carry `photons` photons (uniform by area), at `queries` points taken from those
    photons' positions with their normals; k nearest with the PhotonFilter in the radius that holds
    40 photons on average (r^2 = 40 A / (pi N)), disc estimate, albedo 0.5. Returns a list of dicts
    with the photon arrays (ph), query points (p), normals (n), k, r, paths and rho."""
    out = []
    table = cdir_table()
    for j in range(n_jobs):
        i = j * len(specs) // n_jobs
        meshes = specs[i][0]
        mesh = merge(meshes)
        ph = photons_on_mesh(mesh, photons, seed=seed + j)
        rng = np.random.default_rng(seed + 100 + j)
        sel = rng.choice(photons, queries, replace=False)
        p = ph["pos"][sel].astype(np.float64)
        n = table[ph["normal"][sel]]
        r = float(np.sqrt(40.0 * scene_area(meshes) / (np.pi * photons)))
        out.append({"scene": i, "ph": ph, "p": np.ascontiguousarray(p), "n": np.ascontiguousarray(n), "k": k, "r": r,
                    "paths": 1, "rho": 0.5})
    return out
