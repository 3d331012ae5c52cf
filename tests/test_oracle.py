"""Pins the oracle (CPU restatement) to what fixes it independently of its own formulas:
known answers, brute force as the reference's own tests use it, closed forms and invariants.

Reference tests followed:
  UnitTests/MainTests/Raytracer/Core/TriangleAccelerator.test.h (Intersect / IntersectTest vs brute
      force, Epsilon, ObjectInstancing, GetWorldBounds)
  UnitTests/MainTests/Raytracer/Core/KDTree.test.h (nearest / k-nearest vs brute force, filters,
      duplicates, empty tree)
  UnitTests/MainTests/Math/CompressedDirection.test.h (codec error <= 0.017, normalised)
"""
import numpy as np
import pytest

import oracle
from skwarka_amd import workloads as wl
from tests import scenes

MISS = 0xFFFFFFFF


def test_mt19937_known_answer():
    # C++11 [rand.predef]: the 10000th output of a default-constructed std::mt19937 is 4123659995.
    # boost::mt19937 (RandomGenerator.h:30) is the same engine with the same default seed 5489.
    out = oracle.mt19937(5489, 10000)
    assert int(out[-1]) == 4123659995
    assert np.array_equal(out, wl.mt19937_raw(5489, 10000))


def test_reference_random_generator_range():
    rg = wl.RefRandom()
    v = rg(-10.0, 10.0, n=10000)
    assert v.min() >= -10.0 and v.max() < 10.0
    # RandomGenerator<double>::operator() = raw * (range * 2^-32): the first raw value of seed 5489
    first = wl.mt19937_raw(5489, 1)[0]
    assert wl.RefRandom()(2.0) == first * (2.0 * 2.0 ** -32)


def test_compressed_direction_codec():
    # CompressedDirection.test.h:27-53 on a coarser grid: decode error <= 0.017 and unit length.
    table = oracle.cdir_table()
    assert np.allclose(np.linalg.norm(table, axis=1), 1.0, atol=1e-15)
    assert np.array_equal(table, wl.cdir_table())
    g = np.linspace(-1, 1, 41)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    v = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    v = v[np.linalg.norm(v, axis=1) > 0.1]
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    ids = wl.cdir_encode(v)
    for k in range(0, len(v), 997):
        assert oracle.cdir_encode(*v[k]) == ids[k]
    err = np.linalg.norm(table[ids] - v, axis=1)
    assert err.max() <= 0.017


def _brute_check(acc, rays):
    hits = acc.closest_hit(rays)
    bh = acc.brute_closest(rays)
    hit_a = hits["prim"] != MISS
    hit_b = bh["prim"] != MISS
    assert np.array_equal(hit_a, hit_b)
    assert np.array_equal(hits["prim"], bh["prim"])
    assert np.array_equal(hits["tri"], bh["tri"])
    assert np.all(np.abs(hits["t"][hit_a] - bh["t"][hit_a]) <= 1e-9)
    occ = acc.any_hit(rays)
    assert np.array_equal(occ.astype(bool), hit_b)
    return hits


def test_accelerator_vs_brute_force_reference_fixture():
    # TriangleAccelerator.test.h:35-99, 137-162: four tetra-spheres, 1000 rays from a
    # default-seeded RandomGenerator<double>.
    meshes = scenes.reference_spheres()
    acc = oracle.Accel(meshes)
    b = acc.bounds()
    allv = np.concatenate([m["verts"] for m in meshes]).astype(np.float64)
    assert np.array_equal(b[:3], allv.min(0)) and np.array_equal(b[3:], allv.max(0))
    rays = wl.reference_test_rays(b, 1000)
    hits = _brute_check(acc, rays)
    assert (hits["prim"] != MISS).sum() > 100


def test_accelerator_vs_brute_force_icospheres():
    meshes = scenes.small_spheres(1)
    acc = oracle.Accel(meshes)
    rays = scenes.random_rays(2000, -2.5, 2.5, seed=5)
    _brute_check(acc, rays)


def test_accelerator_invariants():
    meshes = scenes.small_spheres(2, n=10, subdiv=3)
    acc = oracle.Accel(meshes)
    bbox, meta, tp, tt = acc.dump()
    info = acc.info()
    # every triangle in exactly one leaf
    leaves = meta[meta[:, 0] == 0]
    cover = np.zeros(info["triangles"], np.int64)
    for l in leaves:
        cover[l[2]:l[3]] += 1
        # leaf size <= 4 unless forced (depth / middle mask)
    assert np.all(cover == 1)
    assert np.mean((leaves[:, 3] - leaves[:, 2]) <= 4) > 0.95
    # (prim, tri) pairs are a permutation of all triangles
    pairs = set(zip(tp.tolist(), tt.tolist()))
    assert len(pairs) == info["triangles"]
    # each leaf box is the union of its triangles' boxes
    flat = {}
    for pi, m in enumerate(meshes):
        for ti, t in enumerate(m["tris"]):
            flat[(pi, ti)] = m["verts"][t]
    for l, bb in zip(meta, bbox):
        if l[0] != 0:
            continue
        pts = np.concatenate([flat[(tp[i], tt[i])] for i in range(l[2], l[3])])
        assert np.array_equal(bb[:3], pts.min(0)) and np.array_equal(bb[3:], pts.max(0))


def test_epsilon_no_self_hit():
    # TriangleAccelerator.test.h:101-135 (reduced: 200 x 50 outgoing rays).
    meshes = scenes.reference_spheres()[:3]
    acc = oracle.Accel(meshes)
    rg = wl.RefRandom()
    b = acc.bounds()
    rays = wl.reference_test_rays(b, 200, rg)
    hits, geom = acc.closest_hit(rays, geom=True)
    n_checked = 0
    for i in np.nonzero(hits["prim"] != MISS)[0]:
        d = np.stack([rg(1.0, n=50), rg(1.0, n=50), rg(1.0, n=50)], 1)
        d /= np.sqrt((d * d).sum(1))[:, None]
        out = wl.make_rays(np.repeat(geom[i]["p"][None], 50, 0), d)
        for j in range(50):
            out[j]["tmin"] = acc.next_min_t(hits[i]["prim"], geom[i], d[j])
        h2 = acc.closest_hit(out)
        same = (h2["prim"] == hits[i]["prim"]) & (h2["tri"] == hits[i]["tri"])
        assert not same.any()
        n_checked += 1
    assert n_checked > 20


def test_transformed_primitives_match_pretransformed():
    # TriangleAccelerator.test.h:164-205 (reduced to 500 tetrahedra / 2000 rays): primitives with
    # mesh-to-world transforms agree with the same triangles transformed beforehand.
    inst_meshes, inst_prims, flat = scenes.instanced_tetrahedra(500)
    separate = [inst_meshes[0]] * 0 + [scenes.tetra_mesh() for _ in inst_prims]
    acc1 = oracle.Accel(separate, [(i, m) for i, (_, m) in enumerate(inst_prims)])
    acc2 = oracle.Accel(flat)
    rng = np.random.default_rng(7)
    o = rng.uniform(0, 100, (2000, 3))
    d = rng.uniform(0, 1, (2000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = wl.make_rays(o, d)
    h1, g1 = acc1.closest_hit(rays, geom=True)
    h2, g2 = acc2.closest_hit(rays, geom=True)
    assert np.array_equal(h1["prim"] != MISS, h2["prim"] != MISS)
    hit = h1["prim"] != MISS
    assert hit.sum() > 50
    assert np.all(np.abs(h1["t"][hit] - h2["t"][hit]) <= 1e-3)
    assert np.array_equal(h1["tri"][hit], h2["tri"][hit])
    assert np.all(np.abs(g1["p"][hit] - g2["p"][hit]) <= 1e-3)


def test_instancing_matches_pretransformed():
    # One mesh shared by many primitives takes the instanced path (TriangleAccelerator.cpp:54-102).
    inst_meshes, inst_prims, flat = scenes.instanced_tetrahedra(300, seed=4)
    acc1 = oracle.Accel(inst_meshes, inst_prims)
    assert acc1.info()["instances"] == 300
    acc2 = oracle.Accel(flat)
    rng = np.random.default_rng(8)
    o = rng.uniform(0, 100, (2000, 3))
    d = rng.uniform(0, 1, (2000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = wl.make_rays(o, d)
    h1 = acc1.closest_hit(rays)
    h2 = acc2.closest_hit(rays)
    hit = h1["prim"] != MISS
    assert np.array_equal(hit, h2["prim"] != MISS)
    assert hit.sum() > 50
    assert np.array_equal(h1["prim"][hit], h2["prim"][hit])
    assert np.array_equal(h1["tri"][hit], h2["tri"][hit])
    assert np.all(np.abs(h1["t"][hit] - h2["t"][hit]) <= 1e-3)
    occ = acc1.any_hit(rays)
    assert np.array_equal(occ.astype(bool), hit)


def test_degenerate_triangles_are_dropped_and_renumbered():
    # TriangleMesh.cpp:134-147,180-186: out-of-range or repeated indices are removed (stably).
    m = wl.icosphere(1)
    bad = np.array([[0, 0, 1], [1, 2, 10_000]], np.uint32)
    tris = np.concatenate([bad[:1], m["tris"][:10], bad[1:], m["tris"][10:]])
    acc_bad = oracle.Accel([{"verts": m["verts"], "tris": tris}])
    acc_ok = oracle.Accel([{"verts": m["verts"], "tris": m["tris"]}])
    rays = scenes.random_rays(500, -0.5, 0.5, seed=2)
    h1, h2 = acc_bad.closest_hit(rays), acc_ok.closest_hit(rays)
    assert np.array_equal(h1, h2)


def test_closest_hit_geometry_closed_form():
    # Sphere of tetra-subdivision: the hit point lies on its triangle's plane, the geometric normal
    # is the normalised edge cross product and the barycentrics reproduce the point.
    meshes = [wl.tetra_sphere((0, 0, 0), 1.0, 3)]
    acc = oracle.Accel(meshes)
    rays = scenes.random_rays(300, -0.3, 0.3, seed=9)
    h, g = acc.closest_hit(rays, geom=True)
    v = meshes[0]["verts"].astype(np.float64)
    t = meshes[0]["tris"]
    for i in range(len(rays)):
        tri = v[t[h[i]["tri"]]]
        p = (1 - h[i]["b1"] - h[i]["b2"]) * tri[0] + h[i]["b1"] * tri[1] + h[i]["b2"] * tri[2]
        assert np.allclose(p, g[i]["p"], atol=1e-12)
        assert np.allclose(rays[i]["o"] + rays[i]["d"] * h[i]["t"], g[i]["p"], atol=1e-12)
        n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        assert np.allclose(n / np.linalg.norm(n), g[i]["ng"], atol=1e-12)
        assert np.allclose(n, g[i]["cross"], atol=0)


def _kd_points(n=10000, seed=11, dtype=np.float64):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-500, 500, (n, 3))
    p = np.concatenate([p, p[:100]])  # KDTree.test.h:57-59: 100 duplicates
    return p.astype(dtype)


def _custom_filter(p):
    s = np.trunc(p.sum(1)).astype(np.int64)
    return np.fmod(s, 2) == 1


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("filt", [0, 1])
def test_kdtree_knn_vs_brute_force(dtype, filt):
    # KDTree.test.h:160-256: sorted distance multiset equals brute force; d^2 recomputes exactly.
    pts = _kd_points(dtype=dtype)
    tree = oracle.KDTree(pts)
    rng = np.random.default_rng(12 + filt)
    q = rng.uniform(-500, 500, (300, 3))
    r = rng.uniform(0, 1000, 300) * rng.uniform(0, 1, 300) ** 3
    k = rng.integers(1, 501, 300).astype(np.uint32)
    ids, d2, cnt = tree.knn(q, r, k, filt)
    p64 = pts.astype(np.float64)
    ok = _custom_filter(p64) if filt else np.ones(len(pts), bool)
    for i in range(len(q)):
        dd = p64 - q[i]
        dist = dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1] + dd[:, 2] * dd[:, 2]
        cand = np.sort(dist[ok & (dist < r[i] * r[i])])[:k[i]]
        got = d2[i, :cnt[i]]
        assert np.array_equal(np.sort(got), cand)
        sel = ids[i, :cnt[i]]
        rec = p64[sel] - q[i]
        assert np.array_equal(rec[:, 0] * rec[:, 0] + rec[:, 1] * rec[:, 1] + rec[:, 2] * rec[:, 2], got)


@pytest.mark.parametrize("filt", [0, 1])
def test_kdtree_nearest_vs_brute_force(filt):
    # KDTree.test.h:79-158 (ties accepted by equal distance).
    pts = _kd_points()
    tree = oracle.KDTree(pts)
    rng = np.random.default_rng(21 + filt)
    q = rng.uniform(-500, 500, (500, 3))
    r = rng.uniform(0, 1000, 500)
    got = tree.nearest(q, r, filt)
    ok = _custom_filter(pts) if filt else np.ones(len(pts), bool)
    for i in range(len(q)):
        dd = pts - q[i]
        dist = dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1] + dd[:, 2] * dd[:, 2]
        dist = np.where(ok & (dist < r[i] * r[i]), dist, np.inf)
        if not np.isfinite(dist.min()):
            assert got[i] == MISS
        else:
            assert got[i] != MISS and dist[got[i]] == dist.min()


def test_kdtree_empty_and_single():
    # KDTree.test.h:258-272
    empty = oracle.KDTree(np.zeros((0, 3), np.float64))
    ids, d2, cnt = empty.knn(np.zeros((3, 3)), 10.0, 5)
    assert np.all(cnt == 0)
    assert np.all(empty.nearest(np.zeros((3, 3)), 10.0) == MISS)
    one = oracle.KDTree(np.array([[1.0, 2.0, 3.0]]))
    ids, d2, cnt = one.knn(np.array([[1.0, 2.0, 4.0]]), 2.0, 5)
    assert cnt[0] == 1 and ids[0, 0] == 0 and d2[0, 0] == 1.0


def test_kdtree_structure_median_split():
    # KDTree.h:427-473: node i holds the median of its range on the longest axis; the left child is
    # i + 1; the right child index sits in bits 3..31; leaves have axis bits = 3.
    pts = _kd_points(2000, seed=5)
    tree = oracle.KDTree(pts)
    perm, split, word = tree.dump()
    tp = pts[perm]
    n = len(pts)

    def check(b, e):
        if b + 1 == e:
            assert word[b] & 3 == 3
            return
        axis = word[b] & 3
        assert axis < 3
        ext = tp[b:e].max(0) - tp[b:e].min(0)
        assert axis == int(np.argmax(ext)) or ext[axis] == ext.max()
        mid = (b + e) // 2
        assert split[b] == np.float32(tp[b, axis])
        left = tp[b + 1:mid + 1, axis]
        right = tp[mid + 1:e, axis]
        assert np.all(left <= tp[b, axis]) and np.all(right >= tp[b, axis])
        assert bool(word[b] & 4) == (b < mid)
        if mid + 1 < e:
            assert word[b] >> 3 == mid + 1
            check(mid + 1, e)
        else:
            assert word[b] >> 3 == (1 << 29) - 1
        if b < mid:
            check(b + 1, mid + 1)

    check(0, n)


def _photon_scene(n=20000, seed=1):
    mesh = wl.merge(wl.scene_c0())
    return mesh, wl.photons_on_mesh(mesh, n, seed)


def test_photon_irradiance_closed_form():
    # PhotonLTEIntegrator.cpp:127-169 recomputed from the returned set.
    mesh, ph = _photon_scene(30000)
    pm = oracle.PhotonMap(ph["pos"], ph["weight"], ph["dir"], ph["normal"])
    table = wl.cdir_table()
    rng = np.random.default_rng(3)
    sel = rng.choice(len(ph["pos"]), 200, replace=False)
    p = ph["pos"][sel].astype(np.float64) + rng.normal(0, 1e-3, (200, 3))
    n = table[ph["normal"][sel]]
    for k, r in ((100, 0.3), (8, 0.5)):
        e_ext, e_int, ids, d2, cnt = pm.irradiance(p, n, k, r, paths=3)
        assert cnt.max() > 0
        cos_t = 0.87
        sin_t = np.sqrt(1 - cos_t * cos_t)
        for i in range(len(p)):
            c = cnt[i]
            s = ids[i, :c]
            to = ph["pos"][s].astype(np.float64) - p[i]
            # every member passes PhotonFilter and lies strictly inside the radius
            assert np.all(table[ph["normal"][s]] @ n[i] > cos_t)
            assert np.all(np.abs(to @ n[i]) < sin_t * np.sqrt((to * to).sum(1)))
            assert np.all(d2[i, :c] < r * r)
            ext = np.zeros(3)
            itn = np.zeros(3)
            for j in s:
                if table[ph["dir"][j]] @ n[i] > 0.0:
                    ext += ph["weight"][j].astype(np.float64)
                else:
                    itn += ph["weight"][j].astype(np.float64)
            mx = d2[i, :c].max() if c else 0.0
            if c < k or mx == 0.0:
                mx = r * r
            else:
                mx *= c / (c - 0.5)
            inv = 1.0 / (np.pi * 3 * mx)
            if c:
                assert np.allclose(e_ext[i], (ext * inv).astype(np.float32), rtol=1e-6, atol=0)
                assert np.allclose(e_int[i], (itn * inv).astype(np.float32), rtol=1e-6, atol=0)


def _uniform_plane_photons(n=200000, seed=8):
    # photons uniformly on the square [-1,1]^2 of the plane z = 0, normal +z, arriving from +z,
    # equal weights: irradiance E = n * w / 4 everywhere away from the edges
    rng = np.random.default_rng(seed)
    pos = np.c_[rng.uniform(-1, 1, (n, 2)), np.zeros(n)].astype(np.float32)
    up = wl.cdir_encode(np.array([[0.0, 0.0, 1.0]]))[0]
    codes = np.full(n, up, np.uint16)
    w = np.full((n, 3), 2e-6, np.float32)
    q = np.c_[rng.uniform(-0.7, 0.7, (600, 2)), np.zeros(600)]
    qn = np.tile([0.0, 0.0, 1.0], (600, 1))
    e = n * 2e-6 / 4.0
    return oracle.PhotonMap(pos, w, codes, codes), q, qn, e


def test_cone_filter_is_unbiased_on_uniform_density():
    # Jensen's cone filter normalised by 1 - 2/(3 kc) integrates to one over the disc, so a uniform
    # photon density gives back its irradiance (the cone variant has no reference counterpart).
    pm, q, qn, e = _uniform_plane_photons()
    disc = pm.irradiance(q, qn, 100, 0.2, with_sets=False)[0]
    for kc in (1.0, 1.5):
        cone = pm.cone(q, qn, 100, 0.2, 1, kc)
        assert abs(cone.mean() / e - 1.0) < 0.03, (kc, cone.mean() / e)
    assert abs(disc.mean() / e - 1.0) < 0.03
    # a set that is not full uses r: a radius covering few photons
    cone_small = pm.cone(q, qn, 100, 0.005, 1, 1.0)
    assert np.all(np.isfinite(cone_small))


def test_caustic_simpson_kernel_on_uniform_density():
    # Simpson's kernel 3/pi (1 - d^2/R^2)^2 integrates to R^2 over the disc; with a full set the
    # running maximum starts at the heap top (the farthest photon), so a uniform density gives
    # (rho/pi) E (n - 0.5)/n for a Lambertian receiver seen from its side, and 0 from the other.
    pm, q, qn, e = _uniform_plane_photons()
    wo = np.tile([0.3, 0.0, np.sqrt(1 - 0.09)], (len(q), 1))
    rad = pm.caustic(q, qn, wo, 100, 0.2, 1, 0.5)
    expect = 0.5 / np.pi * e * (99.5 / 100.0)
    assert abs(rad.mean() / expect - 1.0) < 0.03, rad.mean() / expect
    assert np.all(pm.caustic(q, qn, -wo, 100, 0.2, 1, 0.5) == 0.0)


def test_photon_lookup_counts_closed_forms():
    """The oracle's per-lookup work counts (the C1 roofline's N_visit / N_filter / N_found): with a
    set that never fills the radius never shrinks, so the photons inside it are exactly the
    brute-force count within r; an unbounded radius visits every node; a zero radius visits one
    root-to-leaf path and finds nothing."""
    rng = np.random.default_rng(5)
    n = 3000
    pos = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    nrm = wl.cdir_encode(np.asarray([v / np.linalg.norm(v) for v in rng.normal(size=(n, 3))]))
    dirs = wl.cdir_encode(np.asarray([v / np.linalg.norm(v) for v in rng.normal(size=(n, 3))]))
    w = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    om = oracle.PhotonMap(pos, w, dirs, nrm)
    q = rng.uniform(-1, 1, (200, 3))
    qn = rng.normal(size=(200, 3))
    qn /= np.linalg.norm(qn, axis=1)[:, None]
    r = 0.3
    c = om.lookup_counts(q, qn, n + 1, r)
    d2 = ((pos.astype(np.float64)[None, :, :] - q[:, None, :]) ** 2).sum(-1)
    assert np.array_equal(c[:, 1], (d2 < r * r).sum(1))
    _, _, _, _, found = om.irradiance(q, qn, n + 1, r, with_sets=False)
    assert np.array_equal(c[:, 2], found) and (found <= c[:, 1]).all()
    assert (c[:, 0] >= c[:, 1]).all()
    big = om.lookup_counts(q[:5], qn[:5], n + 1, 100.0)
    assert (big[:, 0] == n).all() and (big[:, 1] == n).all()
    zero = om.lookup_counts(q[:20], qn[:20], 4, 0.0)
    assert (zero[:, 1] == 0).all() and (zero[:, 2] == 0).all()
    depth = int(np.ceil(np.log2(n + 1)))
    assert (zero[:, 0] >= depth - 1).all() and (zero[:, 0] <= depth).all()
