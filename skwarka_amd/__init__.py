"""Python access to libskwarka_rt (include/skwarka_rt.h) through ctypes.

The library is the product: host tree builders, gfx950 kernels and the C ABI. This module only
loads it, describes its record layouts as numpy dtypes and wraps the handles. Device buffers are
passed as integer device addresses (for example torch CUDA tensors' data_ptr()). There is no CPU
path here: loading fails loudly when the library has not been built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libskwarka_rt.so")

RAY_DTYPE = np.dtype([("o", "<f8", (3,)), ("d", "<f8", (3,)), ("tmin", "<f8"), ("tmax", "<f8")])
HIT_DTYPE = np.dtype([("prim", "<u4"), ("tri", "<u4"), ("t", "<f8"), ("b1", "<f8"), ("b2", "<f8")])
GEOM_DTYPE = np.dtype([("p", "<f8", (3,)), ("ng", "<f8", (3,)), ("ns", "<f8", (3,)),
                       ("cross", "<f8", (3,)), ("dot", "<f8")])
assert RAY_DTYPE.itemsize == 64 and HIT_DTYPE.itemsize == 32 and GEOM_DTYPE.itemsize == 104

MISS = 0xFFFFFFFF
SCENE_HOST_ONLY = 1
SCENE_DEVICE_BUILD = 2

# Every symbol include/skwarka_rt.h declares.
EXPORTS = [
    "skw_scene_create", "skw_scene_destroy", "skw_scene_bounds", "skw_scene_info",
    "skw_scene_export_tree", "skw_closest_hit", "skw_closest_hit_stats", "skw_closest_hit_noalloc", "skw_any_hit", "skw_hit_geometry",
    "skw_photon_map_create", "skw_photon_map_destroy", "skw_photon_map_export", "skw_photon_map_info",
    "skw_photon_irradiance", "skw_photon_irradiance_stats", "skw_photon_knn", "skw_photon_nearest", "skw_photon_cone_irradiance",
    "skw_photon_caustic_radiance", "skw_last_error",
    "skw_device_count", "skw_dl_scene_create", "skw_dl_scene_destroy", "skw_dl_render",
    "skw_dl_render_profiled", "skw_dl_param_count", "skw_dl_render_replay", "skw_dl_adjoint",
    "skw_dl_render_tape", "skw_dl_tape_adjoint", "skw_dl_tape_destroy",
    "skw_batch_create", "skw_batch_create_with_gathers", "skw_batch_destroy", "skw_batch_configure", "skw_batch_autotune", "skw_batch_frame",
    "skw_batch_hits", "skw_batch_latency", "skw_batch_check", "skw_batch_rays", "skw_batch_ray_count",
    "skw_irradiance_map_create", "skw_irradiance_map_create_from", "skw_irradiance_map_destroy",
    "skw_irradiance_map_info", "skw_irradiance_map_export", "skw_dl_final_gather", "skw_pm_render",
    "skw_photon_shoot", "skw_photon_set_info", "skw_photon_set_export", "skw_photon_set_map",
    "skw_photon_set_destroy", "skw_photon_map_build",
    "skw_scene_save", "skw_scene_load", "skw_photon_map_save", "skw_photon_map_load", "skw_irradiance_map_save",
    "skw_irradiance_map_load", "skw_pbrt_load", "skw_pbrt_descs", "skw_pbrt_warnings", "skw_pbrt_warning",
    "skw_pbrt_destroy", "skw_sort_pairs_u32",
]

_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_U32 = ctypes.c_uint32


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("n_meshes", _U32), ("mesh_vert_offset", _P), ("verts", _P), ("normals", _P),
        ("mesh_tri_offset", _P), ("tris", _P), ("mesh_flags", _P),
        ("n_prims", _U32), ("prim_mesh", _P), ("prim_xform", _P),
    ]


class DLDesc(ctypes.Structure):
    _fields_ = [("mat_type", _P), ("kd", _P), ("ks", _P), ("roughness", _P), ("n_lights", _U32),
                ("light_prim", _P), ("light_radiance", _P), ("n_groups", _U32), ("prim_group", _P)]


class Camera(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_double * 3), ("rot", ctypes.c_double * 9), ("x_tan", ctypes.c_double),
                ("y_tan", ctypes.c_double)]


class DLParams(ctypes.Structure):
    _fields_ = [("width", _U32), ("height", _U32), ("spp", _U32), ("light_samples", _U32),
                ("bsdf_samples", _U32), ("seed", _U64)]


class DLProfile(ctypes.Structure):
    _fields_ = [("stage_ms", ctypes.c_double * 6), ("rays", _U64 * 3), ("node_tests", _U64 * 3),
                ("tri_tests", _U64 * 3), ("launches", _U64)]

    STAGES = ("camera", "primary_trace", "shade", "occlusion_trace", "bsdf_trace", "resolve")

    def as_dict(self):
        return {"stage_ms": dict(zip(self.STAGES, list(self.stage_ms))), "rays": list(self.rays),
                "node_tests": list(self.node_tests), "tri_tests": list(self.tri_tests),
                "launches": int(self.launches)}


class IrrMapDesc(ctypes.Structure):
    _fields_ = [("caustic", _P), ("direct", _P), ("indirect", _P), ("caustic_paths", _U64),
                ("direct_paths", _U64), ("indirect_paths", _U64), ("scene_area", ctypes.c_double)]


class FGParams(ctypes.Structure):
    _fields_ = [("width", _U32), ("height", _U32), ("spp", _U32), ("gather_samples", _U32), ("seed", _U64)]


class PMParams(ctypes.Structure):
    _fields_ = [("width", _U32), ("height", _U32), ("spp", _U32), ("light_samples", _U32), ("bsdf_samples", _U32),
                ("gather_samples", _U32), ("seed", _U64)]


class ShootParams(ctypes.Structure):
    _fields_ = [("max_paths", _U64), ("max_photons", _U64 * 3), ("paths_per_round", _U32), ("seed", _U64)]


class BatchView(ctypes.Structure):
    _fields_ = [("cam", Camera), ("width", _U32), ("height", _U32), ("spp", _U32), ("seed", _U64)]


class BatchGather(ctypes.Structure):
    _fields_ = [("map", _P), ("d_points", _P), ("d_normals", _P), ("nq", _U64), ("k", _U32), ("r", ctypes.c_double),
                ("paths", _U64), ("rho", ctypes.c_double), ("d_e_ext", _P), ("d_e_int", _P), ("d_radiance", _P)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("waves_per_block", _U32), ("stack_entries", _U32), ("chunk_rays", _U32), ("waves_per_cu", _U32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


_lib = None


def lib():
    """Loads libskwarka_rt.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libskwarka_rt.so is not built (run __graft_entry__.build() or "
                           "python skwarka_amd/build.py)")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "skw_scene_create": ([ctypes.POINTER(SceneDesc), _U32, ctypes.POINTER(_P)], ctypes.c_int),
        "skw_scene_destroy": ([_P], ctypes.c_int),
        "skw_scene_bounds": ([_P, _P], ctypes.c_int),
        "skw_scene_info": ([_P, _P], ctypes.c_int),
        "skw_scene_export_tree": ([_P, _P, _P, _P, _P, _P], ctypes.c_int),
        "skw_closest_hit": ([_P, _P, _U64, _P, _P], ctypes.c_int),
        "skw_closest_hit_stats": ([_P, _P, _U64, _P, _P, _P, _P], ctypes.c_int),
        "skw_closest_hit_noalloc": ([_P, _P, _U64, _P, _P], ctypes.c_int),
        "skw_any_hit": ([_P, _P, _U64, _P, _P], ctypes.c_int),
        "skw_hit_geometry": ([_P, _P, _P, _U64, _P, _P], ctypes.c_int),
        "skw_photon_map_create": ([_U64, _P, _P, _P, _P, _U32, ctypes.POINTER(_P)], ctypes.c_int),
        "skw_photon_map_destroy": ([_P], ctypes.c_int),
        "skw_photon_map_export": ([_P, _P, _P, _P], ctypes.c_int),
        "skw_photon_map_info": ([_P, _P, _P], ctypes.c_int),
        "skw_photon_irradiance": ([_P, _P, _P, _U64, _U32, ctypes.c_double, _U64, ctypes.c_double,
                                   _P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
        "skw_photon_irradiance_stats": ([_P, _P, _P, _U64, _U32, ctypes.c_double, _U64, ctypes.c_double,
                                         _P, _P, _P, _P, _P], ctypes.c_int),
        "skw_photon_knn": ([_P, _P, _P, _U64, _U32, ctypes.c_double, ctypes.c_int, _P, _P, _P, _P],
                           ctypes.c_int),
        "skw_photon_nearest": ([_P, _P, _P, _U64, ctypes.c_double, ctypes.c_int, _P, _P], ctypes.c_int),
        "skw_photon_cone_irradiance": ([_P, _P, _P, _U64, _U32, ctypes.c_double, _U64, ctypes.c_double, _P, _P],
                                       ctypes.c_int),
        "skw_photon_caustic_radiance": ([_P, _P, _P, _P, _U64, _U32, ctypes.c_double, _U64, ctypes.c_double, _P,
                                         _P], ctypes.c_int),
        "skw_last_error": ([], ctypes.c_char_p),
        "skw_device_count": ([], ctypes.c_int),
        "skw_sort_pairs_u32": ([_P, _P, _P, _P, _U64, ctypes.c_int, ctypes.c_int, _P], ctypes.c_int),
        "skw_dl_scene_create": ([_P, ctypes.POINTER(DLDesc), ctypes.POINTER(_P)], ctypes.c_int),
        "skw_dl_scene_destroy": ([_P], ctypes.c_int),
        "skw_dl_param_count": ([_P, _P], ctypes.c_int),
        "skw_dl_render_profiled": ([_P, _P, _P, _P, _P, ctypes.c_int, _P, _P], ctypes.c_int),
        "skw_dl_render_replay": ([_P, _P, _P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
        "skw_dl_adjoint": ([_P, _P, _P, _P, _P, _P, _P, _P, _P], ctypes.c_int),
        "skw_dl_render_tape": ([_P, _P, _P, _P, _P, _P, _U64, ctypes.POINTER(_P), _P, _P], ctypes.c_int),
        "skw_dl_tape_adjoint": ([_P, _P, _P, _P], ctypes.c_int),
        "skw_dl_tape_destroy": ([_P], ctypes.c_int),
        "skw_dl_render": ([_P, _P, ctypes.POINTER(Camera), ctypes.POINTER(DLParams), _P, _P, _P], ctypes.c_int),
        "skw_batch_create": ([_P, _P, _U32, ctypes.POINTER(_P)], ctypes.c_int),
        "skw_batch_create_with_gathers": ([_P, _P, _U32, _P, _U32, ctypes.POINTER(_P)], ctypes.c_int),
        "skw_batch_destroy": ([_P], ctypes.c_int),
        "skw_batch_configure": ([_P, ctypes.POINTER(BatchConfig)], ctypes.c_int),
        "skw_batch_autotune": ([_P, _U32, ctypes.POINTER(BatchConfig), ctypes.POINTER(ctypes.c_double)],
                               ctypes.c_int),
        "skw_batch_frame": ([_P, _P], ctypes.c_int),
        "skw_batch_hits": ([_P, _U32, ctypes.POINTER(_P), ctypes.POINTER(_U64)], ctypes.c_int),
        "skw_batch_latency": ([_P, _P], ctypes.c_int),
        "skw_batch_check": ([_P], ctypes.c_int),
        "skw_batch_rays": ([_P, _U32, _P, _P], ctypes.c_int),
        "skw_batch_ray_count": ([_P, ctypes.POINTER(_U64)], ctypes.c_int),
        "skw_irradiance_map_create": ([ctypes.POINTER(IrrMapDesc), ctypes.POINTER(_P), _P], ctypes.c_int),
        "skw_irradiance_map_create_from": ([_U64, _P, _P, _P, _P, ctypes.c_double, ctypes.POINTER(_P)],
                                           ctypes.c_int),
        "skw_irradiance_map_destroy": ([_P], ctypes.c_int),
        "skw_irradiance_map_info": ([_P, _P, _P], ctypes.c_int),
        "skw_irradiance_map_export": ([_P, _P, _P, _P, _P], ctypes.c_int),
        "skw_photon_shoot": ([_P, _P, ctypes.POINTER(ShootParams), ctypes.POINTER(_P), _P], ctypes.c_int),
        "skw_photon_set_info": ([_P, _P, _P, _P], ctypes.c_int),
        "skw_photon_set_export": ([_P, _U32, _P, _P, _P, _P], ctypes.c_int),
        "skw_photon_set_map": ([_P, _U32, ctypes.POINTER(_P), _P], ctypes.c_int),
        "skw_photon_map_build": ([_U64, _P, _P, _P, _P, ctypes.POINTER(_P), _P], ctypes.c_int),
        "skw_scene_save": ([_P, ctypes.c_char_p], ctypes.c_int),
        "skw_scene_load": ([ctypes.c_char_p, _U32, ctypes.POINTER(_P)], ctypes.c_int),
        "skw_photon_map_save": ([_P, ctypes.c_char_p], ctypes.c_int),
        "skw_photon_map_load": ([ctypes.c_char_p, _U32, ctypes.POINTER(_P)], ctypes.c_int),
        "skw_irradiance_map_save": ([_P, ctypes.c_char_p], ctypes.c_int),
        "skw_irradiance_map_load": ([ctypes.c_char_p, ctypes.POINTER(_P)], ctypes.c_int),
        "skw_photon_set_destroy": ([_P], ctypes.c_int),
        "skw_pbrt_load": ([ctypes.c_char_p, ctypes.POINTER(_P)], ctypes.c_int),
        "skw_pbrt_descs": ([_P, ctypes.POINTER(SceneDesc), ctypes.POINTER(DLDesc), ctypes.POINTER(Camera),
                            ctypes.POINTER(_U32), ctypes.POINTER(_U32)], ctypes.c_int),
        "skw_pbrt_warnings": ([_P, ctypes.POINTER(_U32)], ctypes.c_int),
        "skw_pbrt_warning": ([_P, _U32, ctypes.POINTER(ctypes.c_char_p)], ctypes.c_int),
        "skw_pbrt_destroy": ([_P], ctypes.c_int),
        "skw_dl_final_gather": ([_P, _P, _P, _P, ctypes.c_double, ctypes.POINTER(Camera), ctypes.POINTER(FGParams),
                                 _P, _P, _P, _P], ctypes.c_int),
        "skw_pm_render": ([_P, _P, _P, _P, _P, ctypes.c_double, ctypes.POINTER(Camera), ctypes.POINTER(PMParams),
                           _P, _P, _P], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


class SkwError(RuntimeError):
    pass


def check(status, what=""):
    if status != 0:
        msg = lib().skw_last_error().decode(errors="replace")
        raise SkwError("%s failed (status %d): %s" % (what, status, msg))


def sort_pairs_u32(d_keys_in, d_keys_out, d_vals_in, d_vals_out, n, begin_bit=0, end_bit=32, stream=0):
    """Stable sort of n (key, value) uint32 pairs (device pointers) by key bits [begin_bit, end_bit)."""
    check(lib().skw_sort_pairs_u32(d_keys_in, d_keys_out, d_vals_in, d_vals_out, n, begin_bit, end_bit,
                                   stream or None), "skw_sort_pairs_u32")


def _ptr(a):
    return None if a is None else a.ctypes.data


MESH_INVERT_NORMALS = 1
MESH_SHADING_NORMALS = 2


def mesh_arrays(meshes):
    """Flattens a list of mesh dicts into the arrays of skw_scene_desc (also used for the oracle)."""
    vo = np.zeros(len(meshes) + 1, np.uint64)
    to = np.zeros(len(meshes) + 1, np.uint64)
    for i, m in enumerate(meshes):
        vo[i + 1] = vo[i] + len(m["verts"])
        to[i + 1] = to[i] + len(m["tris"])
    verts = np.ascontiguousarray(np.concatenate([np.asarray(m["verts"], np.float32).reshape(-1, 3) for m in meshes])
                                 if meshes else np.zeros((0, 3), np.float32))
    tris = np.ascontiguousarray(np.concatenate([np.asarray(m["tris"], np.uint32).reshape(-1, 3) for m in meshes])
                                if meshes else np.zeros((0, 3), np.uint32))
    flags = np.zeros(len(meshes), np.uint8)
    normals = None
    if any(m.get("normals") is not None for m in meshes):
        parts = []
        for i, m in enumerate(meshes):
            if m.get("normals") is not None:
                parts.append(np.asarray(m["normals"], np.float32).reshape(-1, 3))
                flags[i] |= MESH_SHADING_NORMALS
            else:
                parts.append(np.zeros((len(m["verts"]), 3), np.float32))
        normals = np.ascontiguousarray(np.concatenate(parts))
    for i, m in enumerate(meshes):
        if m.get("invert_normals"):
            flags[i] |= MESH_INVERT_NORMALS
    return {"vert_off": vo, "tri_off": to, "verts": verts, "tris": tris, "normals": normals, "flags": flags}


def scene_desc(meshes, prims=None):
    """(SceneDesc, arrays it points into) for a list of mesh dicts and (mesh, 4x4 or None) primitives
    (None: one identity primitive per mesh); keep the arrays alive while the description is used."""
    if prims is None:
        prims = [(i, None) for i in range(len(meshes))]
    arrays = mesh_arrays(meshes)
    pm = np.array([p[0] for p in prims], np.uint32)
    xf = None
    if any(p[1] is not None for p in prims):
        xf = np.zeros((len(prims), 32), np.float64)
        for i, (_, m) in enumerate(prims):
            m = np.eye(4) if m is None else np.asarray(m, np.float64)
            xf[i, :16] = m.reshape(-1)
            xf[i, 16:] = np.linalg.inv(m).reshape(-1)
    desc = SceneDesc(len(meshes), _ptr(arrays["vert_off"]), _ptr(arrays["verts"]), _ptr(arrays["normals"]),
                     _ptr(arrays["tri_off"]), _ptr(arrays["tris"]), _ptr(arrays["flags"]), len(prims), _ptr(pm),
                     _ptr(xf))
    return desc, (arrays, pm, xf)


class Scene:
    """A scene handle: the ternary triangle tree of TriangleAccelerator, built on the host and
    (unless host_only) uploaded to the current HIP device.

    meshes: list of dicts with 'verts' (V,3) float32, 'tris' (T,3) uint32, optional 'normals'
            (V,3) float32 (shading normals, used when present) and 'invert_normals' bool.
    prims:  list of (mesh index, 4x4 mesh_to_world or None). None for all prims = identity.
    device_build: build the tree on the device (SKW_SCENE_DEVICE_BUILD), node for node the host
            builder's tree; scenes with instanced meshes raise (SKW_UNSUPPORTED).
    """

    def __init__(self, meshes, prims=None, host_only=False, device_build=False):
        L = lib()
        desc, self._keep = scene_desc(meshes, prims)
        h = _P()
        flags = (SCENE_HOST_ONLY if host_only else 0) | (SCENE_DEVICE_BUILD if device_build else 0)
        check(L.skw_scene_create(ctypes.byref(desc), flags, ctypes.byref(h)), "skw_scene_create")
        self.handle = h

    @classmethod
    def from_desc(cls, desc, host_only=False, device_build=False):
        """A scene from a filled SceneDesc whose arrays the caller keeps alive for the call."""
        h = _P()
        flags = (SCENE_HOST_ONLY if host_only else 0) | (SCENE_DEVICE_BUILD if device_build else 0)
        check(lib().skw_scene_create(ctypes.byref(desc), flags, ctypes.byref(h)), "skw_scene_create")
        sc = cls.__new__(cls)
        sc.handle = h
        return sc

    def save(self, path):
        """skw_scene_save: the built scene (tree, triangles, meshes, primitives) to a file."""
        check(lib().skw_scene_save(self.handle, os.fsencode(path)), "skw_scene_save")

    @classmethod
    def load(cls, path, host_only=False):
        """skw_scene_load: a saved scene, without rebuilding its tree."""
        h = _P()
        check(lib().skw_scene_load(os.fsencode(path), SCENE_HOST_ONLY if host_only else 0, ctypes.byref(h)),
              "skw_scene_load")
        sc = cls.__new__(cls)
        sc.handle = h
        return sc

    def close(self):
        if self.handle:
            lib().skw_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        out = np.zeros(6, np.uint64)
        check(lib().skw_scene_info(self.handle, out.ctypes.data), "skw_scene_info")
        keys = ["nodes", "triangles", "max_depth", "instances", "leaves", "stack_need"]
        return {k: int(v) for k, v in zip(keys, out)}

    def bounds(self):
        out = np.zeros(6, np.float64)
        check(lib().skw_scene_bounds(self.handle, out.ctypes.data), "skw_scene_bounds")
        return out

    def export_tree(self):
        info = self.info()
        n = info["nodes"]
        bbox = np.zeros((n, 6), np.float32)
        meta = np.zeros((n, 6), np.uint64)
        tp = np.zeros(info["triangles"], np.uint32)
        tt = np.zeros(info["triangles"], np.uint32)
        k = _U64()
        check(lib().skw_scene_export_tree(self.handle, bbox.ctypes.data, meta.ctypes.data, tp.ctypes.data,
                                          tt.ctypes.data, ctypes.byref(k)), "skw_scene_export_tree")
        return bbox[:k.value], meta[:k.value], tp, tt

    def closest_hit(self, d_rays, n, d_hits, stream=0):
        check(lib().skw_closest_hit(self.handle, d_rays, n, d_hits, stream or None), "skw_closest_hit")

    def closest_hit_stats(self, d_rays, n, d_hits, d_pops, d_tests, stream=0):
        check(lib().skw_closest_hit_stats(self.handle, d_rays, n, d_hits, d_pops, d_tests, stream or None),
              "skw_closest_hit_stats")

    def closest_hit_noalloc(self, d_rays, n, d_hits, stream=0):
        check(lib().skw_closest_hit_noalloc(self.handle, d_rays, n, d_hits, stream or None), "skw_closest_hit_noalloc")

    def any_hit(self, d_rays, n, d_occ, stream=0):
        check(lib().skw_any_hit(self.handle, d_rays, n, d_occ, stream or None), "skw_any_hit")

    def hit_geometry(self, d_rays, d_hits, n, d_geom, stream=0):
        check(lib().skw_hit_geometry(self.handle, d_rays, d_hits, n, d_geom, stream or None), "skw_hit_geometry")


class PhotonMap:
    """KDTree<Photon>: pos (N,3) float32, weight (N,3) float32, dir/normal CompressedDirection ids."""

    def __init__(self, pos, weight, dir_ids, normal_ids, host_only=False):
        L = lib()
        self._pos = np.ascontiguousarray(pos, np.float32)
        self._w = np.ascontiguousarray(weight, np.float32)
        self._d = np.ascontiguousarray(dir_ids, np.uint16)
        self._n = np.ascontiguousarray(normal_ids, np.uint16)
        self.n = len(self._pos)
        h = _P()
        check(L.skw_photon_map_create(self.n, _ptr(self._pos), _ptr(self._w), _ptr(self._d), _ptr(self._n),
                                      SCENE_HOST_ONLY if host_only else 0, ctypes.byref(h)), "skw_photon_map_create")
        self.handle = h

    def save(self, path):
        check(lib().skw_photon_map_save(self.handle, os.fsencode(path)), "skw_photon_map_save")

    @classmethod
    def load(cls, path, host_only=False):
        h = _P()
        check(lib().skw_photon_map_load(os.fsencode(path), SCENE_HOST_ONLY if host_only else 0, ctypes.byref(h)),
              "skw_photon_map_load")
        pm = cls.__new__(cls)
        pm.handle = h
        n, depth = ctypes.c_uint64(), ctypes.c_uint32()
        check(lib().skw_photon_map_info(h, ctypes.byref(n), ctypes.byref(depth)), "skw_photon_map_info")
        pm.n = n.value
        return pm

    @classmethod
    def build_on_device(cls, n, d_pos, d_weight=None, d_dir=None, d_normal=None, stream=0):
        """skw_photon_map_build: the kd-tree built on the device from device arrays."""
        h = _P()
        check(lib().skw_photon_map_build(n, d_pos, d_weight, d_dir, d_normal, ctypes.byref(h), stream or None),
              "skw_photon_map_build")
        pm = cls.__new__(cls)
        pm.handle = h
        pm.n = n
        return pm

    def close(self):
        if self.handle:
            lib().skw_photon_map_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def export(self):
        perm = np.zeros(self.n, np.uint32)
        split = np.zeros(self.n, np.float32)
        word = np.zeros(self.n, np.uint32)
        check(lib().skw_photon_map_export(self.handle, perm.ctypes.data, split.ctypes.data, word.ctypes.data),
              "skw_photon_map_export")
        return perm, split, word

    def irradiance(self, d_p, d_n, nq, k, r, paths, rho, d_eext, d_eint, d_rad=None, d_ids=None, d_d2=None,
                   d_count=None, stream=0):
        check(lib().skw_photon_irradiance(self.handle, d_p, d_n, nq, k, r, paths, rho, d_eext, d_eint, d_rad, d_ids,
                                          d_d2, d_count, stream or None), "skw_photon_irradiance")

    def irradiance_stats(self, d_p, d_n, nq, k, r, paths, rho, d_eext, d_eint, d_rad, d_stats, stream=0):
        """irradiance() with each lookup's work counted: d_stats gets (visits, photons inside the
        radius, set size) per query as uint32."""
        check(lib().skw_photon_irradiance_stats(self.handle, d_p, d_n, nq, k, r, paths, rho, d_eext, d_eint, d_rad,
                                                d_stats, stream or None), "skw_photon_irradiance_stats")

    def knn(self, d_p, d_n, nq, k, r, filt, d_ids, d_d2, d_count, stream=0):
        check(lib().skw_photon_knn(self.handle, d_p, d_n, nq, k, r, filt, d_ids, d_d2, d_count, stream or None),
              "skw_photon_knn")

    def cone_irradiance(self, d_p, d_n, nq, k, r, paths, kc, d_e, stream=0):
        check(lib().skw_photon_cone_irradiance(self.handle, d_p, d_n, nq, k, r, paths, kc, d_e, stream or None),
              "skw_photon_cone_irradiance")

    def caustic_radiance(self, d_p, d_n, d_wo, nq, k, r, paths, rho, d_rad, stream=0):
        check(lib().skw_photon_caustic_radiance(self.handle, d_p, d_n, d_wo, nq, k, r, paths, rho, d_rad,
                                                stream or None), "skw_photon_caustic_radiance")

    def nearest(self, d_p, d_n, nq, r, filt, d_id, stream=0):
        check(lib().skw_photon_nearest(self.handle, d_p, d_n, nq, r, filt, d_id, stream or None),
              "skw_photon_nearest")


class IrradianceMap:
    """The irradiance photon map of the photon integrator's final gather.

    IrradianceMap(direct=, indirect=, caustic=None, *_paths=, scene_area=) estimates it on the
    device from device PhotonMaps; IrradianceMap(explicit=(pos, normal_ids, e_ext, e_int,
    max_lookup_dist)) takes given irradiance photons."""

    def __init__(self, direct=None, indirect=None, caustic=None, direct_paths=1, indirect_paths=1, caustic_paths=1,
                 scene_area=0.0, explicit=None, stream=0):
        L = lib()
        h = _P()
        self._keep = (direct, indirect, caustic)
        if explicit is not None:
            pos, nid, ee, ei, md = explicit
            pos = np.ascontiguousarray(pos, np.float32)
            nid = np.ascontiguousarray(nid, np.uint16)
            ee = np.ascontiguousarray(ee, np.float32)
            ei = np.ascontiguousarray(ei, np.float32)
            check(L.skw_irradiance_map_create_from(len(nid), _ptr(pos), _ptr(nid), _ptr(ee), _ptr(ei), md,
                                                   ctypes.byref(h)), "skw_irradiance_map_create_from")
        else:
            desc = IrrMapDesc(caustic.handle if caustic is not None else None, direct.handle, indirect.handle,
                              caustic_paths, direct_paths, indirect_paths, scene_area)
            check(L.skw_irradiance_map_create(ctypes.byref(desc), ctypes.byref(h), stream or None),
                  "skw_irradiance_map_create")
        self.handle = h
        n = ctypes.c_uint64()
        md = ctypes.c_double()
        check(L.skw_irradiance_map_info(h, ctypes.byref(n), ctypes.byref(md)), "skw_irradiance_map_info")
        self.n = n.value
        self.max_lookup_dist = md.value

    def save(self, path):
        check(lib().skw_irradiance_map_save(self.handle, os.fsencode(path)), "skw_irradiance_map_save")

    @classmethod
    def load(cls, path):
        h = _P()
        check(lib().skw_irradiance_map_load(os.fsencode(path), ctypes.byref(h)), "skw_irradiance_map_load")
        m = cls.__new__(cls)
        m.handle = h
        m._keep = ()
        n = ctypes.c_uint64()
        md = ctypes.c_double()
        check(lib().skw_irradiance_map_info(h, ctypes.byref(n), ctypes.byref(md)), "skw_irradiance_map_info")
        m.n = n.value
        m.max_lookup_dist = md.value
        return m

    def close(self):
        if self.handle:
            lib().skw_irradiance_map_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def export(self):
        """(pos (n,3) float32, normal ids (n,) uint16, e_ext (n,3) float32, e_int (n,3) float32)."""
        pos = np.zeros((self.n, 3), np.float32)
        nid = np.zeros(self.n, np.uint16)
        ee = np.zeros((self.n, 3), np.float32)
        ei = np.zeros((self.n, 3), np.float32)
        check(lib().skw_irradiance_map_export(self.handle, pos.ctypes.data, nid.ctypes.data, ee.ctypes.data,
                                              ei.ctypes.data), "skw_irradiance_map_export")
        return pos, nid, ee, ei


class PhotonSet:
    """Photons shot by DLScene.shoot_photons: maps 0 caustic, 1 direct, 2 indirect."""

    def __init__(self, handle):
        self.handle = handle
        counts, paths, total = np.zeros(3, np.uint64), np.zeros(3, np.uint64), ctypes.c_uint64()
        check(lib().skw_photon_set_info(handle, counts.ctypes.data, paths.ctypes.data, ctypes.byref(total)),
              "skw_photon_set_info")
        self.counts = [int(x) for x in counts]
        self.paths = [int(x) for x in paths]
        self.total_paths = total.value

    def export(self, which):
        n = self.counts[which]
        m = {"pos": np.zeros((n, 3), np.float32), "weight": np.zeros((n, 3), np.float32),
             "dir": np.zeros(n, np.uint16), "normal": np.zeros(n, np.uint16)}
        check(lib().skw_photon_set_export(self.handle, which, m["pos"].ctypes.data, m["weight"].ctypes.data,
                                          m["dir"].ctypes.data, m["normal"].ctypes.data), "skw_photon_set_export")
        return m

    def photon_map(self, which):
        """The device PhotonMap (KDTree<Photon>) of one map."""
        h = _P()
        check(lib().skw_photon_set_map(self.handle, which, ctypes.byref(h), None), "skw_photon_set_map")
        pm = PhotonMap.__new__(PhotonMap)
        pm.handle = h
        pm.n = self.counts[which]
        return pm

    def close(self):
        if self.handle:
            lib().skw_photon_set_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _np_view(ptr, dtype, n):
    """Copy of n items of dtype at a host address (empty when n is 0)."""
    if n == 0 or not ptr:
        return np.zeros(0, dtype)
    buf = (ctypes.c_char * (n * np.dtype(dtype).itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype).copy()


class PbrtScene:
    """A pbrt-v2 scene file read by skw_pbrt_load (the subset the traced path uses; everything else
    is listed in `warnings`). scene() / dl_scene() build the handles straight from the import's
    descriptions; arrays() returns numpy copies of them."""

    def __init__(self, path):
        h = _P()
        check(lib().skw_pbrt_load(os.fsencode(path), ctypes.byref(h)), "skw_pbrt_load")
        self.handle = h

    def close(self):
        if self.handle:
            lib().skw_pbrt_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def warnings(self):
        n = _U32()
        check(lib().skw_pbrt_warnings(self.handle, ctypes.byref(n)), "skw_pbrt_warnings")
        out = []
        for i in range(n.value):
            msg = ctypes.c_char_p()
            check(lib().skw_pbrt_warning(self.handle, i, ctypes.byref(msg)), "skw_pbrt_warning")
            out.append(msg.value.decode(errors="replace"))
        return out

    def _descs(self, want_camera=False):
        sd, dd, cam, w, h = SceneDesc(), DLDesc(), Camera(), _U32(), _U32()
        check(lib().skw_pbrt_descs(self.handle, ctypes.byref(sd), ctypes.byref(dd),
                                   ctypes.byref(cam) if want_camera else None, ctypes.byref(w), ctypes.byref(h)),
              "skw_pbrt_descs")
        return sd, dd, cam, w.value, h.value

    def camera(self):
        """(Camera, width, height); raises SkwError when the file has no Camera."""
        _, _, cam, w, h = self._descs(True)
        return cam, w, h

    def scene(self, host_only=False):
        sd, _, _, _, _ = self._descs()
        return Scene.from_desc(sd, host_only)

    def dl_scene(self, scene):
        a = self.arrays()
        return DLScene(scene, a["mat"], a["kd"], a["ks"], a["rough"], a["light_prim"], a["light_radiance"])

    def arrays(self):
        """meshes (list of dicts as Scene takes), prims [(mesh, 4x4)], prim_xform (n, 32), mat, kd, ks,
        rough, light_prim, light_radiance, width, height."""
        sd, dd, _, w, h = self._descs()
        nm, npr = sd.n_meshes, sd.n_prims
        vo = _np_view(sd.mesh_vert_offset, np.uint64, nm + 1)
        to = _np_view(sd.mesh_tri_offset, np.uint64, nm + 1)
        nv, nt = int(vo[-1]) if nm else 0, int(to[-1]) if nm else 0
        verts = _np_view(sd.verts, np.float32, 3 * nv).reshape(-1, 3)
        normals = _np_view(sd.normals, np.float32, 3 * nv).reshape(-1, 3) if sd.normals else None
        tris = _np_view(sd.tris, np.uint32, 3 * nt).reshape(-1, 3)
        flags = _np_view(sd.mesh_flags, np.uint8, nm)
        meshes = []
        for i in range(nm):
            m = {"verts": verts[vo[i]:vo[i + 1]], "tris": tris[to[i]:to[i + 1]],
                 "invert_normals": bool(flags[i] & MESH_INVERT_NORMALS)}
            if flags[i] & MESH_SHADING_NORMALS:
                m["normals"] = normals[vo[i]:vo[i + 1]]
            meshes.append(m)
        pm = _np_view(sd.prim_mesh, np.uint32, npr)
        xf = _np_view(sd.prim_xform, np.float64, 32 * npr).reshape(-1, 32)
        nl = dd.n_lights
        return {"meshes": meshes, "prims": [(int(pm[i]), xf[i, :16].reshape(4, 4)) for i in range(npr)],
                "prim_xform": xf, "mat": _np_view(dd.mat_type, np.uint8, npr),
                "kd": _np_view(dd.kd, np.float64, 3 * npr).reshape(-1, 3),
                "ks": _np_view(dd.ks, np.float64, 3 * npr).reshape(-1, 3),
                "rough": _np_view(dd.roughness, np.float64, npr),
                "light_prim": _np_view(dd.light_prim, np.uint32, nl),
                "light_radiance": _np_view(dd.light_radiance, np.float64, 3 * nl).reshape(-1, 3),
                "width": w, "height": h}


def camera(origin, target_dir, up, x_fov_deg, width, height):
    """skw_camera for a pinhole looking along target_dir (look-at frame of Transform.cpp:138-165,
    tangents of PerspectiveCamera.cpp:20-34). Returns (Camera, cam15) where cam15 is the flat
    layout the oracle takes: origin, rotation (row-major), x_tan, y_tan."""
    d = np.asarray(target_dir, np.float64)
    d = d * (1.0 / np.sqrt(np.dot(d, d)))
    right = np.cross(d, np.asarray(up, np.float64))
    right = right * (1.0 / np.sqrt(np.dot(right, right)))
    new_up = np.cross(right, d)
    rot = np.stack([right, -new_up, d], 1)
    x_tan = float(np.tan(np.deg2rad(x_fov_deg) / 2.0))
    y_tan = x_tan * float(height) / float(width)
    c = Camera()
    for i in range(3):
        c.origin[i] = float(origin[i])
    for i in range(9):
        c.rot[i] = float(rot.reshape(-1)[i])
    c.x_tan, c.y_tan = x_tan, y_tan
    cam15 = np.concatenate([np.asarray(origin, np.float64), rot.reshape(-1), [x_tan, y_tan]])
    return c, cam15


class DLScene:
    """Materials and area lights for skw_dl_render over a Scene."""

    def __init__(self, scene, mat_type, kd, ks, roughness, light_prim, light_radiance, prim_group=None, n_groups=0):
        L = lib()
        self._mt = np.ascontiguousarray(mat_type, np.uint8)
        self._kd = np.ascontiguousarray(kd, np.float64)
        self._ks = np.ascontiguousarray(ks, np.float64)
        self._r = np.ascontiguousarray(roughness, np.float64)
        self._lp = np.ascontiguousarray(light_prim, np.uint32)
        self._lr = np.ascontiguousarray(light_radiance, np.float64)
        self._grp = None if prim_group is None else np.ascontiguousarray(prim_group, np.uint32)
        desc = DLDesc(_ptr(self._mt), _ptr(self._kd), _ptr(self._ks), _ptr(self._r), len(self._lp), _ptr(self._lp),
                      _ptr(self._lr), n_groups if self._grp is not None else 0, _ptr(self._grp))
        h = _P()
        check(L.skw_dl_scene_create(scene.handle, ctypes.byref(desc), ctypes.byref(h)), "skw_dl_scene_create")
        self.handle = h
        self.scene = scene
        n = ctypes.c_uint32()
        check(L.skw_dl_param_count(h, ctypes.byref(n)), "skw_dl_param_count")
        self.n_params = n.value

    def close(self):
        if self.handle:
            lib().skw_dl_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, cam, width, height, spp, d_rgb, light_samples=4, bsdf_samples=4, seed=0, stream=0,
               count_rays=True):
        """Renders into d_rgb. With count_rays the call waits for the frame and returns the
        primary / shadow / BSDF ray counts; without, it only enqueues the frame and returns None."""
        prm = DLParams(width, height, spp, light_samples, bsdf_samples, seed)
        rays = np.zeros(3, np.uint64) if count_rays else None
        check(lib().skw_dl_render(self.scene.handle, self.handle, ctypes.byref(cam), ctypes.byref(prm), d_rgb,
                                  rays.ctypes.data if count_rays else None, stream or None), "skw_dl_render")
        return rays

    def shoot_photons(self, max_paths, max_photons, paths_per_round=1 << 20, seed=0, stream=0):
        """skw_photon_shoot from this scene's area lights; max_photons = (caustic, direct, indirect)."""
        prm = ShootParams(max_paths, (_U64 * 3)(*max_photons), paths_per_round, seed)
        h = _P()
        check(lib().skw_photon_shoot(self.scene.handle, self.handle, ctypes.byref(prm), ctypes.byref(h),
                                     stream or None), "skw_photon_shoot")
        return PhotonSet(h)

    def final_gather(self, irr, indirect, scene_area, cam, width, height, spp, gather_samples, d_rgb=None,
                     d_radiance=None, seed=0, stream=0, count_rays=True):
        """skw_dl_final_gather: the photon integrator's final gather at every sample's camera hit.
        Returns (camera rays, gather rays) when count_rays, else only enqueues the frame."""
        prm = FGParams(width, height, spp, gather_samples, seed)
        rays = np.zeros(2, np.uint64) if count_rays else None
        check(lib().skw_dl_final_gather(self.scene.handle, self.handle, irr.handle,
                                        indirect.handle if indirect is not None else None, scene_area,
                                        ctypes.byref(cam), ctypes.byref(prm), d_rgb, d_radiance,
                                        rays.ctypes.data if count_rays else None, stream or None),
              "skw_dl_final_gather")
        return rays

    def pm_render(self, irr, caustic, indirect, scene_area, cam, width, height, spp, d_rgb, light_samples=4,
                  bsdf_samples=4, gather_samples=8, seed=0, stream=0, count_rays=True):
        """skw_pm_render: Le + direct lighting + caustic + final gather per sample. Returns (camera,
        shadow, BSDF, gather rays) when count_rays, else only enqueues the frame. irr may be None."""
        prm = PMParams(width, height, spp, light_samples, bsdf_samples, gather_samples, seed)
        rays = np.zeros(4, np.uint64) if count_rays else None
        check(lib().skw_pm_render(self.scene.handle, self.handle, irr.handle if irr is not None else None,
                                  caustic.handle if caustic is not None else None,
                                  indirect.handle if indirect is not None else None, scene_area, ctypes.byref(cam),
                                  ctypes.byref(prm), d_rgb, rays.ctypes.data if count_rays else None, stream or None),
              "skw_pm_render")
        return rays

    def render_profiled(self, cam, width, height, spp, d_rgb, light_samples=4, bsdf_samples=4, seed=0,
                        count_tests=False, stream=0):
        prm = DLParams(width, height, spp, light_samples, bsdf_samples, seed)
        prof = DLProfile()
        check(lib().skw_dl_render_profiled(self.scene.handle, self.handle, ctypes.byref(cam), ctypes.byref(prm),
                                           d_rgb, 1 if count_tests else 0, ctypes.byref(prof), stream or None),
              "skw_dl_render_profiled")
        return prof.as_dict()

    def _theta(self, theta):
        t = np.ascontiguousarray(theta, np.float64)
        if t.shape != (self.n_params,):
            raise ValueError("theta must hold %d values" % self.n_params)
        return t

    def render_replay(self, theta_eval, theta_sample, cam, width, height, spp, d_image, light_samples=4,
                      bsdf_samples=4, seed=0, stream=0, count_rays=True):
        """Replayed estimator F(theta_eval; theta_sample) into d_image (width*height*3 doubles).
        count_rays: as in render()."""
        te, ts = self._theta(theta_eval), self._theta(theta_sample)
        prm = DLParams(width, height, spp, light_samples, bsdf_samples, seed)
        rays = np.zeros(3, np.uint64) if count_rays else None
        check(lib().skw_dl_render_replay(self.scene.handle, self.handle, te.ctypes.data, ts.ctypes.data,
                                         ctypes.byref(cam), ctypes.byref(prm), d_image,
                                         rays.ctypes.data if count_rays else None, stream or None),
              "skw_dl_render_replay")
        return rays

    def render_tape(self, theta, cam, width, height, spp, d_image, max_tape_bytes=1 << 34, light_samples=4,
                    bsdf_samples=4, seed=0, stream=0, tape=None, count_rays=True):
        """Replayed render at theta into d_image, keeping per-sample derivatives: returns a DLTape
        (tape: an earlier DLTape of the same frame shape to overwrite). count_rays: as in render()."""
        th = self._theta(theta)
        prm = DLParams(width, height, spp, light_samples, bsdf_samples, seed)
        rays = np.zeros(3, np.uint64) if count_rays else None
        t = _P(tape.handle.value if tape is not None else None)
        check(lib().skw_dl_render_tape(self.scene.handle, self.handle, th.ctypes.data, ctypes.byref(cam),
                                       ctypes.byref(prm), d_image, max_tape_bytes, ctypes.byref(t),
                                       rays.ctypes.data if count_rays else None, stream or None),
              "skw_dl_render_tape")
        if tape is not None:
            tape.rays = rays
            return tape
        return DLTape(t, rays)

    def adjoint(self, theta, cam, width, height, spp, d_adjoint, d_grad, light_samples=4, bsdf_samples=4, seed=0,
                stream=0, count_rays=True):
        """d_grad (n_params doubles) = sum d_adjoint * dImage/dtheta, replaying the paths at theta.
        count_rays: as in render()."""
        th = self._theta(theta)
        prm = DLParams(width, height, spp, light_samples, bsdf_samples, seed)
        rays = np.zeros(3, np.uint64) if count_rays else None
        check(lib().skw_dl_adjoint(self.scene.handle, self.handle, th.ctypes.data, ctypes.byref(cam),
                                   ctypes.byref(prm), d_adjoint, d_grad, rays.ctypes.data if count_rays else None,
                                   stream or None), "skw_dl_adjoint")
        return rays


class DLTape:
    """Per-sample derivatives of one replayed frame (skw_dl_render_tape)."""

    def __init__(self, handle, rays):
        self.handle = handle
        self.rays = rays

    def adjoint(self, d_adjoint, d_grad, stream=0):
        check(lib().skw_dl_tape_adjoint(self.handle, d_adjoint, d_grad, stream or None), "skw_dl_tape_adjoint")

    def close(self):
        if self.handle:
            lib().skw_dl_tape_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many scenes traced in one persistent launch from a device work queue (skw_batch_*).
    views: list of (Camera, width, height, spp, seed), one per scene. gathers (optional): list of
    (PhotonMap, d_points, d_normals, nq, k, r, paths, rho, d_e_ext, d_e_int, d_radiance) jobs."""

    def __init__(self, scenes, views, gathers=()):
        L = lib()
        self.scenes = list(scenes)
        self.maps = [g[0] for g in gathers]
        n, m = len(self.scenes), len(gathers)
        hs = (_P * max(n, 1))(*[s.handle for s in self.scenes])
        vs = (BatchView * max(n, 1))()
        for i, (cam, w, h, spp, seed) in enumerate(views):
            vs[i].cam = cam
            vs[i].width, vs[i].height, vs[i].spp, vs[i].seed = w, h, spp, seed
        gs = (BatchGather * max(m, 1))()
        for j, (pm, dp, dn, nq, k, r, paths, rho, ee, ei, rad) in enumerate(gathers):
            gs[j] = BatchGather(pm.handle, dp, dn, nq, k, r, paths, rho, ee, ei, rad)
        self._hs, self._vs, self._gs = hs, vs, gs
        b = _P()
        check(L.skw_batch_create_with_gathers(hs, vs, n, gs, m, ctypes.byref(b)), "skw_batch_create_with_gathers")
        self.handle = b
        self.n = n + m

    def close(self):
        if self.handle:
            lib().skw_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, waves_per_block, stack_entries, chunk_rays, waves_per_cu):
        cfg = BatchConfig(waves_per_block, stack_entries, chunk_rays, waves_per_cu)
        check(lib().skw_batch_configure(self.handle, ctypes.byref(cfg)), "skw_batch_configure")

    def autotune(self, frames=3):
        cfg = BatchConfig()
        ms = ctypes.c_double()
        check(lib().skw_batch_autotune(self.handle, frames, ctypes.byref(cfg), ctypes.byref(ms)),
              "skw_batch_autotune")
        return cfg.as_dict(), ms.value

    def frame(self, stream=0):
        check(lib().skw_batch_frame(self.handle, stream or None), "skw_batch_frame")

    def hits(self, scene):
        p = _P()
        n = _U64()
        check(lib().skw_batch_hits(self.handle, scene, ctypes.byref(p), ctypes.byref(n)), "skw_batch_hits")
        return p.value, n.value

    def latency_ns(self):
        out = np.zeros(self.n, np.float64)
        check(lib().skw_batch_latency(self.handle, out.ctypes.data), "skw_batch_latency")
        return out

    def check(self):
        """Waits for the last frame; raises if a wave of it stalled and left early."""
        check(lib().skw_batch_check(self.handle), "skw_batch_check")

    def rays(self, scene, d_rays, stream=0):
        check(lib().skw_batch_rays(self.handle, scene, d_rays, stream or None), "skw_batch_rays")

    def ray_count(self):
        n = _U64()
        check(lib().skw_batch_ray_count(self.handle, ctypes.byref(n)), "skw_batch_ray_count")
        return n.value
