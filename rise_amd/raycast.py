"""Python mirror of RISE's object-manager ray queries over the C ABI.

``ObjectManager`` follows ``IObjectManager`` (Interfaces/IObjectManager.h:38-68
in the reference): meshes are added, ``prepare_for_rendering`` builds and
uploads the acceleration structure (ObjectManager::PrepareForRendering,
Managers/ObjectManager.cpp:404-423), and ``intersect_ray`` /
``intersect_shadow_ray`` answer closest-hit and any-hit queries, here for a
whole batch of rays resident in device memory.  Errors raise
:class:`rise_amd._lib.RiseHipError` (the reference returns bool and logs).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib


def _stream_handle(stream) -> int | None:
    if stream is None:
        import torch

        return torch.cuda.current_stream().cuda_stream
    return int(getattr(stream, "cuda_stream", stream))


ARITH_IEEE = 0
"""Hits as the reference's source computes them, every operation rounded."""
ARITH_X86_FMA = 1
"""Hits as the reference's Linux build (g++ -O3 -march=x86-64-v3) computes them."""


FACES_NONE, FACES_FRONT, FACES_BACK, FACES_BOTH = 0, 1, 2, 3


class Hit(ctypes.Structure):
    """rise_hip_hit."""
    _fields_ = [("prim_id", ctypes.c_int32), ("mesh", ctypes.c_int32), ("range", ctypes.c_double),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("point", ctypes.c_float * 3),
                ("geom_normal", ctypes.c_float * 3), ("shading_normal", ctypes.c_float * 3),
                ("uv", ctypes.c_float * 2)]


HIT_DTYPE = np.dtype([("prim_id", "<i4"), ("mesh", "<i4"), ("range", "<f8"), ("alpha", "<f8"), ("beta", "<f8"),
                      ("point", "<f4", (3,)), ("geom_normal", "<f4", (3,)), ("shading_normal", "<f4", (3,)),
                      ("uv", "<f4", (2,)), ("_pad", "<i4")])
assert HIT_DTYPE.itemsize == ctypes.sizeof(Hit) == 80


def hit_fields(records) -> np.ndarray:
    """rise_hip_hit records (bytes: a numpy or torch uint8 buffer) as a
    structured numpy array."""
    if hasattr(records, "cpu"):
        records = records.cpu().numpy()
    return np.frombuffer(np.ascontiguousarray(records).tobytes(), dtype=HIT_DTYPE)


OBJECT_DTYPE = np.dtype([("mesh", "<i4"), ("reserved", "<i4"), ("to_world", "<f8", (16,)),
                         ("to_object", "<f8", (16,))])
"""rise_hip_object_instance: mesh index, 0, and the object's two Matrix4s
(member order _00.._33, row-vector convention, translation in _30.._32)."""


def object_matrices(linear=None, translation=(0.0, 0.0, 0.0), to_object=None):
    """The Matrix4 pair (to_world, to_object) of an object, each 16 float64 in
    member order.  linear: 3x3 acting on row vectors (p' = p @ linear +
    translation), identity when None; to_object: its inverse when given (RISE
    computes it with Matrix4Ops::Inverse), else numpy's inverse."""
    m = np.eye(4)
    if linear is not None:
        m[:3, :3] = np.asarray(linear, dtype=np.float64)
    m[3, :3] = np.asarray(translation, dtype=np.float64)
    inv = np.linalg.inv(m) if to_object is None else np.asarray(to_object, dtype=np.float64).reshape(4, 4)
    return m.reshape(16).copy(), inv.reshape(16).copy()


def make_objects(entries) -> np.ndarray:
    """OBJECT_DTYPE array from (mesh, to_world, to_object) triples."""
    out = np.zeros(len(entries), dtype=OBJECT_DTYPE)
    for k, (mesh, tw, to) in enumerate(entries):
        out[k]["mesh"] = mesh
        out[k]["to_world"] = np.asarray(tw, dtype=np.float64).reshape(16)
        out[k]["to_object"] = np.asarray(to, dtype=np.float64).reshape(16)
    return out


class ObjectManager:
    """Resident triangle scene answering batched ray queries on the GPU.

    Meshes added with add_mesh alone are world-space geometry (identity
    objects).  Once add_object places a mesh, the scene is a list of objects
    (rise_hip_scene_upload_objects): each object is one mesh with its
    transform, intersected in its own frame as Object::IntersectRay does.

    arith selects the reference build whose hits are reproduced bit for bit
    (ARITH_IEEE or ARITH_X86_FMA; rise_hip_scene_set_arithmetic)."""

    def __init__(self, arith: int = ARITH_IEEE) -> None:
        self._meshes: list[tuple] = []  # (triangles, ids, double-sided, normals, coordinates)
        self._objects: list[tuple[int, np.ndarray, np.ndarray]] = []
        self._scene = ctypes.c_void_p(None)
        self._arith = arith
        self._from_cache = False

    def add_object(self, mesh: int, to_world=None, to_object=None) -> int:
        """Places mesh `mesh` (an add_mesh index) by the Matrix4 to_world (16
        float64 in member order, or 4x4; identity when None) with inverse
        to_object (numpy's inverse when None).  Returns the object's index,
        which hit records report in their `mesh` field."""
        if not 0 <= mesh < len(self._meshes):
            raise ValueError("mesh index out of range")
        if to_world is None:
            tw, to = object_matrices()
        else:
            tw = np.asarray(to_world, dtype=np.float64).reshape(16).copy()
            to = (np.linalg.inv(tw.reshape(4, 4)).reshape(16) if to_object is None
                  else np.asarray(to_object, dtype=np.float64).reshape(16).copy())
        self._objects.append((int(mesh), tw, to))
        return len(self._objects) - 1

    def add_mesh(self, tris: np.ndarray, ids: np.ndarray | None = None, double_sided: bool = False,
                 normals: np.ndarray | None = None, coords: np.ndarray | None = None) -> None:
        """tris: (n, 9) float32 world-space triangles, or float64 for a mesh
        whose double vertices the certifier uses (rise_hip_mesh_desc.
        vertices_f64); ids: optional (n,) int32; double_sided: the mesh is
        double-sided (RISE_HIP_MESH_DOUBLE_SIDED); normals (n, 9) / coords
        (n, 6): per-corner vertex normals and texture coordinates (float64)
        for the hit records."""
        if self._from_cache:
            raise RuntimeError("a scene loaded from a cache is complete; it takes no more meshes")
        f64 = np.asarray(tris).dtype == np.float64
        t = np.ascontiguousarray(tris, dtype=np.float64 if f64 else np.float32).reshape(-1, 9)
        i = None if ids is None else np.ascontiguousarray(ids, dtype=np.int32).reshape(-1)
        if i is not None and i.shape[0] != t.shape[0]:
            raise ValueError("ids must have one entry per triangle")
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float64).reshape(t.shape[0], 9)
        crd = None if coords is None else np.ascontiguousarray(coords, np.float64).reshape(t.shape[0], 6)
        self._meshes.append((t, i, bool(double_sided), nrm, crd))

    def prepare_for_rendering(self) -> None:
        if self._from_cache:
            raise RuntimeError("a scene loaded from a cache is already prepared")
        L = _lib.lib()
        self.release()
        descs = (_lib.MeshDesc * max(1, len(self._meshes)))()
        for k, (t, i, ds, nrm, crd) in enumerate(self._meshes):
            _set_desc(descs[k], t, i, ds, nrm, crd)
        out = ctypes.c_void_p(None)
        if self._objects:
            objs = make_objects(self._objects)
            _lib.check(L.rise_hip_scene_upload_objects(descs, len(self._meshes), objs.ctypes.data, len(objs),
                                                       ctypes.byref(out)))
        else:
            _lib.check(L.rise_hip_scene_upload(descs, len(self._meshes), ctypes.byref(out)))
        self._scene = out
        self.set_arithmetic(self._arith)

    def set_arithmetic(self, arith: int) -> None:
        """ARITH_IEEE or ARITH_X86_FMA for the queries that follow."""
        self._arith = arith
        if self._scene.value:
            _lib.check(_lib.lib().rise_hip_scene_set_arithmetic(self._scene, int(arith)))

    @property
    def handle(self) -> ctypes.c_void_p:
        if not self._scene.value:
            raise RuntimeError("prepare_for_rendering() has not been called")
        return self._scene

    def stats(self) -> dict:
        s = _lib.SceneStats()
        _lib.check(_lib.lib().rise_hip_scene_get_stats(self.handle, ctypes.byref(s)))
        return s.as_dict()

    def intersect_ray(self, rays, prim_id=None, t=None, stream=None, counters=None):
        """rays: (6, n) float32 CUDA tensor.  Returns (prim_id int32, t float32)."""
        import torch

        n = rays.shape[1]
        _check_rays(rays)
        if prim_id is None:
            prim_id = torch.empty(n, dtype=torch.int32, device=rays.device)
        if t is None:
            t = torch.empty(n, dtype=torch.float32, device=rays.device)
        L = _lib.lib()
        if counters is None:
            st = L.rise_hip_intersect_closest(
                self.handle, rays.data_ptr(), n, prim_id.data_ptr(), t.data_ptr(), _stream_handle(stream)
            )
        else:
            st = L.rise_hip_intersect_closest_counted(
                self.handle, rays.data_ptr(), n, prim_id.data_ptr(), t.data_ptr(), counters.data_ptr(),
                _stream_handle(stream),
            )
        _lib.check(st)
        return prim_id, t

    def intersect_shadow_ray(self, rays, how_far=None, occluded=None, stream=None, counters=None):
        """Any-hit within how_far (world distance, (n,) float32 tensor or None)."""
        import torch

        n = rays.shape[1]
        _check_rays(rays)
        if occluded is None:
            occluded = torch.empty(n, dtype=torch.uint8, device=rays.device)
        hf = 0 if how_far is None else how_far.data_ptr()
        L = _lib.lib()
        if counters is None:
            st = L.rise_hip_intersect_any(self.handle, rays.data_ptr(), hf, n, occluded.data_ptr(), _stream_handle(stream))
        else:
            st = L.rise_hip_intersect_any_counted(
                self.handle, rays.data_ptr(), hf, n, occluded.data_ptr(), counters.data_ptr(), _stream_handle(stream)
            )
        _lib.check(st)
        return occluded

    def intersect_ray_ex(self, rays, faces: int = FACES_BOTH, hits=None, stream=None):
        """rise_hip_intersect_closest_ex: closest hit under a face selection
        (FACES_*) with the intersection record.  Returns a (n,) uint8 CUDA
        tensor view of rise_hip_hit records; hit_fields() splits it."""
        import torch

        n = rays.shape[1]
        _check_rays(rays)
        if hits is None:
            hits = torch.empty(n * ctypes.sizeof(Hit), dtype=torch.uint8, device=rays.device)
        _lib.check(_lib.lib().rise_hip_intersect_closest_ex(self.handle, rays.data_ptr(), n, int(faces),
                                                            hits.data_ptr(), _stream_handle(stream)))
        return hits

    def intersect_shadow_ray_ex(self, rays, faces: int = FACES_BOTH, how_far=None, occluded=None, stream=None):
        """rise_hip_intersect_any_ex: any hit within how_far under a face selection."""
        import torch

        n = rays.shape[1]
        _check_rays(rays)
        if occluded is None:
            occluded = torch.empty(n, dtype=torch.uint8, device=rays.device)
        hf = 0 if how_far is None else how_far.data_ptr()
        _lib.check(_lib.lib().rise_hip_intersect_any_ex(self.handle, rays.data_ptr(), hf, n, int(faces),
                                                        occluded.data_ptr(), _stream_handle(stream)))
        return occluded

    def save_cache(self, path) -> None:
        """Saves the built 8-wide tree (rise_hip_scene_cache_write)."""
        L = _lib.lib()
        L.rise_hip_scene_cache_write.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        _lib.check(L.rise_hip_scene_cache_write(self.handle, os.fsencode(path)))

    @classmethod
    def load_cache(cls, path, arith: int = ARITH_IEEE) -> "ObjectManager":
        """A resident scene from a saved tree, with no build (rise_hip_scene_cache_read).
        The scene is complete: add_mesh and prepare_for_rendering raise."""
        L = _lib.lib()
        L.rise_hip_scene_cache_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]
        om = cls(arith)
        out = ctypes.c_void_p(None)
        _lib.check(L.rise_hip_scene_cache_read(os.fsencode(path), ctypes.byref(out)))
        om._scene = out
        om._from_cache = True
        om.set_arithmetic(arith)
        return om

    def release(self) -> None:
        if self._scene.value:
            _lib.lib().rise_hip_scene_release(self._scene)
            self._scene = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def _set_desc(desc, t: np.ndarray, ids, double_sided: bool = False, normals=None, coords=None) -> None:
    """Fills a rise_hip_mesh_desc from float32 or float64 (double vertices) triangles."""
    desc.flags = 1 if double_sided else 0
    desc.normals = None if normals is None else normals.ctypes.data
    desc.coords = None if coords is None else coords.ctypes.data
    if t.dtype == np.float64:
        desc.vertices = None
        desc.vertices_f64 = t.ctypes.data
    else:
        desc.vertices = t.ctypes.data
        desc.vertices_f64 = None
    desc.n_tris = t.shape[0]
    desc.tri_ids = None if ids is None else ids.ctypes.data


def _check_rays(rays) -> None:
    import torch

    if rays.dtype != torch.float32 or rays.dim() != 2 or rays.shape[0] != 6 or not rays.is_contiguous():
        raise ValueError("rays must be a contiguous (6, n) float32 tensor")
    if not rays.is_cuda:
        raise ValueError("rays must be in device memory")


def host_trace(tris: np.ndarray, rays: np.ndarray, mode: int = 0, tmax: np.ndarray | None = None,
               arith: int = ARITH_IEEE):
    """Runs the kernels' traversal code on host threads (validation only).
    Returns (prim_id, t, counters, stats) for mode 0, (occluded, counters, stats)
    for mode 1."""
    mode = (mode & 1) | (2 if arith == ARITH_X86_FMA else 0)
    L = _lib.lib()
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    n = rays.shape[1]
    counters = np.zeros(2, dtype=np.uint64)
    stats = _lib.SceneStats()
    tm = None if tmax is None else np.ascontiguousarray(tmax, dtype=np.float32)
    if (mode & 1) == 0:
        pid = np.empty(n, dtype=np.int32)
        t = np.empty(n, dtype=np.float32)
        _lib.check(
            L.rise_hip_debug_trace_host(
                tris.ctypes.data, tris.shape[0], rays.ctypes.data, None, n, mode, pid.ctypes.data, t.ctypes.data,
                None, counters.ctypes.data, ctypes.byref(stats),
            )
        )
        return pid, t, counters, stats.as_dict()
    occ = np.empty(n, dtype=np.uint8)
    _lib.check(
        L.rise_hip_debug_trace_host(
            tris.ctypes.data, tris.shape[0], rays.ctypes.data, None if tm is None else tm.ctypes.data, n, mode,
            None, None, occ.ctypes.data, counters.ctypes.data, ctypes.byref(stats),
        )
    )
    return occ, counters, stats.as_dict()


def host_trace_ex(tris: np.ndarray, rays: np.ndarray, mode: int = 0, faces: int = FACES_BOTH,
                  tmax: np.ndarray | None = None, arith: int = ARITH_IEEE):
    """rise_hip_debug_trace_ex_host: the kernels' traversal with a face
    selection on host threads (validation only).  Mode 0: structured hit
    records (HIT_DTYPE); mode 1: occlusion bytes."""
    mode = (mode & 1) | (2 if arith == ARITH_X86_FMA else 0)
    L = _lib.lib()
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    n = rays.shape[1]
    tm = None if tmax is None else np.ascontiguousarray(tmax, dtype=np.float32)
    if (mode & 1) == 0:
        out = np.zeros(n, dtype=HIT_DTYPE)
        _lib.check(L.rise_hip_debug_trace_ex_host(tris.ctypes.data, tris.shape[0], rays.ctypes.data, None, n, mode,
                                                  int(faces), out.ctypes.data, None))
        return out
    occ = np.empty(n, dtype=np.uint8)
    _lib.check(L.rise_hip_debug_trace_ex_host(tris.ctypes.data, tris.shape[0], rays.ctypes.data,
                                              None if tm is None else tm.ctypes.data, n, mode, int(faces), None,
                                              occ.ctypes.data))
    return occ


def host_trace_objects(meshes, objects, rays: np.ndarray, mode: int = 0, faces: int = FACES_BOTH,
                       tmax: np.ndarray | None = None, arith: int = ARITH_IEEE, ids=None, double_sided=None,
                       normals=None, coords=None):
    """rise_hip_debug_objects_host: a scene of objects (meshes: list of (n, 9)
    float32 arrays, or float64 for double vertices; objects: OBJECT_DTYPE array) queried with the kernels'
    per-ray code on host threads (validation only).  Mode 0: hit records
    (HIT_DTYPE, `mesh` = object index); mode 1: occlusion bytes."""
    mode = (mode & 1) | (2 if arith == ARITH_X86_FMA else 0)
    L = _lib.lib()
    arrs = [np.ascontiguousarray(m, dtype=np.float64 if np.asarray(m).dtype == np.float64 else np.float32)
            .reshape(-1, 9) for m in meshes]
    id_arrs = [None if ids is None or ids[k] is None else np.ascontiguousarray(ids[k], np.int32)
               for k in range(len(arrs))]
    descs = (_lib.MeshDesc * max(1, len(arrs)))()
    nrms = [None if normals is None or normals[k] is None else
            np.ascontiguousarray(normals[k], np.float64).reshape(-1, 9) for k in range(len(arrs))]
    crds = [None if coords is None or coords[k] is None else
            np.ascontiguousarray(coords[k], np.float64).reshape(-1, 6) for k in range(len(arrs))]
    for k, t in enumerate(arrs):
        _set_desc(descs[k], t, id_arrs[k], bool(double_sided[k]) if double_sided is not None else False,
                  nrms[k], crds[k])
    objs = np.ascontiguousarray(objects, dtype=OBJECT_DTYPE)
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    n = rays.shape[1]
    tm = None if tmax is None else np.ascontiguousarray(tmax, dtype=np.float32)
    if (mode & 1) == 0:
        out = np.zeros(n, dtype=HIT_DTYPE)
        _lib.check(L.rise_hip_debug_objects_host(descs, len(arrs), objs.ctypes.data, len(objs), rays.ctypes.data,
                                                 None, n, mode, int(faces), out.ctypes.data, None))
        return out
    occ = np.empty(n, dtype=np.uint8)
    _lib.check(L.rise_hip_debug_objects_host(descs, len(arrs), objs.ctypes.data, len(objs), rays.ctypes.data,
                                             None if tm is None else tm.ctypes.data, n, mode, int(faces), None,
                                             occ.ctypes.data))
    return occ
