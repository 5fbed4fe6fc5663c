"""RISE's binary mesh files (.risemesh v1-v5, with the BVH2 cache block) over
the C ABI of include/rise_hip_risemesh.h.

``RiseMesh(path)`` reads a file (TriangleMeshGeometryIndexed::Deserialize,
Geometry/TriangleMeshGeometryIndexed.cpp:758-1000 in the reference);
``object_manager()`` makes its triangles a resident ray-cast scene whose
8-wide tree comes from the file's BVH2 cache when it is valid, as the
reference skips its SAH rebuild.  ``write_risemesh`` writes version 5
(TriangleMeshGeometryIndexed::Serialize, :606-747) with a BVH2 block the
reference loads.  Errors raise RiseHipError.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from .raycast import ObjectManager


class RiseMeshInfo(ctypes.Structure):
    _fields_ = [
        ("version", ctypes.c_int32),
        ("face_normals", ctypes.c_int32),
        ("double_sided", ctypes.c_int32),
        ("has_bvh_cache", ctypes.c_int32),
        ("bvh_cache_valid", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("n_points", ctypes.c_int64),
        ("n_normals", ctypes.c_int64),
        ("n_coords", ctypes.c_int64),
        ("n_colors", ctypes.c_int64),
        ("n_tris", ctypes.c_int64),
        ("n_bvh_nodes", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_P = ctypes.c_void_p
_SIG = {
    "rise_hip_risemesh_read": ([ctypes.c_char_p, ctypes.POINTER(_P)], ctypes.c_int),
    "rise_hip_risemesh_get_info": ([_P, ctypes.POINTER(RiseMeshInfo)], ctypes.c_int),
    "rise_hip_risemesh_get_triangles": ([_P, _P], ctypes.c_int),
    "rise_hip_risemesh_write": ([ctypes.c_char_p, _P, ctypes.c_int64, _P, ctypes.c_int32, ctypes.c_int32],
                                ctypes.c_int),
    "rise_hip_risemesh_upload": ([_P, ctypes.POINTER(_P)], ctypes.c_int),
    "rise_hip_risemesh_release": ([_P], ctypes.c_int),
    "rise_hip_debug_trace_host_risemesh": ([_P, _P, _P, ctypes.c_int64, ctypes.c_int32, _P, _P, _P, _P,
                                            ctypes.POINTER(_lib.SceneStats)], ctypes.c_int),
}


def _lib_risemesh():
    L = _lib.lib()
    for name, (a, r) in _SIG.items():
        fn = getattr(L, name)
        fn.argtypes = a
        fn.restype = r
    return L


class RiseMesh:
    """A .risemesh file read into host memory."""

    def __init__(self, path: str | os.PathLike):
        h = _P(None)
        _lib.check(_lib_risemesh().rise_hip_risemesh_read(os.fsencode(path), ctypes.byref(h)))
        self._h = h

    def info(self) -> dict:
        out = RiseMeshInfo()
        _lib.check(_lib_risemesh().rise_hip_risemesh_get_info(self._h, ctypes.byref(out)))
        return out.as_dict()

    def triangles(self) -> np.ndarray:
        """(n, 9) float32 triangles in file order (ids 0..n-1)."""
        n = self.info()["n_tris"]
        out = np.empty((n, 9), np.float32)
        _lib.check(_lib_risemesh().rise_hip_risemesh_get_triangles(self._h, out.ctypes.data))
        return out

    def object_manager(self) -> ObjectManager:
        """A resident ray-cast scene of the mesh (GPU)."""
        out = _P(None)
        _lib.check(_lib_risemesh().rise_hip_risemesh_upload(self._h, ctypes.byref(out)))
        om = ObjectManager()
        om._scene = out
        return om

    def trace_host(self, rays: np.ndarray, mode: int = 0, tmax: np.ndarray | None = None):
        """The kernels' traversal on host threads over the mesh's tree (no GPU):
        (prim_id, t, stats) for mode 0, (occluded, stats) for mode 1."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[1]
        st = _lib.SceneStats()
        tm = None if tmax is None else np.ascontiguousarray(tmax, np.float32)
        L = _lib_risemesh()
        if mode == 0:
            pid = np.empty(n, np.int32)
            t = np.empty(n, np.float32)
            _lib.check(L.rise_hip_debug_trace_host_risemesh(self._h, rays.ctypes.data, None, n, 0, pid.ctypes.data,
                                                            t.ctypes.data, None, None, ctypes.byref(st)))
            return pid, t, st.as_dict()
        occ = np.empty(n, np.uint8)
        _lib.check(L.rise_hip_debug_trace_host_risemesh(self._h, rays.ctypes.data,
                                                        None if tm is None else tm.ctypes.data, n, 1, None, None,
                                                        occ.ctypes.data, None, ctypes.byref(st)))
        return occ, st.as_dict()

    def release(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib_risemesh().rise_hip_risemesh_release(self._h)
            self._h = _P(None)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def write_risemesh(path: str | os.PathLike, tris: np.ndarray, uvs: np.ndarray | None = None,
                   double_sided: bool = False, with_bvh_cache: bool = True) -> None:
    """Writes (n, 9) float32 triangles as a version-5 .risemesh file; uvs: optional
    (n, 6) per-corner texture coordinates."""
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    u = None if uvs is None else np.ascontiguousarray(uvs, np.float32).reshape(-1, 6)
    if u is not None and u.shape[0] != t.shape[0]:
        raise ValueError("uvs must have one row per triangle")
    _lib.check(_lib_risemesh().rise_hip_risemesh_write(os.fsencode(path), t.ctypes.data, t.shape[0],
                                                       None if u is None else u.ctypes.data,
                                                       1 if double_sided else 0, 1 if with_bvh_cache else 0))
