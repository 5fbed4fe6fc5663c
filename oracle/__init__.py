"""Loaders for the oracle libraries.  TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this package, and only as the checker (or the measured CPU baseline);
the product path never does.

* ``liboracle.so`` (oracle/Makefile): the C restatement (rt_oracle.c).
* ``_ref/libriseref_ieee.so`` / ``_ref/libriseref.so`` (oracle/ref/Makefile):
  the reference's own mesh intersection path compiled from its sources.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_DIR = HERE / "_ref"
REFERENCE_TREE = Path("/root/reference")

_P = ctypes.c_void_p
_I64 = ctypes.c_int64


def build(with_ref: bool = True) -> None:
    """Builds liboracle.so, and oracle/_ref when the reference tree exists."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if with_ref and REFERENCE_TREE.exists():
        subprocess.run(["make", "-s", "-j8", "-C", str(HERE / "ref"), f"REF={REFERENCE_TREE}"], check=True)


def _load(path: Path) -> ctypes.CDLL:
    if not path.exists():
        raise FileNotFoundError(f"{path} not built")
    return ctypes.CDLL(str(path))


_oracle = None


def oracle_lib() -> ctypes.CDLL:
    global _oracle
    if _oracle is None:
        L = _load(ORACLE_SO)
        L.rise_oracle_ray_triangle.argtypes = [_P] * 5 + [_P, _P, _P]
        L.rise_oracle_ray_triangle.restype = ctypes.c_int
        L.rise_oracle_normalize.argtypes = [_P, _P]
        L.rise_oracle_closest_brute.argtypes = [_P, _I64, _P, _I64, _P, _P, _P, ctypes.c_int]
        L.rise_oracle_any_brute.argtypes = [_P, _I64, _P, _P, _I64, _P, ctypes.c_int]
        _oracle = L
    return _oracle


def ray_triangle(o, d, v0, eA, eB):
    """RayTriangleIntersection on double inputs; returns (hit, t, alpha, beta)."""
    L = oracle_lib()
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (o, d, v0, eA, eB)]
    out = np.zeros(3, dtype=np.float64)
    hit = L.rise_oracle_ray_triangle(
        *[a.ctypes.data for a in arrs],
        out[0:].ctypes.data, out[1:].ctypes.data, out[2:].ctypes.data,
    )
    return bool(hit), float(out[0]), float(out[1]), float(out[2])


def closest(tris: np.ndarray, rays: np.ndarray, threads: int = 0):
    """Brute-force closest hit.  Returns (prim_id int32, t float32, t64 float64)."""
    L = oracle_lib()
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    n = rays.shape[1]
    pid = np.empty(n, dtype=np.int32)
    t = np.empty(n, dtype=np.float32)
    t64 = np.empty(n, dtype=np.float64)
    L.rise_oracle_closest_brute(tris.ctypes.data, tris.shape[0], rays.ctypes.data, n, pid.ctypes.data,
                                t.ctypes.data, t64.ctypes.data, threads)
    return pid, t, t64


def any_hit(tris: np.ndarray, rays: np.ndarray, tmax: np.ndarray | None = None, threads: int = 0) -> np.ndarray:
    L = oracle_lib()
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    n = rays.shape[1]
    occ = np.empty(n, dtype=np.uint8)
    tm = None if tmax is None else np.ascontiguousarray(tmax, dtype=np.float32)
    L.rise_oracle_any_brute(tris.ctypes.data, tris.shape[0], rays.ctypes.data,
                            None if tm is None else tm.ctypes.data, n, occ.ctypes.data, threads)
    return occ


class RefMesh:
    """The reference's TriangleMeshGeometryIndexed (BVH4 + float filter +
    double certifier) built from its own sources (oracle/_ref)."""

    def __init__(self, tris: np.ndarray, variant: str = "ieee"):
        name = "libriseref_ieee.so" if variant == "ieee" else "libriseref.so"
        L = _load(REF_DIR / name)
        L.rise_ref_mesh_create.argtypes = [_P, _I64]
        L.rise_ref_mesh_create.restype = _P
        L.rise_ref_mesh_release.argtypes = [_P]
        L.rise_ref_closest.argtypes = [_P, _P, _I64, _P, _P, _P, ctypes.c_int]
        L.rise_ref_any.argtypes = [_P, _P, _P, _I64, _P, ctypes.c_int]
        self._L = L
        self._tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
        self._h = L.rise_ref_mesh_create(self._tris.ctypes.data, self._tris.shape[0])

    def closest(self, rays: np.ndarray, threads: int = 0):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = rays.shape[1]
        pid = np.empty(n, dtype=np.int32)
        t = np.empty(n, dtype=np.float32)
        t64 = np.empty(n, dtype=np.float64)
        self._L.rise_ref_closest(self._h, rays.ctypes.data, n, pid.ctypes.data, t.ctypes.data, t64.ctypes.data,
                                 threads)
        return pid, t, t64

    def any_hit(self, rays: np.ndarray, tmax: np.ndarray | None = None, threads: int = 0) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = rays.shape[1]
        occ = np.empty(n, dtype=np.uint8)
        tm = None if tmax is None else np.ascontiguousarray(tmax, dtype=np.float32)
        self._L.rise_ref_any(self._h, rays.ctypes.data, None if tm is None else tm.ctypes.data, n,
                             occ.ctypes.data, threads)
        return occ

    def close(self) -> None:
        if self._h:
            self._L.rise_ref_mesh_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ref_available(variant: str = "ieee") -> bool:
    return (REF_DIR / ("libriseref_ieee.so" if variant == "ieee" else "libriseref.so")).exists()
