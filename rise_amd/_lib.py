"""ctypes binding of ``librise_hip.so`` (the C ABI in ``include/rise_hip.h``).

The library is loaded from this package directory only.  If it has not been
built, :func:`lib` raises: there is no CPU fallback for any GPU entry point.
"""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB_PATH = PKG / "librise_hip.so"
INCLUDE = ROOT / "include"

OK = 0
STATUS_NAMES = {
    0: "OK",
    1: "INVALID_ARGUMENT",
    2: "OUT_OF_MEMORY",
    3: "DEVICE",
    4: "BUILD",
    5: "NOT_SUPPORTED",
}


class RiseHipError(RuntimeError):
    """A C-ABI call returned a status other than RISE_HIP_OK."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("vertices", ctypes.c_void_p),
        ("n_tris", ctypes.c_int64),
        ("tri_ids", ctypes.c_void_p),
        ("vertices_f64", ctypes.c_void_p),
        ("flags", ctypes.c_uint32),
        ("normals", ctypes.c_void_p),
        ("coords", ctypes.c_void_p),
    ]


class SceneStats(ctypes.Structure):
    _fields_ = [
        ("n_tris", ctypes.c_int64),
        ("n_nodes", ctypes.c_int64),
        ("n_bvh2_nodes", ctypes.c_int64),
        ("depth", ctypes.c_int32),
        ("node_bytes", ctypes.c_int32),
        ("tri_bytes", ctypes.c_int32),
        ("bvh_source", ctypes.c_int32),
        ("build_ms", ctypes.c_double),
        ("sah_cost", ctypes.c_double),
        ("device_bytes", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_SIZE = ctypes.c_size_t

_SIGNATURES = {
    "rise_hip_last_error": ([], ctypes.c_char_p),
    "rise_hip_version": ([], ctypes.c_char_p),
    "rise_hip_scene_upload": ([_P, _SIZE, ctypes.POINTER(_P)], ctypes.c_int),
    "rise_hip_scene_release": ([_P], ctypes.c_int),
    "rise_hip_scene_get_stats": ([_P, ctypes.POINTER(SceneStats)], ctypes.c_int),
    "rise_hip_scene_set_arithmetic": ([_P, ctypes.c_int], ctypes.c_int),
    "rise_hip_intersect_closest": ([_P, _P, _I64, _P, _P, _P], ctypes.c_int),
    "rise_hip_intersect_any": ([_P, _P, _P, _I64, _P, _P], ctypes.c_int),
    "rise_hip_intersect_closest_noalloc": ([_P, _P, _I64, _P, _P, _P], ctypes.c_int),
    "rise_hip_intersect_closest_ex": ([_P, _P, _I64, _I32, _P, _P], ctypes.c_int),
    "rise_hip_intersect_any_ex": ([_P, _P, _P, _I64, _I32, _P, _P], ctypes.c_int),
    "rise_hip_debug_trace_ex_host": ([_P, _I64, _P, _P, _I64, _I32, _I32, _P, _P], ctypes.c_int),
    "rise_hip_intersect_closest_counted": ([_P, _P, _I64, _P, _P, _P, _P], ctypes.c_int),
    "rise_hip_intersect_any_counted": ([_P, _P, _P, _I64, _P, _P, _P], ctypes.c_int),
    "rise_hip_debug_trace_host": (
        [_P, _I64, _P, _P, _I64, _I32, _P, _P, _P, _P, ctypes.POINTER(SceneStats)],
        ctypes.c_int,
    ),
    "rise_hip_debug_node_accept_host": ([_P, _P, ctypes.c_double, _P], ctypes.c_int),
    "rise_hip_scene_upload_objects": ([_P, _SIZE, _P, _SIZE, ctypes.POINTER(_P)], ctypes.c_int),
    "rise_hip_debug_objects_host": ([_P, _SIZE, _P, _SIZE, _P, _P, _I64, _I32, _I32, _P, _P], ctypes.c_int),
    "rise_hip_debug_node_accept_obj_host": ([_P, _P, _P, ctypes.c_double, _P], ctypes.c_int),
}

_lib: ctypes.CDLL | None = None


def lib() -> ctypes.CDLL:
    """Loads and returns the library, raising if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -m rise_amd.build` (no CPU fallback exists)"
        )
    L = ctypes.CDLL(str(LIB_PATH))
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(status: int) -> None:
    """Raises RiseHipError with the library's message if status != OK."""
    if status != OK:
        msg = lib().rise_hip_last_error()
        raise RiseHipError(status, msg.decode() if msg else "")


def declared_symbols() -> list[str]:
    """Function names declared in include/*.h (the exported C ABI)."""
    names: list[str] = []
    for h in sorted(INCLUDE.glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(rise_hip_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))
