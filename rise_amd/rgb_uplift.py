"""RGB to spectrum uplift (Jakob & Hanika 2019) over the C ABI of
include/rise_hip_rgb_uplift.h.

``RgbLut.generate(res)`` regenerates the reference's Rec.709 table (its
tools/JakobHanikaLUTGen.cpp solver), ``RgbLut.read(path)`` loads a ".coeff"
file, ``lookup(rgb)`` is RGBToSpectrumTable::operator()
(Utilities/Color/RGBToSpectrumTable.cpp:168-243), and ``uplift_spd(rgb)``
returns the sigmoid spectrum a uniform RGB painter becomes
(RGBAlbedoSpectrum::FromRGB) as a ``pt_scene.Spd``.  Errors raise RiseHipError.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from .pt_scene import Spd


class RgbLutStats(ctypes.Structure):
    _fields_ = [
        ("cells", ctypes.c_int64),
        ("converged", ctypes.c_int64),
        ("failed", ctypes.c_int64),
        ("max_residual", ctypes.c_double),
        ("mean_residual", ctypes.c_double),
        ("build_ms", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = ctypes.c_void_p
_D3 = ctypes.c_double * 3
_F3 = ctypes.c_float * 3
_SIG = {
    "rise_hip_rgb_lut_generate": ([ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(_P), ctypes.POINTER(RgbLutStats)],
                                  ctypes.c_int),
    "rise_hip_rgb_lut_read": ([ctypes.c_char_p, ctypes.POINTER(_P)], ctypes.c_int),
    "rise_hip_rgb_lut_write": ([_P, ctypes.c_char_p], ctypes.c_int),
    "rise_hip_rgb_lut_get": ([_P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(_P)], ctypes.c_int),
    "rise_hip_rgb_lut_lookup": ([_P, _D3, _F3], ctypes.c_int),
    "rise_hip_rgb_lut_release": ([_P], ctypes.c_int),
    "rise_hip_rgb_lut_default": ([ctypes.POINTER(_P)], ctypes.c_int),
    "rise_hip_sigmoid_spectrum_eval": ([_F3, ctypes.c_double, ctypes.c_double], ctypes.c_double),
}


def _lib_uplift():
    L = _lib.lib()
    for name, (a, r) in _SIG.items():
        fn = getattr(L, name)
        fn.argtypes = a
        fn.restype = r
    return L


class RgbLut:
    """A Jakob-Hanika table (owned unless it is the library's default)."""

    def __init__(self, handle, owned: bool = True, stats: dict | None = None):
        self._h = handle
        self._owned = owned
        self.stats = stats

    @classmethod
    def generate(cls, resolution: int = 64, threads: int = 0) -> "RgbLut":
        h, st = _P(None), RgbLutStats()
        _lib.check(_lib_uplift().rise_hip_rgb_lut_generate(resolution, threads, ctypes.byref(h), ctypes.byref(st)))
        return cls(h, True, st.as_dict())

    @classmethod
    def read(cls, path: str | os.PathLike) -> "RgbLut":
        h = _P(None)
        _lib.check(_lib_uplift().rise_hip_rgb_lut_read(os.fsencode(path), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def default(cls) -> "RgbLut":
        """The library's 64^3 table (generated on first use)."""
        h = _P(None)
        _lib.check(_lib_uplift().rise_hip_rgb_lut_default(ctypes.byref(h)))
        return cls(h, owned=False)

    def write(self, path: str | os.PathLike) -> None:
        _lib.check(_lib_uplift().rise_hip_rgb_lut_write(self._h, os.fsencode(path)))

    def table(self) -> np.ndarray:
        """(3, res, res, res, 3) float32 copy: [max channel, iz, ix, iy, coefficient]."""
        r, p = ctypes.c_int32(), _P(None)
        _lib.check(_lib_uplift().rise_hip_rgb_lut_get(self._h, ctypes.byref(r), ctypes.byref(p)))
        n = r.value
        a = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_float)), (9 * n ** 3,))
        return a.reshape(3, n, n, n, 3).copy()

    def lookup(self, rgb) -> np.ndarray:
        out = _F3()
        _lib.check(_lib_uplift().rise_hip_rgb_lut_lookup(self._h, _D3(*map(float, rgb)), out))
        return np.array(out[:], np.float32)

    def release(self) -> None:
        if self._owned and getattr(self, "_h", None) is not None and self._h.value:
            _lib_uplift().rise_hip_rgb_lut_release(self._h)
        self._h = _P(None)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def sigmoid_eval(coeffs, scale: float, nm: float) -> float:
    """RGBSigmoidPolynomial::Eval(nm) * scale (double)."""
    return _lib_uplift().rise_hip_sigmoid_spectrum_eval(_F3(*map(float, coeffs)), float(scale), float(nm))


def uplift_spd(rgb, lut: RgbLut | None = None) -> Spd:
    """The spectrum of a uniform linear-Rec.709 RGB albedo (scale 1)."""
    c = (lut or RgbLut.default()).lookup(rgb)
    return Spd(380.0, 780.0, np.array([c[0], c[1], c[2], 1.0], np.float32), sigmoid=True)
