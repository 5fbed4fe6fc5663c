"""The output stage over the C ABI of include/rise_hip_output.h: a rendered
XYZ film written as an image file (the reference's FileRasterizerOutput with
its colour pipeline, display transforms and writers).

``write_image(film, path, ...)`` resolves a CUDA film on the GPU (or a numpy
film on host threads) and writes TGA, PPM, PNG, HDR or EXR.  Errors raise
RiseHipError.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib

FORMATS = {"TGA": 0, "PPM": 1, "PNG": 2, "HDR": 3, "EXR": 6}
COLOR_SPACES = {"Rec709RGB_Linear": 0, "sRGB": 1}
DISPLAY = {"default": -1, "none": 0, "reinhard": 1, "aces": 2, "agx": 3, "hable": 4}


class OutputDesc(ctypes.Structure):
    _fields_ = [
        ("format", ctypes.c_int32),
        ("bpp", ctypes.c_int32),
        ("color_space", ctypes.c_int32),
        ("display_transform", ctypes.c_int32),
        ("exposure_ev", ctypes.c_double),
        ("with_alpha", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_SIG = {
    "rise_hip_output_write": ([_P, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(OutputDesc), ctypes.c_char_p, _P],
                              ctypes.c_int),
    "rise_hip_output_write_host": ([_P, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(OutputDesc),
                                    ctypes.c_char_p], ctypes.c_int),
    "rise_hip_output_extension": ([ctypes.c_int32], ctypes.c_char_p),
    "rise_hip_output_resolve": ([_P, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(OutputDesc), _P, _P],
                                ctypes.c_int),
}


def resolve(film, out, desc: OutputDesc, stream=None) -> None:
    """The per-pixel stage alone on the GPU: film (H, W, 3) float32 CUDA tensor
    into `out` (a CUDA tensor of the format's pixel bytes)."""
    import torch

    h, w = film.shape[:2]
    s = torch.cuda.current_stream().cuda_stream if stream is None else int(getattr(stream, "cuda_stream", stream))
    _lib.check(_lib_output().rise_hip_output_resolve(film.data_ptr(), w, h, ctypes.byref(desc), out.data_ptr(), s))


def _lib_output():
    L = _lib.lib()
    for name, (a, r) in _SIG.items():
        fn = getattr(L, name)
        fn.argtypes = a
        fn.restype = r
    return L


def output_desc(fmt: str = "TGA", bpp: int | None = None, color_space: str = "sRGB", display: str = "default",
                exposure: float = 0.0, with_alpha: bool = True) -> OutputDesc:
    """The file_rasterizeroutput parameters (defaults as the reference's parser:
    TGA, 8 bits, sRGB, ACES for LDR formats and none for HDR ones)."""
    f = FORMATS[fmt]
    if bpp is None:
        bpp = 32 if f == FORMATS["EXR"] else 8
    return OutputDesc(f, bpp, COLOR_SPACES[color_space], DISPLAY[display], float(exposure), 1 if with_alpha else 0, 0)


def write_image(film, path: str | os.PathLike, desc: OutputDesc | None = None, stream=None) -> None:
    """film: (H, W, 3) float32 mean XYZ, a CUDA tensor (resolved on the GPU) or a
    numpy array (resolved on host threads)."""
    desc = desc or output_desc()
    L = _lib_output()
    if isinstance(film, np.ndarray):
        f = np.ascontiguousarray(film, np.float32)
        h, w = f.shape[:2]
        _lib.check(L.rise_hip_output_write_host(f.ctypes.data, w, h, ctypes.byref(desc), os.fsencode(path)))
        return
    import torch

    if film.dtype != torch.float32 or not film.is_cuda or not film.is_contiguous() or film.dim() != 3:
        raise ValueError("film must be a contiguous (H, W, 3) float32 CUDA tensor")
    h, w = film.shape[:2]
    s = torch.cuda.current_stream().cuda_stream if stream is None else int(getattr(stream, "cuda_stream", stream))
    _lib.check(L.rise_hip_output_write(film.data_ptr(), w, h, ctypes.byref(desc), os.fsencode(path), s))


def extension(fmt: str) -> str:
    return _lib_output().rise_hip_output_extension(FORMATS[fmt]).decode()
