"""Writes the output-stage fixtures under tests/golden/output/ with the
reference's own colour pipeline and writers (ColorUtils::XYZtoRec709RGB,
DisplayTransformWriter, PPMWriter, TGAWriter, HDRWriter), compiled from the
reference sources by oracle/ref/Makefile into oracle/_ref/libriseref_ieee.so
and driven by oracle/ref/ref_harness.cpp (rise_ref_write_image).  Run in a
checkout that has the reference tree and the built oracle/_ref:

    python tests/golden/make_output_golden.py

film.npy is the input: a 23 x 37 XYZ film with black, saturated (outside the
Rec.709 gamut) and over-bright pixels.  Each image is named
<format>_<colour space>_<display transform>_<exposure EV>.<ext>.
"""
from __future__ import annotations

import ctypes
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
OUT = Path(__file__).resolve().parent / "output"

CASES = [  # (type, ext, colour space, display transform, EV)
    (0, "tga", 1, 2, 0.0),
    (0, "tga", 0, 4, 1.5),
    (1, "ppm", 1, 0, 0.0),
    (1, "ppm", 1, 3, -1.0),
    (1, "ppm", 0, 1, 0.5),
    (3, "hdr", 0, 0, 0.0),
    (3, "hdr", 1, 0, 0.0),
]


def film():
    g = np.random.default_rng(1)
    f = g.gamma(0.7, 0.6, (23, 37, 3)).astype(np.float32)
    f[0, :5] = 0
    f[1, :4] = [[5, 0.1, 0.1], [0.0, 1, 0], [0.1, 0.1, 3], [0.2, 0.3, 0.4]]
    f[2, :3] *= 20
    return f


def name(t, ext, cs, dt, ev):
    return f"{ext}_{['linear', 'srgb'][cs]}_{['none', 'reinhard', 'aces', 'agx', 'hable'][dt]}_{ev:+.1f}.{ext}"


def main():
    L = ctypes.CDLL(str(ROOT / "oracle" / "_ref" / "libriseref_ieee.so"))
    L.rise_ref_write_image.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_char_p]
    OUT.mkdir(exist_ok=True)
    f = film()
    np.save(OUT / "film.npy", f)
    for t, ext, cs, dt, ev in CASES:
        p = OUT / name(t, ext, cs, dt, ev)
        if L.rise_ref_write_image(f.ctypes.data, f.shape[1], f.shape[0], t, cs, ev, dt, str(p).encode()) != 0:
            sys.exit(f"reference writer failed for {p.name}")


if __name__ == "__main__":
    main()
