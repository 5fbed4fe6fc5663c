"""The output stage (include/rise_hip_output.h) on host threads, no GPU.

* PPM, TGA and HDR files equal, byte for byte, what the reference's own
  colour pipeline and writers produce for the same XYZ film
  (ColorUtils::XYZtoRec709RGB with its gamut clip, DisplayTransformWriter,
  PPMWriter / TGAWriter / HDRWriter compiled into oracle/_ref): the
  committed fixtures tests/golden/output/ (tests/golden/make_output_golden.py)
  and, when oracle/_ref is built, every combination of colour space, display
  transform and exposure live.  The HDR header's generator comment line is
  the only difference.
* PNG (written here directly; the reference uses libpng): the zlib stream
  inflates to the same RGBA as the TGA path; 16-bit samples are big-endian
  and agree with the 8-bit ones.
* EXR (written here directly; the reference uses OpenEXR): parsed back, the
  float channels are the reference's linear Rec.709 values rounded to float,
  half channels their round-to-nearest-even halves.
"""
import ctypes
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

import oracle
from rise_amd._lib import RiseHipError
from rise_amd.output import extension, output_desc, write_image

OUT = Path(__file__).resolve().parent / "golden" / "output"
DT = ["none", "reinhard", "aces", "agx", "hable"]


def _film():
    return np.load(OUT / "film.npy")


def _body(data: bytes, ext: str) -> bytes:
    return data.split(b"\n", 2)[2] if ext == "hdr" else data


@pytest.mark.parametrize("path", sorted(p.name for p in OUT.glob("*.*") if p.suffix != ".npy"))
def test_matches_reference_writer_fixtures(tmp_path, path):
    ext, cs, dt, ev = Path(path).stem.split("_")[0], *Path(path).stem.split("_")[1:]
    fmt = {"tga": "TGA", "ppm": "PPM", "hdr": "HDR"}[ext]
    d = output_desc(fmt, 8, "sRGB" if cs == "srgb" else "Rec709RGB_Linear", dt, float(ev))
    write_image(_film(), tmp_path / path, d)
    assert _body((tmp_path / path).read_bytes(), ext) == _body((OUT / path).read_bytes(), ext)


def _ref_lib():
    L = ctypes.CDLL(str(oracle.REF_DIR / "libriseref_ieee.so"))
    L.rise_ref_write_image.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_char_p]
    L.rise_ref_xyz_to_rec709.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    return L


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref (the reference's own build) not built")
def test_every_ldr_combination_matches_reference_live(tmp_path):
    L = _ref_lib()
    g = np.random.default_rng(4)
    f = np.ascontiguousarray(np.concatenate([_film(), g.gamma(0.5, 2.0, (9, 37, 3)).astype(np.float32)], 0))
    h, w = f.shape[:2]
    for t, fmt in ((0, "TGA"), (1, "PPM")):
        for cs in (0, 1):
            for dt in range(5):
                for ev in (0.0, 1.5, -2.0):
                    ours, ref = tmp_path / "o.bin", tmp_path / "r.bin"
                    write_image(f, ours, output_desc(fmt, 8, ["Rec709RGB_Linear", "sRGB"][cs], DT[dt], ev))
                    assert L.rise_ref_write_image(f.ctypes.data, w, h, t, cs, ev, dt, str(ref).encode()) == 0
                    assert ours.read_bytes() == ref.read_bytes(), (fmt, cs, dt, ev)


def _png_rgba(data: bytes):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    p, idat, ihdr = 8, b"", None
    while p < len(data):
        n, = struct.unpack(">I", data[p:p + 4])
        typ, body = data[p + 4:p + 8], data[p + 8:p + 8 + n]
        assert struct.unpack(">I", data[p + 8 + n:p + 12 + n])[0] == zlib.crc32(typ + body)
        if typ == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif typ == b"IDAT":
            idat += body
        p += 12 + n
    w, h, depth, ctype = ihdr[:4]
    assert ctype == 6
    raw = zlib.decompress(idat)
    bps = depth // 8
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + w * 4 * bps)
    assert np.all(rows[:, 0] == 0)
    px = rows[:, 1:]
    if bps == 2:
        return px.reshape(h, w, 4, 2).astype(np.uint16) @ np.array([256, 1], np.uint16)
    return px.reshape(h, w, 4)


def test_png_8_and_16_bit(tmp_path):
    f = _film()
    write_image(f, tmp_path / "a.png", output_desc("PNG", 8, "sRGB", "aces"))
    write_image(f, tmp_path / "a.tga", output_desc("TGA", 8, "sRGB", "aces"))
    tga = np.frombuffer((tmp_path / "a.tga").read_bytes()[18:], np.uint8).reshape(f.shape[0], f.shape[1], 4)
    tga = tga[::-1][..., [2, 1, 0, 3]]  # bottom-up BGRA -> top-down RGBA
    np.testing.assert_array_equal(_png_rgba((tmp_path / "a.png").read_bytes()), tga)
    write_image(f, tmp_path / "b.png", output_desc("PNG", 16, "sRGB", "aces"))
    p16 = _png_rgba((tmp_path / "b.png").read_bytes()).astype(np.int64)
    assert np.all(p16[..., 3] == 65535)
    # floor(v * 255) against floor(v * 65535) / 257: equal up to the boundary
    assert np.abs(p16[..., :3] / 257.0 - tga[..., :3]).max() < 1.0 + 1e-9


def _exr_read(data: bytes):
    assert data[:4] == b"\x76\x2f\x31\x01" and data[4] == 2
    p, attrs = 8, {}
    while data[p] != 0:
        e = data.index(b"\0", p)
        name = data[p:e].decode()
        e2 = data.index(b"\0", e + 1)
        n, = struct.unpack("<i", data[e2 + 1:e2 + 5])
        attrs[name] = data[e2 + 5:e2 + 5 + n]
        p = e2 + 5 + n
    p += 1
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    chans, q = [], 0
    cl = attrs["channels"]
    while cl[q] != 0:
        e = cl.index(b"\0", q)
        chans.append((cl[q:e].decode(), struct.unpack("<i", cl[e + 1:e + 5])[0]))
        q = e + 17
    assert attrs["compression"] == b"\0"
    offs = struct.unpack(f"<{h}Q", data[p:p + 8 * h])
    out = {c: np.zeros((h, w), np.float32) for c, _ in chans}
    for y, o in enumerate(offs):
        yy, size = struct.unpack("<ii", data[o:o + 8])
        assert yy == y
        q = o + 8
        for c, t in chans:
            dt = np.float16 if t == 1 else np.float32
            n = w * np.dtype(dt).itemsize
            out[c][y] = np.frombuffer(data[q:q + n], dt).astype(np.float32)
            q += n
        assert q == o + 8 + size
    return out


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref (the reference's own build) not built")
def test_exr_float_and_half(tmp_path):
    f = _film()
    rgb = np.zeros(f.shape, np.float64)
    _ref_lib().rise_ref_xyz_to_rec709(f.ctypes.data, f.shape[0] * f.shape[1], rgb.ctypes.data)
    write_image(f, tmp_path / "a.exr", output_desc("EXR", 32, "Rec709RGB_Linear"))
    e = _exr_read((tmp_path / "a.exr").read_bytes())
    for k, c in enumerate("RGB"):
        np.testing.assert_array_equal(e[c], rgb[..., k].astype(np.float32))
    np.testing.assert_array_equal(e["A"], 1.0)
    write_image(f, tmp_path / "b.exr", output_desc("EXR", 16, "Rec709RGB_Linear", with_alpha=False))
    e = _exr_read((tmp_path / "b.exr").read_bytes())
    assert "A" not in e
    for k, c in enumerate("RGB"):
        np.testing.assert_array_equal(e[c], rgb[..., k].astype(np.float32).astype(np.float16).astype(np.float32))


def _bad_cs():
    d = output_desc("PPM")
    d.color_space = 2  # ROMM RGB: not written here
    return d


def test_errors_and_extensions(tmp_path):
    f = _film()
    # PNGWriter's fallback for other depths is 8 bits
    write_image(f, tmp_path / "x.png", output_desc("PNG", 12))
    write_image(f, tmp_path / "y.png", output_desc("PNG", 8))
    assert (tmp_path / "x.png").read_bytes() == (tmp_path / "y.png").read_bytes()
    with pytest.raises(RiseHipError, match="colour space"):
        write_image(f, tmp_path / "x.ppm", _bad_cs())
    with pytest.raises(RiseHipError, match="cannot create"):
        write_image(f, tmp_path / "no" / "such" / "x.ppm", output_desc("PPM"))
    assert [extension(k) for k in ("TGA", "PPM", "PNG", "HDR", "EXR")] == ["tga", "ppm", "png", "hdr", "exr"]
