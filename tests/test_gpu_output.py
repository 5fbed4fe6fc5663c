"""The output stage on the GPU (rise_hip_output_write), through the C ABI.

* The XYZ film fixture resolved on the GPU gives, byte for byte, the files
  the reference's own writers made from it (tests/golden/output/; HDR
  headers differ only in the generator comment line).
* PNG (8 and 16 bits) and EXR (half and float) from the GPU equal the host
  path's files.
* A rendered Cornell box is written through the scene file's own
  file_rasterizeroutput chunk (PNG); the file equals the host path's on the
  same film.
"""
from pathlib import Path

import numpy as np
import pytest

from rise_amd.output import output_desc, write_image
from rise_amd.scene_file import SceneFile
from tests.conftest import log_case
from tests.test_output import _png_rgba

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
OUT = GOLDEN / "output"


def _body(data: bytes, ext: str) -> bytes:
    return data.split(b"\n", 2)[2] if ext == "hdr" else data


def test_gpu_output_matches_reference_fixtures(cuda, tmp_path):
    import torch

    log_case("output fixtures on the GPU")
    film = torch.from_numpy(np.load(OUT / "film.npy")).to(cuda).contiguous()
    names = sorted(p.name for p in OUT.glob("*.*") if p.suffix != ".npy")
    assert len(names) == 7
    for name in names:
        ext, cs, dt, ev = Path(name).stem.split("_")[0], *Path(name).stem.split("_")[1:]
        d = output_desc({"tga": "TGA", "ppm": "PPM", "hdr": "HDR"}[ext], 8,
                        "sRGB" if cs == "srgb" else "Rec709RGB_Linear", dt, float(ev))
        write_image(film, tmp_path / name, d)
        assert _body((tmp_path / name).read_bytes(), ext) == _body((OUT / name).read_bytes(), ext), name


def test_gpu_resolve_device_to_device(cuda):
    import torch

    from rise_amd.output import resolve

    log_case("output resolve device to device")
    f = np.load(OUT / "film.npy")
    film = torch.from_numpy(f).to(cuda).contiguous()
    px = torch.empty((f.shape[0], f.shape[1], 4), dtype=torch.uint8, device=cuda)
    resolve(film, px, output_desc("PPM", 8, "sRGB", "none", 0.0))
    torch.cuda.synchronize()
    ppm = (OUT / "ppm_srgb_none_+0.0.ppm").read_bytes()
    body = np.frombuffer(ppm[len(b"P6\n37 23\n255\n"):], np.uint8).reshape(f.shape[0], f.shape[1], 3)
    got = px.cpu().numpy()
    np.testing.assert_array_equal(got[..., :3], body)
    assert np.all(got[..., 3] == 255)


@pytest.mark.parametrize("fmt,bpp", [("PNG", 8), ("PNG", 16), ("EXR", 16), ("EXR", 32)])
def test_gpu_png_exr_equal_host(cuda, tmp_path, fmt, bpp):
    import torch

    log_case(f"output {fmt} {bpp} gpu vs host")
    f = np.load(OUT / "film.npy")
    d = output_desc(fmt, bpp, "sRGB" if fmt == "PNG" else "Rec709RGB_Linear")
    write_image(torch.from_numpy(f).to(cuda).contiguous(), tmp_path / "g.bin", d)
    write_image(f, tmp_path / "h.bin", d)
    assert (tmp_path / "g.bin").read_bytes() == (tmp_path / "h.bin").read_bytes()


def test_gpu_render_to_scene_file_output(cuda, tmp_path):
    log_case("cornell render -> file_rasterizeroutput")
    sf = SceneFile(GOLDEN / "scenes" / "cornell.RISEscene")
    r = sf.settings()
    r.width, r.height, r.spp = 64, 48, 4
    pt = sf.rasterizer()
    film, _, _ = pt.rasterize_scene(r)
    (d, name), = sf.outputs()
    assert name == "rendered/cornell.png"  # the chunk's pattern and type PNG
    write_image(film, tmp_path / "g.png", d)
    write_image(film.cpu().numpy(), tmp_path / "h.png", d)
    data = (tmp_path / "g.png").read_bytes()
    assert data == (tmp_path / "h.png").read_bytes()
    px = _png_rgba(data)
    assert px.shape == (48, 64, 4) and px[..., :3].max() > 100  # the lit box is visible
    pt.release()
