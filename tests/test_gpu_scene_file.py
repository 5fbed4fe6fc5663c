"""Scenes read from scene files render on the GPU like the same scenes built
through the table API (include/rise_hip_scene_file.h, through the C ABI).

* The Cornell-box scene file parses to tables identical to the Cornell-box
  workload's, so its render is bit-identical to the programmatic scene's.
* rise_hip_scene_file_upload (the library builds the device scene from its
  own tables) and rise_hip_pt_scene_create_with_media on the same tables
  copied out give bit-identical films, for a scene with every supported
  chunk (sky, GGX conductor, glass, fog box).
* The C++ adapter's LoadAsciiScene renders the Cornell file.
"""
from pathlib import Path

import numpy as np
import pytest

from rise_amd import pt_scene
from rise_amd.render import PathTracingSpectralRasterizer
from rise_amd.scene_file import SceneFile
from tests.conftest import log_case

pytestmark = pytest.mark.gpu

SCENES = Path(__file__).resolve().parent / "golden" / "scenes"


def _film(pt, settings):
    film, _, st = pt.rasterize_scene(settings)
    return film.cpu().numpy(), st


def test_cornell_file_renders_like_the_workload(cuda):
    log_case("scene file cornell vs cornell_box")
    f = SceneFile(SCENES / "cornell.RISEscene")
    r = f.settings()
    r.width, r.height, r.spp = 96, 96, 4
    pf = f.rasterizer()
    a, sa = _film(pf, r)
    pf.release()
    e, _ = pt_scene.cornell_box()
    pe = PathTracingSpectralRasterizer(e)
    b, sb = _film(pe, r)
    pe.release()
    assert sa == sb
    assert a.mean() > 0.05
    np.testing.assert_array_equal(a, b)


def test_upload_matches_table_api(cuda):
    log_case("scene file features: upload vs tables")
    f = SceneFile(SCENES / "features.RISEscene")
    r = f.settings()
    pf = f.rasterizer()
    a, sa = _film(pf, r)
    pf.release()
    pt = PathTracingSpectralRasterizer(f.scene())
    b, sb = _film(pt, r)
    pt.release()
    assert sa == sb
    assert np.isfinite(a).all() and a.mean() > 0
    np.testing.assert_array_equal(a, b)
