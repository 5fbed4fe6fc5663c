"""Scene files (native "RISE ASCII SCENE 7") read by the library's host
parser (include/rise_hip_scene_file.h), no GPU.

* The Cornell-box workload written as a scene file (tests/golden/scenes)
  parses to the programmatic scene of rise_amd.pt_scene.cornell_box():
  same objects, materials, spectra (evaluated on the render's wavelengths),
  camera and render settings.
* Every supported chunk (features.RISEscene): spectral painters follow the
  piecewise-linear sampling of Job::AddSpectralColorPainter (known answers),
  transforms follow scale -> rotate z, y, x -> translate (known answers for
  quarter turns, which fix the order), media, sky and defaults.
* Unsupported input is an error naming the chunk and line.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from rise_amd import pt_scene
from rise_amd._lib import RiseHipError
from rise_amd.scene_file import SceneFile, load_scene

SCENES = Path(__file__).resolve().parent / "golden" / "scenes"


def _spd_eval(spd, nm):
    """SpectralPacket::ValueAtNM step bins (0 outside the range)."""
    out = np.zeros_like(nm)
    inside = (nm >= spd.begin) & (nm <= spd.end)
    idx = np.minimum(((nm - spd.begin) / ((spd.end - spd.begin) / len(spd.amp))).astype(int), len(spd.amp) - 1)
    out[inside] = spd.amp[idx[inside]]
    return out


def test_cornell_file_matches_programmatic_scene():
    s, r = load_scene(SCENES / "cornell.RISEscene")
    e, er = pt_scene.cornell_box()
    assert r == er
    assert s.camera == e.camera
    assert s.sky is None and s.media == []
    assert len(s.objects) == len(e.objects)
    for a, b in zip(s.objects, e.objects):
        assert a.tris.shape == b.tris.shape
        np.testing.assert_array_equal(a.tris, b.tris)
        assert (a.double_sided, a.medium, a.casts_shadows) == (b.double_sided, b.medium, b.casts_shadows)
    nm = np.linspace(er.lambda_begin, er.lambda_end, 341)
    assert len(s.materials) == len(e.materials)
    for a, b in zip(s.materials, e.materials):
        assert (a.kind, a.emission_scale) == (b.kind, pytest.approx(b.emission_scale, rel=1e-7))
        for fa, fb in ((a.reflectance, b.reflectance), (a.emission, b.emission)):
            assert (fa < 0) == (fb < 0)
            if fa >= 0:
                np.testing.assert_array_equal(_spd_eval(s.spds[fa], nm), _spd_eval(e.spds[fb], nm))


def test_spectral_painter_sampling():
    s = SceneFile(SCENES / "features.RISEscene").scene()
    # ramp: cp (700, 1), (400, 0) over [400, 700]: 2 bins sampled at 400, 550
    np.testing.assert_array_equal(s.spds[0].amp, np.float32([0.0, 0.5]))
    assert (s.spds[0].begin, s.spds[0].end) == (400.0, 700.0)
    # from a file: (380 .3) (500 .8) (600 .25) (780 .17), 4 bins at 380, 480, 580, 680
    want = [0.3, 0.3 + 0.5 * 100 / 120, 0.8 - 0.55 * 80 / 100, 0.25 - 0.08 * 80 / 180]
    np.testing.assert_allclose(s.spds[1].amp, want, rtol=1e-6)
    # two points, 2 bins at 380 and 580: 1.9 and the midpoint 3.25
    np.testing.assert_allclose(s.spds[2].amp, [1.9, 3.25], rtol=1e-7)
    # a number is a uniform painter: one value at every wavelength
    flat = s.spds[s.materials[2].ior]
    assert flat.amp.tolist() == [1.5] and flat.begin <= 300 and flat.end >= 900


def test_materials_media_sky_and_settings():
    f = SceneFile(SCENES / "features.RISEscene")
    info = f.info()
    assert info == {"n_objects": 4, "n_materials": 4, "n_spds": 5, "n_media": 2, "n_tris": 38, "has_sky": 1,
                    "n_skipped_chunks": 1, "n_outputs": 1, "n_directional_lights": 0}
    (d, name), = f.outputs()
    # type EXR; bpp, colour space, display transform and alpha at the parser's defaults
    assert name == "rendered/features.exr"
    assert (d.format, d.bpp, d.color_space, d.display_transform, d.exposure_ev, d.with_alpha) == (6, 8, 1, -1, 0.0, 1)
    s, r = f.scene(), f.settings()
    ground, gold, glass, shell = s.materials
    assert ground.kind == pt_scene.LAMBERTIAN and ground.reflectance == 0 and ground.emission == -1
    assert gold.kind == pt_scene.GGX and gold.reflectance == -1 and gold.alpha == pytest.approx(0.2)
    assert (gold.ior, gold.extinction) == (1, 2) and s.spds[gold.specular].amp.tolist() == [1.0]
    assert glass.kind == pt_scene.DIELECTRIC and glass.extinction == -1
    assert s.spds[shell.ior].amp.tolist() == [1.0]
    fog, haze = s.media
    assert fog.sigma_a == pytest.approx((0.05,) * 3) and fog.sigma_s == pytest.approx((0.5, 0.4, 0.3))
    assert fog.g == pytest.approx(0.3) and haze.g == 0.0
    assert s.sky == pt_scene.Sky(30.0, 120.0, 4.0, 0.5, create_sun=False)
    assert [(o.medium, o.casts_shadows, o.double_sided) for o in s.objects] == [
        (-1, True, True), (-1, True, False), (-1, True, False), (0, False, False)]
    assert s.camera == pt_scene.Camera((0.0, 1.0, 4.0), (0.0, 0.5, 0.0), (0.0, 1.0, 0.0), 45.0)
    assert (r.width, r.height, r.spp, r.lambda_begin, r.lambda_end) == (64, 48, 4, 390.0, 710.0)
    assert (r.rr_min_depth, r.max_diffuse_bounce, r.max_glossy_bounce) == (4, 6, 255)
    assert r.rr_threshold == pytest.approx(0.1) and r.sky_background is False and r.seed == 1


def _cube_vertex_map(tris):
    """The unit-cube fixture's corner (x, y, z) -> transformed point, from the
    first vertex of each triangle slot (triangle order is the file's)."""
    idx = [(0, 1, 3), (0, 3, 2), (4, 6, 7), (4, 7, 5), (0, 4, 5), (0, 5, 1), (2, 3, 7), (2, 7, 6), (0, 2, 6),
           (0, 6, 4), (1, 5, 7), (1, 7, 3)]
    out = {}
    for t, ids in enumerate(idx):
        for j, v in enumerate(ids):
            out[(v >> 2 & 1, v >> 1 & 1, v & 1)] = tris[t, 3 * j:3 * j + 3].astype(np.float64)
    return out


def test_object_transforms():
    s = SceneFile(SCENES / "features.RISEscene").scene()
    # gold: scale 0.5, 30 degrees about +y (right-handed: x -> -z), then +x 1
    g = _cube_vertex_map(s.objects[1].tris)
    c, sn = math.cos(math.radians(30)), math.sin(math.radians(30))
    np.testing.assert_allclose(g[(1, 0, 0)], [1 + 0.5 * c, 0, -0.5 * sn], atol=1e-6)
    np.testing.assert_allclose(g[(0, 0, 1)], [1 + 0.5 * sn, 0, 0.5 * c], atol=1e-6)
    np.testing.assert_allclose(g[(0, 1, 0)], [1, 0.5, 0], atol=1e-6)
    # glass: orientation (90, 90, 0) = y quarter turn, then x quarter turn:
    # x -> -z -> +y, y -> y -> +z, z -> x -> x
    q = _cube_vertex_map(s.objects[2].tris)
    o = np.array([-1, 0.25, 0])
    np.testing.assert_allclose(q[(1, 0, 0)], o + [0, 0.5, 0], atol=1e-6)
    np.testing.assert_allclose(q[(0, 1, 0)], o + [0, 0, 0.5], atol=1e-6)
    np.testing.assert_allclose(q[(0, 0, 1)], o + [0.5, 0, 0], atol=1e-6)
    # fog box: non-uniform scale then translation
    b = _cube_vertex_map(s.objects[3].tris)
    np.testing.assert_allclose(b[(1, 1, 1)], [2, 1, 2], atol=0)
    np.testing.assert_allclose(b[(0, 0, 0)], [-2, 0, -2], atol=0)


def test_clipped_plane_triangles():
    s = SceneFile(SCENES / "features.RISEscene").scene()
    a, b, c, d = (-5, 0, -5), (-5, 0, 5), (5, 0, 5), (5, 0, -5)
    np.testing.assert_array_equal(s.objects[0].tris, np.float32([a + b + c, a + c + d]))


BASE = """RISE ASCII SCENE 7
lambertian_material
{
\tname\tm
\treflectance\t0.5
}
clippedplane_geometry
{
\tname\tq
\tpta\t0 0 0
\tptb\t1 0 0
\tptc\t1 1 0
\tptd\t0 1 0
}
standard_object
{
\tgeometry\tq
\tmaterial\tm
}
pinhole_camera
{
\tlocation\t0 0 5
}
"""


@pytest.mark.parametrize("extra, needle", [
    ("bogus_chunk\n{\n\tname x\n}\n", "bogus_chunk (line 24)"),
    ("uniformcolor_painter\n{\n\tname c\n\tcolor 1 0 0\n\tcolorspace ROMMRGB_Linear\n}\n", "colorspace"),
    ("directional_light\n{\n\tpower 2\n}\n", "`direction` is zero"),
    ("dielectric_material\n{\n\tname d\n\tior 1.5\n}\n", "smooth dielectrics"),
    ("ggx_material\n{\n\tname g\n\trs 1\n\talphax 0.1\n\talphay 0.3\n}\n", "anisotropic"),
    ("lambertian_material\n{\n\tname e\n\treflectance expr(0.5)\n}\n", "expr"),
    ("standard_object\n{\n\tgeometry q\n\tmaterial nothere\n}\n", "unknown material"),
    ("rawmesh2_geometry\n{\n\tname r\n\tfile does_not_exist.raw2\n}\n", "cannot open"),
    ("homogeneous_medium\n{\n\tname h\n\tphase schlick 0.2\n}\n", "phase"),
    ("> run other.RISEscene\n", "not supported"),
    ("> echo hello\n> set accelerator bvh\nbogus\n{\n}\n", "bogus (line 26)"),  # neutral commands pass
    ("film\n{\n\twidth 8\n", "unterminated"),
])
def test_errors_name_the_problem(tmp_path, extra, needle):
    p = tmp_path / "bad.RISEscene"
    p.write_text(BASE + extra)
    with pytest.raises(RiseHipError) as e:
        SceneFile(p)
    assert needle in str(e.value)


def test_minimal_scene_defaults(tmp_path):
    p = tmp_path / "ok.RISEscene"
    p.write_text(BASE)
    s, r = load_scene(p)
    # film 960 x 540, 32 samples over 380-780 nm, RR from depth 3 at 0.05,
    # unlimited bounces (the path-depth limit bounds them), sky background on
    assert (r.width, r.height, r.spp, r.lambda_begin, r.lambda_end) == (960, 540, 32, 380.0, 780.0)
    assert (r.max_diffuse_bounce, r.max_glossy_bounce, r.max_transmission_bounce) == (255, 255, 255)
    assert (r.rr_min_depth, r.seed, r.sky_background) == (3, 1, True)
    assert s.camera == pt_scene.Camera((0.0, 0.0, 5.0), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 30.0)
    assert s.objects[0].double_sided  # clipped planes default to double-sided


def test_sun_and_directional_lights(tmp_path):
    # hosek_wilkie_skylight's sun at its defaults (create_sun TRUE, power
    # sun_intensity_scale 3.14) and directional_light chunks in file order
    p = tmp_path / "sun.RISEscene"
    p.write_text(BASE + "hosek_wilkie_skylight\n{\n\tsolar_elevation 60\n}\n"
                 "directional_light\n{\n\tdirection 0 1 1\n\tcolor 1 0.5 0.25\n\tpower 2\n}\n"
                 "directional_light\n{\n\tdirection 1 0 0\n}\n")
    f = SceneFile(p)
    assert f.info()["n_directional_lights"] == 2
    s = f.scene()
    assert s.sky == pt_scene.Sky(60.0, 0.0, 3.0, 1.0, create_sun=True, sun_power=pytest.approx(3.14))
    assert s.lights == [pt_scene.DirectionalLight((0.0, 1.0, 1.0), (1.0, 0.5, 0.25), 2.0),
                        pt_scene.DirectionalLight((1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 1.0)]


def test_missing_header_and_file(tmp_path):
    p = tmp_path / "x.RISEscene"
    p.write_text(BASE.replace("RISE ASCII SCENE 7\n", ""))
    with pytest.raises(RiseHipError, match="RISE ASCII SCENE"):
        SceneFile(p)
    with pytest.raises(RiseHipError, match="cannot open"):
        SceneFile(tmp_path / "missing.RISEscene")
