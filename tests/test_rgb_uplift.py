"""RGB to spectrum uplift (include/rise_hip_rgb_uplift.h), no GPU.

* The library regenerates the reference's Jakob-Hanika table: at 16 cells per
  axis it equals, float for float, the table the reference's own generator
  (tools/JakobHanikaLUTGen.cpp, compiled into oracle/_ref) wrote into
  tests/golden/jh_rec709_16.coeff (tests/golden/make_jh_golden.py), read back
  through the ".coeff" reader; at 24 the generator is run live when
  oracle/_ref is built.
* Lookup: at grid points it returns the cell exactly (pins the table layout
  and the z-warp inverse in all three max-channel sub-tables); between them
  the uplifted spectrum reproduces its colour through the forward model (CIE
  1931 at the 5 nm nodes, flat illuminant, XYZ to Rec.709), to the table's
  interpolation error.
* The sigmoid spectrum's closed forms; a uniform RGB painter in a scene file
  becomes that spectrum (sRGB decoded first); the oracle's sigmoid spectra
  give a directly viewed emitter its luminance.
"""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from rise_amd import pt_scene
from rise_amd._lib import RiseHipError
from rise_amd.rgb_uplift import RgbLut, sigmoid_eval, uplift_spd
from rise_amd.scene_file import load_scene

GOLDEN = Path(__file__).resolve().parent / "golden"
M709 = np.array([[3.240479, -1.537150, -0.498535], [-0.969256, 1.875992, 0.041556], [0.055648, -0.204043, 1.057311]])


def _cie_nodes():
    nm = 380.0 + 5.0 * np.arange(81)
    xyz = np.array([oracle.xyz_from_nm(float(x))[1] for x in nm])
    return nm, xyz


def _rgb_of(c, scale=1.0):
    """Forward model: sigmoid spectrum -> linear Rec.709 (flat illuminant, S = 1 -> Y = 1)."""
    nm, xyz = _cie_nodes()
    t = (nm - 380.0) / 400.0
    x = c[0] * t * t + c[1] * t + c[2]
    s = scale * (0.5 + x / (2.0 * np.sqrt(1.0 + x * x)))
    XYZ = (s[:, None] * xyz).sum(0) / xyz[:, 1].sum()
    return M709 @ XYZ


def test_table_equals_reference_generator_output():
    ref = RgbLut.read(GOLDEN / "jh_rec709_16.coeff").table()
    ours = RgbLut.generate(16)
    assert ours.stats["cells"] == 3 * 16 ** 3
    np.testing.assert_array_equal(ours.table(), ref)
    # the reference's own convergence record for this table
    assert (ours.stats["converged"], ours.stats["failed"]) == (11697, 591)


@pytest.mark.skipif(not (oracle.REF_DIR / "jh_lut_gen").exists(), reason="oracle/_ref/jh_lut_gen not built")
def test_table_equals_live_reference_generator(tmp_path):
    out = tmp_path / "r24.coeff"
    r = subprocess.run([str(oracle.REF_DIR / "jh_lut_gen"), "--resolution", "24", "--output", str(out)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode in (0, 2) and out.exists(), r.stderr
    np.testing.assert_array_equal(RgbLut.generate(24).table(), RgbLut.read(out).table())


def test_coeff_file_round_trip_and_errors(tmp_path):
    lut = RgbLut.read(GOLDEN / "jh_rec709_16.coeff")
    lut.write(tmp_path / "x.coeff")
    assert (tmp_path / "x.coeff").read_bytes() == (GOLDEN / "jh_rec709_16.coeff").read_bytes()
    (tmp_path / "bad.coeff").write_bytes(b"RJHX" + bytes(16))
    with pytest.raises(RiseHipError, match="RJHL"):
        RgbLut.read(tmp_path / "bad.coeff")
    (tmp_path / "short.coeff").write_bytes((GOLDEN / "jh_rec709_16.coeff").read_bytes()[:5000])
    with pytest.raises(RiseHipError, match="truncated"):
        RgbLut.read(tmp_path / "short.coeff")
    with pytest.raises(RiseHipError, match="resolution"):
        RgbLut.generate(4)


def test_lookup_at_grid_points_returns_the_cell():
    lut = RgbLut.read(GOLDEN / "jh_rec709_16.coeff")
    tab = lut.table()
    res = 16
    g = np.random.default_rng(2)
    for _ in range(300):
        m, iz, ix, iy = int(g.integers(3)), int(g.integers(1, res)), int(g.integers(res)), int(g.integers(res))
        z = math.sin(0.5 * math.pi * iz / (res - 1)) ** 2
        rgb = [0.0, 0.0, 0.0]
        rgb[m] = z
        rgb[(m + 1) % 3] = ix / (res - 1) * z
        rgb[(m + 2) % 3] = iy / (res - 1) * z
        # ties between channels pick the lowest max-channel index, as the
        # table itself does for those cells' colours
        if ix == res - 1 or iy == res - 1:
            continue
        got = lut.lookup(rgb)
        np.testing.assert_allclose(got, tab[m, iz, ix, iy], rtol=0, atol=2e-6 * (1 + np.abs(tab[m, iz, ix, iy]).max()))


def test_uplift_reproduces_colour():
    lut = RgbLut.default()
    g = np.random.default_rng(7)
    errs = []
    for _ in range(400):
        rgb = g.uniform(0.05, 0.9, 3)
        errs.append(np.abs(_rgb_of(lut.lookup(rgb).astype(np.float64)) - rgb).max())
    errs = np.array(errs)
    # trilinear interpolation between converged cells: small everywhere in
    # this range, and well below an 8-bit step for most colours
    assert np.median(errs) < 2e-3 and errs.max() < 2e-2, (np.median(errs), errs.max())
    # Rec.709 white is D65's colour, not that of a flat spectrum under the
    # flat training illuminant, so greys are not flat; a brighter grey is
    # brighter at every wavelength
    lo, hi = lut.lookup([0.2] * 3), lut.lookup([0.6] * 3)
    for nm in range(380, 781, 20):
        assert sigmoid_eval(lo, 1.0, nm) < sigmoid_eval(hi, 1.0, nm)


def test_sigmoid_closed_forms():
    assert sigmoid_eval([0, 0, 0], 1.0, 550.0) == 0.5
    x = 1.7
    assert abs(sigmoid_eval([0, 0, x], 2.0, 450.0) - 2.0 * (0.5 + x / (2 * math.sqrt(1 + x * x)))) < 1e-7
    # t runs over [0, 1] on [380, 780] nm: c0 t^2 + c1 t at the ends
    assert abs(sigmoid_eval([1.0, -3.0, 0.0], 1.0, 780.0) - (0.5 + -2.0 / (2 * math.sqrt(5.0)))) < 1e-7
    assert sigmoid_eval([0, 0, 5.0], 1.0, 379.9) == 0.0 and sigmoid_eval([0, 0, 5.0], 1.0, 780.1) == 0.0


SCENE = """RISE ASCII SCENE 7
uniformcolor_painter
{
\tname\tlin
\tcolor\t0.6 0.3 0.1
}
uniformcolor_painter
{
\tname\tsrgb
\tcolor\t0.5 0.5 0.5
\tcolorspace\tsRGB
}
lambertian_material
{
\tname\ta
\treflectance\tlin
}
lambertian_material
{
\tname\tb
\treflectance\tsrgb
}
lambertian_material
{
\tname\tc
\treflectance\tlin
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
\tmaterial\ta
}
pinhole_camera
{
\tlocation\t0 0 5
}
"""


def test_scene_file_uniform_painters(tmp_path):
    p = tmp_path / "rgb.RISEscene"
    p.write_text(SCENE)
    s, _ = load_scene(p)
    a, b, c = s.materials
    assert a.reflectance == c.reflectance  # one spectrum per painter
    sa, sb = s.spds[a.reflectance], s.spds[b.reflectance]
    assert sa.sigmoid and sb.sigmoid and sa.amp[3] == 1.0
    lut = RgbLut.default()
    np.testing.assert_array_equal(sa.amp[:3], lut.lookup([0.6, 0.3, 0.1]))
    lin = ((0.5 + 0.055) / 1.055) ** 2.4  # sRGB 0.5 decoded
    np.testing.assert_array_equal(sb.amp[:3], lut.lookup([lin] * 3))
    np.testing.assert_allclose(_rgb_of(sa.amp[:3].astype(np.float64)), [0.6, 0.3, 0.1], atol=5e-3)
    p.write_text(SCENE.replace("\tcolorspace\tsRGB", "\tcolorspace\tACEScg"))
    with pytest.raises(RiseHipError, match="colorspace"):
        load_scene(p)


def test_oracle_sigmoid_emitter_seen_directly():
    s, r = pt_scene.cornell_box()
    c = np.array([2.0, -1.5, 0.3], np.float32)
    s.spds.append(pt_scene.Spd(380.0, 780.0, np.array([c[0], c[1], c[2], 3.0], np.float32), sigmoid=True))
    s.spds.append(pt_scene.Spd.flat(0.0))
    s.materials[3].emission = len(s.spds) - 2
    s.materials[3].emission_scale = math.pi
    s.materials[3].reflectance = len(s.spds) - 1
    s.camera = pt_scene.Camera(position=(0.0, 1.0, 0.0), look_at=(0.0, 2.0, 0.0), up=(0.0, 0.0, 1.0), fov_deg=10.0)
    r.width = r.height = 8
    r.spp = 512
    r.lambda_begin, r.lambda_end = 380.0, 780.0
    sm, sq, _ = oracle.render(s, r)
    n = r.width * r.height * r.spp
    y = sm[..., 1].sum() / n
    se = math.sqrt(max(sq[..., 1].sum() / n - y * y, 0) / n)
    # Expected: the luminance of 3 * S under a flat illuminant (Y of S = 1 is 1).
    nm = np.linspace(380.0, 780.0, 4001)
    ys = np.array([oracle.xyz_from_nm(float(v))[1][1] for v in nm])
    t = (nm - 380.0) / 400.0
    x = c[0] * t * t + c[1] * t + c[2]
    sv = 3.0 * (0.5 + x / (2.0 * np.sqrt(1.0 + x * x)))
    want = np.trapezoid(sv * ys, nm) / np.trapezoid(ys, nm)
    assert abs(y - want) < 4 * se + 1e-3, (y, want, se)
