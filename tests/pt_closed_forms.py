"""Closed forms for the spectral furnace and sky-plane tests (CPU oracle and
GPU).

A closed box whose faces all emit Le(nm) and reflect a Lambertian albedo
rho(nm) has radiance L(nm) = Le(nm) / (1 - rho(nm)) everywhere inside (up to
the rho^(cap + 1) term of the bounce cap).  The HWSS film value of one sample
is mYNormalization / 4 * sum_i ybar(nm_i) L(nm_i) with every nm_i uniform on
[a, b) and mYNormalization = (b - a) / int_a^b ybar
(PixelBasedSpectralIntegratingRasterizer.cpp:53-65,
PathTracingSpectralRasterizer.cpp:156-173), so

    E[Y] = int_a^b ybar(nm) L(nm) dnm / int_a^b ybar(nm) dnm.

The integrals are computed here by the midpoint rule on 0.05 nm cells whose
edges include every step of the spectra and every node of the 5 nm CIE table,
so each cell sees a linear ybar times a constant L: the rule is exact up to
float rounding.  ybar comes from the oracle's CIE table, which is pinned on
its own (the reference's stated integral, its nodes and linear interpolation).
"""
import numpy as np

import oracle

CELL = 0.05


def _grid(a: float, b: float) -> np.ndarray:
    n = int(round((b - a) / CELL))
    return a + (np.arange(n) + 0.5) * CELL


def step_values(spd, nm: np.ndarray) -> np.ndarray:
    """A step spectrum's value: n equal bins over [begin, end], bin index
    floor((nm - begin) / delta) (SpectralPacket::ValueAtNM,
    Utilities/Color/SpectralPacket.h:285-297)."""
    amp = np.asarray(spd.amp, np.float64)
    n = len(amp)
    delta = (spd.end - spd.begin) / n
    idx = np.minimum(np.floor((nm - spd.begin) / delta).astype(np.int64), n - 1)
    out = amp[np.clip(idx, 0, n - 1)]
    return np.where((nm < spd.begin) | (nm > spd.end), 0.0, out)


def ybar(nm: np.ndarray) -> np.ndarray:
    return np.array([oracle.xyz_from_nm(float(x))[1][1] for x in nm])


def furnace_y(albedo, emission, a: float = 380.0, b: float = 720.0) -> float:
    """E[Y] of the spectral furnace, the closed form above."""
    nm = _grid(a, b)
    y = ybar(nm)
    L = step_values(emission, nm) / (1.0 - step_values(albedo, nm))
    return float((y * L).sum() / y.sum())


def furnace_y_hero_only(albedo, emission, a: float = 380.0, b: float = 720.0) -> float:
    """E[Y] an estimator would give if it carried the hero wavelength's
    radiance on every companion (sum_i ybar(nm_i) L(nm_0)): the mistake the
    spectral furnace must be able to tell apart from the closed form."""
    nm0 = _grid(a, b)
    L0 = step_values(emission, nm0) / (1.0 - step_values(albedo, nm0))
    acc = np.zeros_like(nm0)
    for i in range(4):
        nmi = a + np.fmod(nm0 - a + i * (b - a) / 4.0, b - a)
        acc += ybar(nmi)
    y_int = ybar(nm0).sum()
    return float((L0 * acc / 4.0).sum() / y_int)


def spectral_furnace_spectra():
    """Step spectra for the furnace: a 17-bin albedo and a 10-bin emission
    whose bins do not line up, both far from flat."""
    from rise_amd import pt_scene

    rho = pt_scene.Spd(380.0, 720.0, np.array([0.15, 0.8, 0.3, 0.7, 0.2, 0.85, 0.4, 0.6, 0.1, 0.75, 0.5, 0.35,
                                               0.8, 0.25, 0.65, 0.45, 0.55], np.float32))
    le = pt_scene.Spd(380.0, 720.0, np.array([0.5, 2.0, 0.3, 1.5, 3.0, 0.8, 0.2, 2.5, 1.0, 0.4], np.float32))
    return rho, le


def _hemisphere_xyz(sky, n_mu: int = 192, n_phi: int = 384):
    """The sky's cosine-weighted integrals over the upper hemisphere of its
    CIE X, Y and Z: (EX, EY, EZ) with E_sky(nm) = (EX xbar + EY ybar + EZ zbar)
    / (xbar + ybar + zbar) (nm) times the sky's scale -- the form of the
    reference's spectral sky radiance (HosekWilkieSkyModel.cpp:220-256).  The
    per-direction X, Y, Z are read back from the oracle's spectral sky
    radiance (pinned value for value against the reference's model) at three
    wavelengths by solving that 3 x 3 form, so the closed form depends on the
    model only through the pinned function.  Gauss-Legendre in cos(theta),
    midpoint rule in phi."""
    import ctypes

    L = oracle.oracle_lib()
    L.rise_oracle_sky_radiance.restype = ctypes.c_double
    L.rise_oracle_sky_radiance.argtypes = [ctypes.POINTER(oracle._RoScene), ctypes.c_void_p, ctypes.c_double]
    rs = oracle._RoScene()  # the sky's parameters as a render passes them (float32), scale 1
    rs.sky_enabled = 1
    rs.sky_elevation_deg = float(np.float32(sky.elevation_deg))
    rs.sky_azimuth_deg = float(np.float32(sky.azimuth_deg))
    rs.sky_turbidity = float(np.float32(sky.turbidity))
    rs.sky_scale = 1.0
    nms = (450.0, 550.0, 610.0)
    A = np.array([np.array(oracle.xyz_from_nm(nm)[1]) / sum(oracle.xyz_from_nm(nm)[1]) for nm in nms])
    Ainv = np.linalg.inv(A)
    x, w = np.polynomial.legendre.leggauss(n_mu)
    mu, wmu = 0.5 * (x + 1.0), 0.5 * w
    phi = (np.arange(n_phi) + 0.5) * (2.0 * np.pi / n_phi)
    v = (ctypes.c_double * 3)()
    acc = np.zeros(3)
    for m, wm in zip(mu, wmu):
        st = np.sqrt(max(0.0, 1.0 - m * m))
        row = np.zeros(3)
        for p in phi:
            v[0], v[1], v[2] = st * np.sin(p), m, st * np.cos(p)
            Lk = [L.rise_oracle_sky_radiance(ctypes.byref(rs), ctypes.cast(v, ctypes.c_void_p), nm) for nm in nms]
            row += Ainv @ np.array(Lk)
        acc += wm * m * row * (2.0 * np.pi / n_phi)
    return acc


def sky_plane_y(scene, a: float = 380.0, b: float = 720.0, **quad) -> float:
    """E[Y] of the sky-plane scene (pt_scene.sky_plane): albedo(nm) / pi times
    the ground's irradiance, i.e. the ybar-weighted mean over [a, b] of
    albedo(nm) / pi * (E_sky(nm) + sum over directional lights of
    lum(colour) * power * cos(theta_light)), the sun included as a light of
    colour (1, 1, 1) and power sun_intensity_scale (Job.cpp:2736-2746)."""
    nm = _grid(a, b)
    cmf = np.array([oracle.xyz_from_nm(float(x))[1] for x in nm])
    E = np.zeros_like(nm)
    sky = scene.sky
    if sky is not None:
        EX, EY, EZ = _hemisphere_xyz(sky, **quad)
        E += float(np.float32(sky.scale)) * (cmf @ np.array([EX, EY, EZ])) / cmf.sum(axis=1)
    lights = list(scene.lights)
    if sky is not None and sky.create_sun:
        el, az = np.radians(np.float32(sky.elevation_deg)), np.radians(np.float32(sky.azimuth_deg))
        from rise_amd import pt_scene
        lights.append(pt_scene.DirectionalLight((np.cos(el) * np.sin(az), np.sin(el), np.cos(el) * np.cos(az)),
                                                (1.0, 1.0, 1.0), sky.sun_power))
    for lt in lights:
        d = np.asarray(lt.direction, np.float64)
        cos_t = max(0.0, d[1] / np.linalg.norm(d))
        lum = 0.2126 * lt.color[0] + 0.7152 * lt.color[1] + 0.0722 * lt.color[2]
        E += lum * float(np.float32(lt.power)) * cos_t
    rho = scene.spds[scene.materials[0].reflectance]
    f = step_values(rho, nm) / np.pi
    y = cmf[:, 1]
    return float((y * f * E).sum() / y.sum())
