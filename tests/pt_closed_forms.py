"""Closed forms for the spectral furnace tests (CPU oracle and GPU).

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
