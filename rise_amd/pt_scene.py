"""Spectral path-tracing scene description and the seeded test scenes.

A scene is a list of step spectra (the reference's ``spectral_painter``:
a SpectralPacket of n equal bins over [begin, end]), Lambertian materials
(optionally emitting, as ``lambertian_luminaire_material``), world-space
triangle-mesh objects and a pinhole camera.  ``cornell_box()`` builds the
Cornell-box render workload (closed box of 5 quads, 2 axis-aligned boxes,
a 0.5 m^2 area light of radiance 15, 256x256, 16 spp, depth 5, Russian
roulette from depth 3, 4 hero wavelengths over 380-720 nm).

RGB albedos (white 0.75, red 0.63/0.06/0.05, green 0.14/0.45/0.09) are
mapped to 3-band step spectra over [380, 720] nm (bands of 113.3 nm: blue,
green, red), because the reference's RGB-to-spectrum table is not available
(its RGB painters evaluate to a constant 0.5 without it); scenes for the
reference therefore have to be authored with spectral painters, and this is
the mapping used on every side.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Spd:
    """A step spectrum (n bins over [begin, end]) or, with sigmoid True, a
    sigmoid-polynomial spectrum from RGB uplift: amp = (c0, c1, c2, scale),
    S = scale * sigmoid(c0 t^2 + c1 t + c2), t = (nm - 380) / 400, zero
    outside [380, 780] nm."""
    begin: float
    end: float
    amp: np.ndarray
    sigmoid: bool = False

    @staticmethod
    def flat(v: float, begin: float = 380.0, end: float = 720.0) -> "Spd":
        return Spd(begin, end, np.array([v], np.float32))

    @staticmethod
    def rgb_bands(r: float, g: float, b: float, begin: float = 380.0, end: float = 720.0) -> "Spd":
        return Spd(begin, end, np.array([b, g, r], np.float32))


LAMBERTIAN, GGX, DIELECTRIC = 0, 1, 2

# Sampler of the bounce loops (rise_hip_render_desc.sampler): Z-order
# Owen-scrambled Sobol, the reference's default (blue_noise_sampler TRUE), or
# the per-pixel Owen-scrambled Sobol (blue_noise_sampler FALSE), both with the
# sample index shuffled per dimension pair; or the Z-order sampler exactly as
# the reference draws it (unshuffled: biased by about 1% on the spectral
# furnace, see tests/test_oracle_pt.py).
SAMPLER_ZSOBOL, SAMPLER_SOBOL, SAMPLER_ZSOBOL_UNSHUFFLED = 0, 1, 2


@dataclass
class Material:
    """A material; SPD indices refer to Scene.spds (-1: none).

    * LAMBERTIAN (``lambertian_material``): reflectance, optional emission
      (``lambertian_luminaire_material``: L = spd / pi * emission_scale).
    * GGX (``ggx_material``, conductor Fresnel mode, isotropic): reflectance =
      diffuse ``rd`` (-1: none), specular = ``rs``, ior / extinction = the
      conductor's n and k, alpha = roughness.
    * DIELECTRIC (``dielectric_material``, smooth): ior, extinction = ``tau``
      (transmittance per unit length, -1: none)."""
    reflectance: int = -1
    emission: int = -1
    emission_scale: float = 1.0
    kind: int = LAMBERTIAN
    specular: int = -1
    ior: int = -1
    extinction: int = -1
    alpha: float = 0.15


@dataclass
class Object:
    """A triangle-mesh object (``standard_object``): `medium` indexes
    Scene.media for its interior (``interior_medium``, -1: none; entered
    through a dielectric boundary); `casts_shadows` False lets shadow rays
    through it (``casts_shadows FALSE``), the usual setting for a medium's
    shell.  Medium shells must be single-sided and closed, with outward
    normals: shadow rays enter and leave media by the face orientation."""
    tris: np.ndarray            # (n, 9) float32
    material: int
    double_sided: bool = False
    medium: int = -1
    casts_shadows: bool = True


@dataclass
class Medium:
    """``homogeneous_medium`` with an ``hg`` phase function: RGB absorption
    and scattering coefficients (per unit length; the single-wavelength
    coefficients are their luminances) and the Henyey-Greenstein asymmetry g."""
    sigma_a: tuple = (0.05, 0.05, 0.05)
    sigma_s: tuple = (0.5, 0.5, 0.5)
    g: float = 0.3


@dataclass
class Camera:
    position: tuple
    look_at: tuple
    up: tuple = (0.0, 1.0, 0.0)
    fov_deg: float = 40.0


@dataclass
class Sky:
    """``hosek_wilkie_skylight``: the spectral sky radiance map
    (HosekWilkieSpectralRadianceMap; the reference's model is Preetham 1999) --
    sun elevation / azimuth in degrees, turbidity, radiance scale -- sampled by
    next-event estimation, and its matched directional sun light (create_sun,
    colour (1, 1, 1), power sun_intensity_scale), as the chunk's defaults have it."""
    elevation_deg: float = 45.0
    azimuth_deg: float = 0.0
    turbidity: float = 3.0
    scale: float = 1.0
    create_sun: bool = True
    sun_power: float = 3.14


@dataclass
class DirectionalLight:
    """``directional_light``: light arriving from `direction` (toward the
    light) with RGB `color` times `power`."""
    direction: tuple
    color: tuple = (1.0, 1.0, 1.0)
    power: float = 1.0


@dataclass
class RenderSettings:
    width: int
    height: int
    spp: int
    lambda_begin: float = 380.0
    lambda_end: float = 720.0
    max_diffuse_bounce: int = 5
    max_glossy_bounce: int = 5
    max_transmission_bounce: int = 5
    rr_min_depth: int = 3
    rr_threshold: float = 0.05
    seed: int = 1
    sky_background: bool = True   # camera rays that leave the scene see the sky
    sampler: int = 0              # SAMPLER_ZSOBOL (the default), SAMPLER_SOBOL or SAMPLER_ZSOBOL_UNSHUFFLED


@dataclass
class Scene:
    spds: list = field(default_factory=list)
    materials: list = field(default_factory=list)
    objects: list = field(default_factory=list)
    camera: Camera | None = None
    sky: Sky | None = None
    media: list = field(default_factory=list)
    lights: list = field(default_factory=list)   # DirectionalLight

    def flat(self):
        """All triangles (n, 9) float32 and their object index (n,) int32."""
        tris = np.concatenate([o.tris for o in self.objects], 0).astype(np.float32)
        obj = np.concatenate([np.full(len(o.tris), k, np.int32) for k, o in enumerate(self.objects)])
        return np.ascontiguousarray(tris), obj

    @property
    def n_tris(self) -> int:
        return int(sum(len(o.tris) for o in self.objects))


def quad(p0, p1, p2, p3) -> np.ndarray:
    """Two triangles (p0, p1, p2), (p0, p2, p3); face normal = (p1-p0) x (p2-p0)."""
    p0, p1, p2, p3 = (np.asarray(p, np.float64) for p in (p0, p1, p2, p3))
    return np.array([np.concatenate([p0, p1, p2]), np.concatenate([p0, p2, p3])], np.float32)


def box(lo, hi) -> np.ndarray:
    """Axis-aligned box, 12 triangles, outward normals."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = [
        quad((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0)),  # -x
        quad((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1)),  # +x
        quad((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)),  # -y
        quad((x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0)),  # +y
        quad((x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0)),  # -z
        quad((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)),  # +z
    ]
    return np.concatenate(faces, 0)


def cornell_box():
    """The Cornell-box workload: (Scene, RenderSettings).

    Box x in [-1, 1], y in [0, 2], z in [-1, 1], open toward +z where the
    camera looks in; inward normals.  10 wall triangles, 24 box triangles,
    2 light triangles (36 in all)."""
    s = Scene()
    s.spds = [
        Spd.flat(0.75),                      # 0 white
        Spd.rgb_bands(0.63, 0.06, 0.05),     # 1 red
        Spd.rgb_bands(0.14, 0.45, 0.09),     # 2 green
        Spd.flat(15.0),                      # 3 emission: L = 15 / pi * pi = 15
    ]
    s.materials = [
        Material(0),                                         # 0 white
        Material(1),                                         # 1 red
        Material(2),                                         # 2 green
        Material(0, emission=3, emission_scale=math.pi),     # 3 light
    ]
    floor = quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1))          # +y
    ceiling = quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1))        # -y
    back = quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1))         # +z
    left = quad((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1))         # +x
    right = quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1))            # -x
    h = math.sqrt(0.5) / 2.0
    light = quad((-h, 1.98, -h), (h, 1.98, -h), (h, 1.98, h), (-h, 1.98, h))  # -y, area 0.5
    s.objects = [
        Object(np.concatenate([floor, ceiling, back]), 0),
        Object(left, 1),
        Object(right, 2),
        Object(box((0.05, 0.0, -0.05), (0.65, 0.6, 0.55)), 0),
        Object(box((-0.65, 0.0, -0.6), (-0.05, 1.2, 0.0)), 0),
        Object(light, 3),
    ]
    s.camera = Camera(position=(0.0, 1.0, 3.4), look_at=(0.0, 1.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg=40.0)
    r = RenderSettings(width=256, height=256, spp=16)
    return s, r


def furnace_box(albedo=0.5, radiance=1.0, max_bounce: int = 64):
    """White furnace: a closed box whose six faces (six objects, so each face is
    its own luminary) all emit `radiance` and reflect `albedo` (each a number
    for a flat spectrum, or an Spd).  From inside, L(nm) = Le(nm) / (1 -
    albedo(nm)) up to the albedo^(max_bounce + 1) term (the last vertex keeps
    only the NEE share of its MIS pair), so with a large cap the render checks
    MIS, NEE, Russian roulette and the Y normalization (E[Y] = L for a flat
    spectrum; E[Y] = the ybar-weighted mean of L(nm) over the render's
    wavelength range otherwise) against a closed form."""
    s = Scene()
    s.spds = [albedo if isinstance(albedo, Spd) else Spd.flat(albedo),
              radiance if isinstance(radiance, Spd) else Spd.flat(radiance)]
    s.materials = [Material(0, emission=1, emission_scale=math.pi)]
    tris = box((-1, -1, -1), (1, 1, 1))
    # inward normals: swap two vertices of every outward triangle
    tris = np.ascontiguousarray(tris.reshape(-1, 3, 3)[:, [0, 2, 1], :].reshape(-1, 9))
    s.objects = [Object(tris[2 * f: 2 * f + 2].copy(), 0) for f in range(6)]
    s.camera = Camera(position=(0.1, 0.05, 0.2), look_at=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov_deg=60.0)
    r = RenderSettings(width=16, height=16, spp=64, max_diffuse_bounce=max_bounce, max_glossy_bounce=0,
                       max_transmission_bounce=0)
    return s, r


def cloud_cornell(seed: int, n_extra: int):
    """A Cornell-box-style scene with `n_extra` seeded random triangles added:
    the box of ``cornell_box()`` with per-scene albedos, plus a cloud of small
    double-sided triangles (centroids uniform over the upper-middle of the
    box, edge vectors Gaussian with a scale that keeps the cloud's total area
    near 1.5 m^2 whatever its count).  Returns (Scene, RenderSettings) with
    depth 6."""
    rng = np.random.default_rng(seed)
    s, r = cornell_box()
    s.spds[0] = Spd.flat(float(rng.uniform(0.5, 0.8)))
    s.spds[1] = Spd.rgb_bands(*rng.uniform(0.05, 0.7, 3))
    s.spds[2] = Spd.rgb_bands(*rng.uniform(0.05, 0.7, 3))
    s.spds.append(Spd.rgb_bands(*rng.uniform(0.2, 0.8, 3)))
    s.materials.append(Material(len(s.spds) - 1))
    if n_extra > 0:
        c = rng.uniform((-0.8, 0.7, -0.8), (0.8, 1.7, 0.8), (n_extra, 3))
        sigma = math.sqrt(3.0 / n_extra)
        e1 = rng.normal(0.0, sigma, (n_extra, 3))
        e2 = rng.normal(0.0, sigma, (n_extra, 3))
        v0 = c - (e1 + e2) / 3.0
        tris = np.concatenate([v0, v0 + e1, v0 + e2], 1).astype(np.float32)
        s.objects.insert(len(s.objects) - 1, Object(tris, len(s.materials) - 1, double_sided=True))
    r.max_diffuse_bounce = 6
    r.max_glossy_bounce = 6
    r.max_transmission_bounce = 6
    r.seed = seed
    return s, r


def batch_workload(seed: int = 4, n_instances: int = 64, n_scenes: int = 8):
    """The batched-rendering workload: `n_scenes` seeded ``cloud_cornell``
    scenes with 10k-500k added triangles (log-uniform), and `n_instances`
    render requests, each over a scene drawn uniformly, at a resolution drawn
    from {128, 256, 512, 1024}^2 and spp from {1, 4, 16, 64}, depth 6.
    Returns (scenes, [(scene index, RenderSettings)])."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(1e4), math.log(5e5), n_scenes)).astype(np.int64)
    scenes = [cloud_cornell(seed * 1000 + k, int(counts[k]))[0] for k in range(n_scenes)]
    reqs = []
    for i in range(n_instances):
        k = int(rng.integers(n_scenes))
        res = int(rng.choice([128, 256, 512, 1024]))
        spp = int(rng.choice([1, 4, 16, 64]))
        r = RenderSettings(width=res, height=res, spp=spp, max_diffuse_bounce=6, max_glossy_bounce=6,
                           max_transmission_bounce=6, seed=seed * 100000 + i)
        reqs.append((k, r))
    return scenes, reqs


def icosphere(center=(0.0, 0.0, 0.0), radius: float = 1.0, subdiv: int = 2) -> np.ndarray:
    """Icosahedron subdivided `subdiv` times and projected to the sphere:
    20 * 4^subdiv triangles (n, 9) float32, outward face normals."""
    unit = _unit_icosphere(subdiv)
    c = np.tile(np.asarray(center, np.float64), 3)
    return np.ascontiguousarray((unit * radius + c).astype(np.float32))


_ICO_CACHE: dict = {}


def _unit_icosphere(subdiv: int) -> np.ndarray:
    if subdiv in _ICO_CACHE:
        return _ICO_CACHE[subdiv]
    center, radius = (0.0, 0.0, 0.0), 1.0
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius + np.asarray(center, np.float64)
    tris = np.array([np.concatenate([V[a], V[b], V[c]]) for a, b, c in faces], np.float64)
    # orient outward
    cen = tris.reshape(-1, 3, 3).mean(1)
    n = np.cross(tris[:, 3:6] - tris[:, 0:3], tris[:, 6:9] - tris[:, 0:3])
    flip = np.sum(n * cen, 1) < 0
    tris[flip] = tris[flip].reshape(-1, 3, 3)[:, [0, 2, 1], :].reshape(-1, 9)
    _ICO_CACHE[subdiv] = tris
    return tris


def furnace_with(kind: int, radiance: float = 1.0, alpha: float = 0.3, ior: float = 1.5, max_bounce: int = 64):
    """A closed emitting box (radiance `radiance`, albedo 0, every face its own
    luminary) around a sphere of radius 0.5 at the origin: a smooth dielectric
    (index `ior`, no absorption) or a GGX mirror-like conductor (rd = 0,
    rs = 1, n = 1e-3, k = 1e6, so F ~ 1, roughness `alpha`).  Seen through the
    camera inside the box the sphere has radiance = `radiance` x its albedo,
    which is 1 for the dielectric and ~1 for the compensated GGX lobe."""
    s = Scene()
    s.spds = [Spd.flat(0.0), Spd.flat(radiance), Spd.flat(1.0), Spd.flat(ior), Spd.flat(1e-3), Spd.flat(1e6)]
    s.materials = [Material(0, emission=1, emission_scale=math.pi)]
    tris = box((-2, -2, -2), (2, 2, 2))
    tris = np.ascontiguousarray(tris.reshape(-1, 3, 3)[:, [0, 2, 1], :].reshape(-1, 9))
    s.objects = [Object(tris[2 * f: 2 * f + 2].copy(), 0) for f in range(6)]
    if kind == DIELECTRIC:
        s.materials.append(Material(kind=DIELECTRIC, ior=3))
    else:
        s.materials.append(Material(kind=GGX, reflectance=-1, specular=2, ior=4, extinction=5, alpha=alpha))
    s.objects.append(Object(icosphere((0, 0, 0), 0.5, 3), 1))
    s.camera = Camera(position=(0.0, 0.0, 1.8), look_at=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg=20.0)
    r = RenderSettings(width=12, height=12, spp=64, max_diffuse_bounce=max_bounce, max_glossy_bounce=max_bounce,
                       max_transmission_bounce=max_bounce)
    return s, r


def nested_glass(levels: int = 4, max_bounce: int = 64):
    """The emitting box of furnace_with around `levels` concentric glass
    spheres (radii 0.9 down to 0.3, indices 1.3, 1.45, 1.6, ...): a path that
    refracts into the innermost of 4 nests 4 dielectrics, one more than the
    HIP renderer's index-of-refraction stack holds."""
    s, r = furnace_with(DIELECTRIC, max_bounce=max_bounce)
    s.objects = s.objects[:6]
    s.materials = s.materials[:1]
    for k in range(levels):
        s.spds.append(Spd.flat(1.3 + 0.15 * k))
        s.materials.append(Material(kind=DIELECTRIC, ior=len(s.spds) - 1))
        rad = 0.9 - 0.6 * k / max(1, levels - 1) if levels > 1 else 0.9
        s.objects.append(Object(icosphere((0, 0, 0), rad, 3), len(s.materials) - 1))
    s.camera = Camera(position=(0.0, 0.0, 1.8), look_at=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg=30.0)
    r.width = r.height = 16
    return s, r


def medium_shell(lo, hi, medium: int, material: int) -> Object:
    """A closed box shell (outward normals) holding medium `medium`, with a
    dielectric boundary `material` and no shadows of its own."""
    return Object(box(lo, hi), material, medium=medium, casts_shadows=False)


def furnace_medium(sigma_a=(0.0, 0.0, 0.0), sigma_s=(0.8, 0.8, 0.8), g: float = 0.3, half: float = 0.5,
                   max_bounce: int = 64):
    """A closed emitting box (radiance 1, albedo 0, every face its own
    luminary) around a cube of half-size `half` at the origin filled with a
    homogeneous medium behind an index-matched (index 1) dielectric shell.
    A purely scattering medium leaves the uniform radiance field unchanged
    (radiance 1 everywhere); a purely absorbing one shows exp(-sigma_t * chord)."""
    s = Scene()
    s.spds = [Spd.flat(0.0), Spd.flat(1.0)]
    s.materials = [Material(0, emission=1, emission_scale=math.pi), Material(kind=DIELECTRIC, ior=1)]
    tris = box((-2, -2, -2), (2, 2, 2))
    tris = np.ascontiguousarray(tris.reshape(-1, 3, 3)[:, [0, 2, 1], :].reshape(-1, 9))
    s.objects = [Object(tris[2 * f: 2 * f + 2].copy(), 0) for f in range(6)]
    s.media = [Medium(tuple(sigma_a), tuple(sigma_s), g)]
    s.objects.append(medium_shell((-half, -half, -half), (half, half, half), 0, 1))
    s.camera = Camera(position=(0.0, 0.0, 1.8), look_at=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg=30.0)
    r = RenderSettings(width=12, height=12, spp=64, max_diffuse_bounce=max_bounce, max_glossy_bounce=max_bounce,
                       max_transmission_bounce=max_bounce)
    return s, r


def cornell_fog(sigma_a=(0.05, 0.05, 0.05), sigma_s=(0.5, 0.5, 0.5), g: float = 0.3):
    """The Cornell box with a floating box of homogeneous medium (sigma_s 0.5,
    sigma_a 0.05, g 0.3 by default) behind an index-matched dielectric shell
    that casts no shadow, above the short block: x in [0, 0.9], y in [0.7, 1.7],
    z in [-0.5, 0.6]."""
    s, r = cornell_box()
    s.spds.append(Spd.flat(1.0))
    s.materials.append(Material(kind=DIELECTRIC, ior=len(s.spds) - 1))
    s.media = [Medium(tuple(sigma_a), tuple(sigma_s), g)]
    s.objects.insert(5, medium_shell((0.0, 0.7, -0.5), (0.9, 1.7, 0.6), 0, len(s.materials) - 1))
    return s, r


def cornell_materials(glass_subdiv: int = 2):
    """The Cornell box with the tall block made a GGX conductor (gold-like
    n, k over three bands, roughness 0.25, white diffuse 0 / specular 1) and a
    smooth glass sphere (index 1.5, radius 0.3, 20 * 4^glass_subdiv
    triangles) on the floor in front of the short block."""
    s, r = cornell_box()
    s.spds += [
        Spd.flat(1.0),                                   # 4 specular tint
        Spd(380.0, 720.0, np.array([1.3, 0.6, 0.2], np.float32)),   # 5 conductor n (blue, green, red bands)
        Spd(380.0, 720.0, np.array([1.9, 2.4, 3.2], np.float32)),   # 6 conductor k
        Spd.flat(1.5),                                   # 7 glass index
    ]
    s.materials += [
        Material(kind=GGX, reflectance=-1, specular=4, ior=5, extinction=6, alpha=0.25),   # 4 gold-like
        Material(kind=DIELECTRIC, ior=7),                                                  # 5 glass
    ]
    s.objects[4] = Object(s.objects[4].tris, 4)
    s.objects.insert(5, Object(icosphere((0.35, 0.3, 0.85), 0.3, glass_subdiv), 5))
    return s, r


def sky_spheres(n_side: int = 8, seed: int = 3):
    """A ground quad under the spectral sky with a grid of small spheres on
    it, alternately GGX conductors (roughness drawn from U[0.05, 0.5]) and
    smooth glass (index 1.5): a reduced version of the terrain-and-spheres
    workload for parity tests.  No area lights: the sky is the only source."""
    rng = np.random.default_rng(seed)
    s = Scene()
    s.spds = [Spd.flat(0.4), Spd.flat(1.0), Spd(380.0, 720.0, np.array([1.3, 0.6, 0.2], np.float32)),
              Spd(380.0, 720.0, np.array([1.9, 2.4, 3.2], np.float32)), Spd.flat(1.5)]
    s.materials = [Material(0), Material(kind=DIELECTRIC, ior=4)]
    s.objects = [Object(quad((-6, 0, -6), (-6, 0, 6), (6, 0, 6), (6, 0, -6)), 0)]
    for i in range(n_side):
        for j in range(n_side):
            c = (-4 + 8 * (i + 0.5) / n_side, 0.35, -4 + 8 * (j + 0.5) / n_side)
            if (i + j) % 2 == 0:
                s.materials.append(Material(kind=GGX, reflectance=-1, specular=1, ior=2, extinction=3,
                                            alpha=float(rng.uniform(0.05, 0.5))))
                s.objects.append(Object(icosphere(c, 0.35, 1), len(s.materials) - 1))
            else:
                s.objects.append(Object(icosphere(c, 0.35, 1), 1))
    s.camera = Camera(position=(0.0, 4.0, 9.0), look_at=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg=45.0)
    s.sky = Sky(elevation_deg=40.0, azimuth_deg=30.0, turbidity=3.0, scale=0.05)
    r = RenderSettings(width=64, height=48, spp=16, max_diffuse_bounce=8, max_glossy_bounce=8,
                       max_transmission_bounce=8)
    return s, r


def value_noise(nx: int, ny: int, octaves: int = 6, seed: int = 7) -> np.ndarray:
    """Seeded value noise on an nx x ny grid: the sum over octaves k of
    bilinearly interpolated random lattices of 4 * 2^k cells, amplitude 2^-k,
    normalized to [-1, 1]."""
    rng = np.random.default_rng(seed)
    u = np.linspace(0.0, 1.0, nx)
    v = np.linspace(0.0, 1.0, ny)
    h = np.zeros((ny, nx))
    amp, total = 1.0, 0.0
    for k in range(octaves):
        cells = 4 * (1 << k)
        lat = rng.uniform(-1.0, 1.0, (cells + 2, cells + 2))
        x = u * cells
        y = v * cells
        x0 = np.floor(x).astype(int)
        y0 = np.floor(y).astype(int)
        fx = x - x0
        fy = y - y0
        fx = fx * fx * (3 - 2 * fx)
        fy = fy * fy * (3 - 2 * fy)
        a = lat[np.ix_(y0, x0)]
        b = lat[np.ix_(y0, x0 + 1)]
        c = lat[np.ix_(y0 + 1, x0)]
        d = lat[np.ix_(y0 + 1, x0 + 1)]
        top = a + (b - a) * fx[None, :]
        bot = c + (d - c) * fx[None, :]
        h += amp * (top + (bot - top) * fy[:, None])
        total += amp
        amp *= 0.5
    return h / total


def heightfield(nx: int, ny: int, extent=(20.0, 10.0), amplitude: float = 0.2, seed: int = 7) -> np.ndarray:
    """Triangulated heightfield over an nx x ny vertex grid centred at the
    origin (x across, z along, y up): 2 (nx-1)(ny-1) triangles, upward normals."""
    h = amplitude * value_noise(nx, ny, 6, seed)
    xs = np.linspace(-extent[0] / 2, extent[0] / 2, nx)
    zs = np.linspace(-extent[1] / 2, extent[1] / 2, ny)
    X, Z = np.meshgrid(xs, zs)
    P = np.stack([X, h, Z], -1).astype(np.float32)
    p00, p10, p01, p11 = P[:-1, :-1], P[:-1, 1:], P[1:, :-1], P[1:, 1:]
    # (p00, p01, p10) and (p10, p01, p11): normals +y for x across, z along
    t1 = np.concatenate([p00, p01, p10], -1).reshape(-1, 9)
    t2 = np.concatenate([p10, p01, p11], -1).reshape(-1, 9)
    return np.ascontiguousarray(np.concatenate([t1, t2], 0))


def sky_plane(albedo=0.5, elevation_deg: float = 40.0, azimuth_deg: float = 30.0, turbidity: float = 3.0,
              scale: float = 0.1, create_sun: bool = True, sun_power: float = 3.14, lights=()):
    """A Lambertian ground quad (100 m square, `albedo` a number or an Spd)
    under the spectral sky, seen straight down from 1 m through a narrow
    lens, so every camera ray meets the ground and every light path leaves
    it for the sky: the pixel radiance is albedo / pi times the ground's
    irradiance from the sky, the sun and `lights` (DirectionalLight), a
    closed form that checks the environment light sample, its MIS with the
    BSDF sample that escapes, and the directional lights' deterministic
    sample."""
    s = Scene()
    s.spds = [albedo if isinstance(albedo, Spd) else Spd.flat(albedo)]
    s.materials = [Material(0)]
    s.objects = [Object(quad((-50, 0, -50), (-50, 0, 50), (50, 0, 50), (50, 0, -50)), 0)]
    s.camera = Camera(position=(0.0, 1.0, 0.0), look_at=(0.0, 0.0, 0.0), up=(0.0, 0.0, -1.0), fov_deg=20.0)
    s.sky = Sky(elevation_deg=elevation_deg, azimuth_deg=azimuth_deg, turbidity=turbidity, scale=scale,
                create_sun=create_sun, sun_power=sun_power)
    s.lights = list(lights)
    r = RenderSettings(width=16, height=16, spp=64, max_diffuse_bounce=4, max_glossy_bounce=0,
                       max_transmission_bounce=0)
    return s, r


def terrain_spheres(nx: int = 2048, ny: int = 1024, subdiv: int = 4, seed: int = 7, n_side: int = 8):
    """The terrain-and-spheres workload: a 2048 x 1024 heightfield of seeded
    6-octave value noise (amplitude 0.2; 4,188,162 triangles) with 64
    icospheres (subdivision 4, 5,120 triangles each) resting on it, alternately
    GGX conductors (roughness U[0.05, 0.5]) and smooth glass (index 1.5),
    under the spectral sky; 1920 x 1080, 64 spp, depth 8.  Smaller grids,
    subdivisions and n_side x n_side spheres give reduced scenes from the
    same generator."""
    rng = np.random.default_rng(seed)
    s = Scene()
    s.spds = [Spd.rgb_bands(0.25, 0.3, 0.18), Spd.flat(1.0), Spd(380.0, 720.0, np.array([1.3, 0.6, 0.2], np.float32)),
              Spd(380.0, 720.0, np.array([1.9, 2.4, 3.2], np.float32)), Spd.flat(1.5)]
    s.materials = [Material(0), Material(kind=DIELECTRIC, ior=4)]
    s.objects = [Object(heightfield(nx, ny, seed=seed), 0)]
    for i in range(n_side):
        for j in range(n_side):
            c = (-8.0 + 16.0 * (i + 0.5) / n_side, 0.55, -4.0 + 8.0 * (j + 0.5) / n_side)
            if (i + j) % 2 == 0:
                s.materials.append(Material(kind=GGX, reflectance=-1, specular=1, ior=2, extinction=3,
                                            alpha=float(rng.uniform(0.05, 0.5))))
                s.objects.append(Object(icosphere(c, 0.4, subdiv), len(s.materials) - 1))
            else:
                s.objects.append(Object(icosphere(c, 0.4, subdiv), 1))
    s.camera = Camera(position=(0.0, 5.0, 12.0), look_at=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_deg=50.0)
    s.sky = Sky(elevation_deg=35.0, azimuth_deg=40.0, turbidity=3.0, scale=2e-4)
    r = RenderSettings(width=1920, height=1080, spp=64, max_diffuse_bounce=8, max_glossy_bounce=8,
                       max_transmission_bounce=8)
    return s, r


def preview_requests(n_scenes: int, n_requests: int = 256, seed: int = 6):
    """Small render requests over `n_scenes` resident scenes (previews and
    thumbnails: 64^2 or 128^2 pixels, 1, 2 or 4 spp, depth 6), as
    [(scene index, RenderSettings)]: the request sizes where one-at-a-time
    rendering is bound by per-request overhead rather than by the GPU."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_requests):
        res = int(rng.choice([64, 128]))
        r = RenderSettings(width=res, height=res, spp=int(rng.choice([1, 2, 4])), max_diffuse_bounce=6,
                           max_glossy_bounce=6, max_transmission_bounce=6, seed=seed * 100000 + i)
        out.append((int(rng.integers(n_scenes)), r))
    return out


def poisson_arrivals(n: int, rate: float, seed: int = 7) -> np.ndarray:
    """Arrival times (s) of a Poisson stream of n requests at `rate` per second."""
    return np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, n))


def serving_workload(seed: int = 5, n_requests: int = 64, n_scenes: int = 16, rate: float = 200.0):
    """The serving workload: `n_scenes` resident scenes of 50k-2M triangles
    (log-uniform; Cornell-style boxes with a seeded triangle cloud, every third
    cloud a GGX conductor, every third scene also holding a glass sphere and a
    GGX block, every other scene a box of ground fog -- homogeneous medium
    sigma_s 0.5, sigma_a 0.05, g 0.3 behind an index-matched shell that casts
    no shadow, over the floor up to y = 0.6) and a Poisson stream of
    `n_requests` render requests at `rate` per second, each over a scene drawn
    uniformly, at a resolution drawn from {128, 256, 512, 1024}^2 and spp from
    {1, 4, 16, 64}.  Returns (scenes, [(arrival s, scene index, RenderSettings)])."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(5e4), math.log(2e6), n_scenes)).astype(np.int64)
    scenes = []
    for k in range(n_scenes):
        s, _ = cloud_cornell(seed * 1000 + k, int(counts[k]))
        if k % 3 == 1:
            s.spds += [Spd.flat(1.0), Spd.flat(0.2), Spd.flat(3.0)]
            s.materials.append(Material(kind=GGX, reflectance=-1, specular=len(s.spds) - 3, ior=len(s.spds) - 2,
                                        extinction=len(s.spds) - 1, alpha=0.3))
            s.objects[-2] = Object(s.objects[-2].tris, len(s.materials) - 1, double_sided=True)
        elif k % 3 == 2:
            s.spds += [Spd.flat(1.0), Spd(380.0, 720.0, np.array([1.3, 0.6, 0.2], np.float32)),
                       Spd(380.0, 720.0, np.array([1.9, 2.4, 3.2], np.float32)), Spd.flat(1.5)]
            n = len(s.spds)
            s.materials += [Material(kind=GGX, reflectance=-1, specular=n - 4, ior=n - 3, extinction=n - 2, alpha=0.25),
                            Material(kind=DIELECTRIC, ior=n - 1)]
            s.objects[4] = Object(s.objects[4].tris, len(s.materials) - 2)
            s.objects.insert(len(s.objects) - 1, Object(icosphere((0.35, 0.3, 0.85), 0.3, 2), len(s.materials) - 1))
        if k % 2 == 0:
            s.spds.append(Spd.flat(1.0))
            s.materials.append(Material(kind=DIELECTRIC, ior=len(s.spds) - 1))
            s.media = [Medium((0.05, 0.05, 0.05), (0.5, 0.5, 0.5), 0.3)]
            s.objects.insert(len(s.objects) - 1,
                             medium_shell((-0.95, 0.02, -0.95), (0.95, 0.6, 0.95), 0, len(s.materials) - 1))
        scenes.append(s)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    reqs = []
    for i in range(n_requests):
        res = int(rng.choice([128, 256, 512, 1024]))
        spp = int(rng.choice([1, 4, 16, 64]))
        r = RenderSettings(width=res, height=res, spp=spp, max_diffuse_bounce=6, max_glossy_bounce=6,
                           max_transmission_bounce=6, seed=seed * 100000 + i)
        reqs.append((float(t[i]), int(rng.integers(n_scenes)), r))
    return scenes, reqs
