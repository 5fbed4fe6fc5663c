"""RISE scene files (native "RISE ASCII SCENE 7" text) for the spectral path
tracer, over the C ABI of include/rise_hip_scene_file.h.

``SceneFile(path)`` parses a file with the library's reader (host code, no
GPU), the counterpart of the reference's ``Job::LoadAsciiSceneViaCst``
(Job.h:2843; Cst::ParseToCst + Cst::DeriveToJob, chunk parsers in
Parsers/ChunkParserRegistry.cpp).  ``scene()`` / ``settings()`` return the
parsed tables as a ``pt_scene.Scene`` and ``RenderSettings``;
``rasterizer()`` makes the scene resident through
``rise_hip_scene_file_upload``.  Errors raise RiseHipError.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from .pt_scene import Camera, DirectionalLight, Material, Medium, Object, RenderSettings, Scene, Sky, Spd
from .output import OutputDesc, extension
from .render import (DirectionalLightDesc, MaterialDesc, MediumDesc, ObjectDesc, ObjectMedia,
                     PathTracingSpectralRasterizer, RenderDesc, SkyDesc, SpdDesc)


class SceneFileInfo(ctypes.Structure):
    _fields_ = [
        ("n_objects", ctypes.c_int64),
        ("n_materials", ctypes.c_int64),
        ("n_spds", ctypes.c_int64),
        ("n_media", ctypes.c_int64),
        ("n_tris", ctypes.c_int64),
        ("has_sky", ctypes.c_int32),
        ("n_skipped_chunks", ctypes.c_int32),
        ("n_outputs", ctypes.c_int32),
        ("n_directional_lights", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = ctypes.c_void_p
_PP = ctypes.POINTER(_P)
_SIG = {
    "rise_hip_scene_file_parse": ([ctypes.c_char_p, _PP], ctypes.c_int),
    "rise_hip_scene_file_get_info": ([_P, ctypes.POINTER(SceneFileInfo)], ctypes.c_int),
    "rise_hip_scene_file_get_tables": ([_P, _PP, _PP, _PP, _PP, _PP, _PP], ctypes.c_int),
    "rise_hip_scene_file_get_render_desc": ([_P, ctypes.POINTER(RenderDesc)], ctypes.c_int),
    "rise_hip_scene_file_upload": ([_P, _PP], ctypes.c_int),
    "rise_hip_scene_file_get_output": ([_P, ctypes.c_int32, ctypes.POINTER(OutputDesc),
                                        ctypes.POINTER(ctypes.c_char_p)], ctypes.c_int),
    "rise_hip_scene_file_get_directional_lights": ([_P, _PP], ctypes.c_int),
    "rise_hip_scene_file_release": ([_P], ctypes.c_int),
}


def _lib_scene_file():
    L = _lib.lib()
    for name, (a, r) in _SIG.items():
        fn = getattr(L, name)
        fn.argtypes = a
        fn.restype = r
    return L


def _f(x) -> float:
    """A float32 field as the shortest decimal that rounds back to it."""
    return float(str(np.float32(x)))


def _array(ptr, ctype, n):
    if n == 0:
        return []
    return ctypes.cast(ptr, ctypes.POINTER(ctype * n)).contents


class SceneFile:
    """A parsed scene file (host tables owned by the library)."""

    def __init__(self, path: str | os.PathLike):
        L = _lib_scene_file()
        h = _P(None)
        _lib.check(L.rise_hip_scene_file_parse(os.fsencode(path), ctypes.byref(h)))
        self._h = h

    def info(self) -> dict:
        out = SceneFileInfo()
        _lib.check(_lib_scene_file().rise_hip_scene_file_get_info(self._h, ctypes.byref(out)))
        return out.as_dict()

    def render_desc(self) -> RenderDesc:
        d = RenderDesc()
        _lib.check(_lib_scene_file().rise_hip_scene_file_get_render_desc(self._h, ctypes.byref(d)))
        return d

    def scene(self) -> Scene:
        """The parsed scene as a pt_scene.Scene (copies of the tables)."""
        info = self.info()
        p = [_P(None) for _ in range(6)]
        _lib.check(_lib_scene_file().rise_hip_scene_file_get_tables(self._h, *[ctypes.byref(x) for x in p]))
        objs = _array(p[0], ObjectDesc, info["n_objects"])
        om = _array(p[1], ObjectMedia, info["n_objects"])
        mats = _array(p[2], MaterialDesc, info["n_materials"])
        spds = _array(p[3], SpdDesc, info["n_spds"])
        media = _array(p[4], MediumDesc, info["n_media"])
        s = Scene()
        for d in spds:
            sig = d.n == -1
            amp = np.ctypeslib.as_array(ctypes.cast(d.amplitudes, ctypes.POINTER(ctypes.c_float)),
                                        (4 if sig else d.n,)).copy()
            s.spds.append(Spd(_f(d.lambda_begin), _f(d.lambda_end), amp, sigmoid=sig))
        for m in mats:
            s.materials.append(Material(reflectance=m.reflectance_spd, emission=m.emission_spd,
                                        emission_scale=_f(m.emission_scale), kind=m.type,
                                        specular=m.specular_spd, ior=m.ior_spd, extinction=m.extinction_spd,
                                        alpha=_f(m.alpha)))
        for k, o in enumerate(objs):
            tris = np.ctypeslib.as_array(ctypes.cast(o.vertices, ctypes.POINTER(ctypes.c_float)),
                                         (o.n_tris * 9,)).reshape(-1, 9).copy()
            s.objects.append(Object(tris, o.material, bool(o.double_sided), medium=om[k].interior_medium,
                                    casts_shadows=bool(om[k].casts_shadows)))
        for m in media:
            s.media.append(Medium(tuple(_f(x) for x in m.sigma_a), tuple(_f(x) for x in m.sigma_s), _f(m.hg_g)))
        if p[5].value:
            sd = ctypes.cast(p[5], ctypes.POINTER(SkyDesc)).contents
            s.sky = Sky(_f(sd.solar_elevation_deg), _f(sd.solar_azimuth_deg), _f(sd.turbidity), _f(sd.sky_scale),
                        create_sun=bool(sd.create_sun), sun_power=_f(sd.sun_intensity_scale))
        if info["n_directional_lights"]:
            lp = _P(None)
            _lib.check(_lib_scene_file().rise_hip_scene_file_get_directional_lights(self._h, ctypes.byref(lp)))
            for d in _array(lp, DirectionalLightDesc, info["n_directional_lights"]):
                s.lights.append(DirectionalLight(tuple(_f(x) for x in d.direction), tuple(_f(x) for x in d.color),
                                                 _f(d.power)))
        c = self.render_desc().camera
        s.camera = Camera(tuple(_f(x) for x in c.position), tuple(_f(x) for x in c.look_at), tuple(_f(x) for x in c.up),
                          _f(c.fov_deg))
        return s

    def settings(self) -> RenderSettings:
        d = self.render_desc()
        return RenderSettings(width=d.width, height=d.height, spp=d.spp, lambda_begin=_f(d.lambda_begin),
                              lambda_end=_f(d.lambda_end), max_diffuse_bounce=d.max_diffuse_bounce,
                              max_glossy_bounce=d.max_glossy_bounce,
                              max_transmission_bounce=d.max_transmission_bounce, rr_min_depth=d.rr_min_depth,
                              rr_threshold=_f(d.rr_threshold), seed=int(d.seed),
                              sky_background=bool(d.sky_background), sampler=int(d.sampler))

    def outputs(self) -> list:
        """[(OutputDesc, file name)] of the file's file_rasterizeroutput chunks."""
        out = []
        names = {0: "TGA", 1: "PPM", 2: "PNG", 3: "HDR", 6: "EXR"}
        for i in range(self.info()["n_outputs"]):
            d, pat = OutputDesc(), ctypes.c_char_p()
            _lib.check(_lib_scene_file().rise_hip_scene_file_get_output(self._h, i, ctypes.byref(d),
                                                                         ctypes.byref(pat)))
            out.append((d, f"{pat.value.decode()}.{extension(names[d.format])}"))
        return out

    def rasterizer(self) -> PathTracingSpectralRasterizer:
        """The scene made resident by rise_hip_scene_file_upload."""
        out = _P(None)
        _lib.check(_lib_scene_file().rise_hip_scene_file_upload(self._h, ctypes.byref(out)))
        return PathTracingSpectralRasterizer.from_handle(out, self.scene())

    def release(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib_scene_file().rise_hip_scene_file_release(self._h)
            self._h = _P(None)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def load_scene(path: str | os.PathLike):
    """(Scene, RenderSettings) of a scene file."""
    f = SceneFile(path)
    try:
        return f.scene(), f.settings()
    finally:
        f.release()
