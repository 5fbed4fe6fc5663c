"""Python mirror of RISE's spectral path-tracing rasterizer over the C ABI.

``PathTracingSpectralRasterizer`` follows the reference's
``RISE_API_CreatePathTracingSpectralRasterizer`` + ``RasterizeScene``
(RISE_API.h:3112-3141, Interfaces/IRasterizer.h:70): it is created for a
scene, and ``rasterize_scene`` renders the configured samples of every pixel
into an XYZ film on the GPU (mean XYZ per pixel; the reference resolves the
same XYZ sums to RGB afterwards).  Errors raise RiseHipError.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .pt_scene import RenderSettings, Scene


class SpdDesc(ctypes.Structure):
    _fields_ = [
        ("lambda_begin", ctypes.c_float),
        ("lambda_end", ctypes.c_float),
        ("n", ctypes.c_int32),
        ("amplitudes", ctypes.c_void_p),
    ]


class MaterialDesc(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32),
        ("reflectance_spd", ctypes.c_int32),
        ("emission_spd", ctypes.c_int32),
        ("emission_scale", ctypes.c_float),
        ("specular_spd", ctypes.c_int32),
        ("ior_spd", ctypes.c_int32),
        ("extinction_spd", ctypes.c_int32),
        ("alpha", ctypes.c_float),
    ]


class ObjectDesc(ctypes.Structure):
    _fields_ = [
        ("vertices", ctypes.c_void_p),
        ("n_tris", ctypes.c_int64),
        ("material", ctypes.c_int32),
        ("double_sided", ctypes.c_int32),
    ]


class MediumDesc(ctypes.Structure):
    _fields_ = [
        ("sigma_a", ctypes.c_float * 3),
        ("sigma_s", ctypes.c_float * 3),
        ("hg_g", ctypes.c_float),
        ("reserved", ctypes.c_int32),
    ]


class ObjectMedia(ctypes.Structure):
    _fields_ = [("interior_medium", ctypes.c_int32), ("casts_shadows", ctypes.c_int32)]


class SkyDesc(ctypes.Structure):
    _fields_ = [
        ("solar_elevation_deg", ctypes.c_float),
        ("solar_azimuth_deg", ctypes.c_float),
        ("turbidity", ctypes.c_float),
        ("sky_scale", ctypes.c_float),
        ("create_sun", ctypes.c_int32),
        ("sun_intensity_scale", ctypes.c_float),
    ]


class DirectionalLightDesc(ctypes.Structure):
    _fields_ = [("direction", ctypes.c_float * 3), ("color", ctypes.c_float * 3), ("power", ctypes.c_float)]


class CameraDesc(ctypes.Structure):
    _fields_ = [
        ("position", ctypes.c_float * 3),
        ("look_at", ctypes.c_float * 3),
        ("up", ctypes.c_float * 3),
        ("fov_deg", ctypes.c_float),
    ]


class RenderDesc(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("spp", ctypes.c_int32),
        ("spp_offset", ctypes.c_int32),
        ("spp_total", ctypes.c_int32),
        ("lambda_begin", ctypes.c_float),
        ("lambda_end", ctypes.c_float),
        ("max_diffuse_bounce", ctypes.c_int32),
        ("max_glossy_bounce", ctypes.c_int32),
        ("max_transmission_bounce", ctypes.c_int32),
        ("rr_min_depth", ctypes.c_int32),
        ("rr_threshold", ctypes.c_float),
        ("seed", ctypes.c_uint64),
        ("camera", CameraDesc),
        ("sky_background", ctypes.c_int32),
        ("sampler", ctypes.c_int32),
    ]


class RenderStats(ctypes.Structure):
    _fields_ = [
        ("extension_rays", ctypes.c_int64),
        ("shadow_rays", ctypes.c_int64),
        ("paths", ctypes.c_int64),
        ("bounces", ctypes.c_int64),
        ("ior_overflows", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class RenderRequest(ctypes.Structure):
    _fields_ = [
        ("scene", ctypes.c_void_p),
        ("desc", RenderDesc),
        ("film_xyz", ctypes.c_void_p),
        ("film_xyz_sq", ctypes.c_void_p),
    ]


class BatchTuning(ctypes.Structure):
    _fields_ = [
        ("max_paths", ctypes.c_int32),
        ("trace_block", ctypes.c_int32),
        ("stack", ctypes.c_int32),
        ("persistent", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_int64),
        ("max_requests", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("tuning", BatchTuning),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in ("submitted", "completed", "admitted", "waiting", "passes", "paths",
                                              "extension_rays", "shadow_rays", "hbm_in_use", "cancelled",
                                              "timed_out", "failed", "alloc_retries", "ior_overflows")]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = ctypes.c_void_p
_SIG = {
    "rise_hip_pt_scene_create": ([_P, ctypes.c_size_t, _P, ctypes.c_size_t, _P, ctypes.c_size_t, ctypes.POINTER(_P)],
                                 ctypes.c_int),
    "rise_hip_pt_scene_create_with_media": ([_P, ctypes.c_size_t, _P, ctypes.c_size_t, _P, ctypes.c_size_t, _P,
                                             ctypes.c_size_t, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "rise_hip_pt_scene_release": ([_P], ctypes.c_int),
    "rise_hip_pt_scene_set_sky": ([_P, ctypes.POINTER(SkyDesc)], ctypes.c_int),
    "rise_hip_pt_scene_set_directional_lights": ([_P, ctypes.POINTER(DirectionalLightDesc), ctypes.c_size_t],
                                                 ctypes.c_int),
    "rise_hip_render": ([_P, ctypes.POINTER(RenderDesc), _P, _P, ctypes.POINTER(RenderStats), _P], ctypes.c_int),
    "rise_hip_batch_create": ([_P, ctypes.c_size_t, ctypes.POINTER(_P)], ctypes.c_int),
    "rise_hip_batch_create_tuned": ([_P, ctypes.c_size_t, ctypes.POINTER(BatchTuning), ctypes.POINTER(_P)],
                                    ctypes.c_int),
    "rise_hip_batch_get_tuning": ([_P, ctypes.POINTER(BatchTuning)], ctypes.c_int),
    "rise_hip_batch_launch": ([_P, _P], ctypes.c_int),
    "rise_hip_batch_get_stats": ([_P, _P, ctypes.POINTER(RenderStats)], ctypes.c_int),
    "rise_hip_batch_release": ([_P], ctypes.c_int),
    "rise_hip_server_create": ([ctypes.POINTER(ServerConfig), _P, ctypes.POINTER(_P)], ctypes.c_int),
    "rise_hip_server_submit": ([_P, ctypes.POINTER(RenderRequest), ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
    "rise_hip_server_poll": ([_P, _P, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "rise_hip_server_drain": ([_P, _P, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "rise_hip_server_get_stats": ([_P, ctypes.POINTER(ServerStats)], ctypes.c_int),
    "rise_hip_server_poll_results": ([_P, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "rise_hip_server_drain_results": ([_P, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "rise_hip_server_cancel": ([_P, ctypes.c_uint64], ctypes.c_int),
    "rise_hip_server_set_timeout": ([_P, ctypes.c_uint64, ctypes.c_int64], ctypes.c_int),
    "rise_hip_server_fail_next_allocations": ([_P, ctypes.c_int32], ctypes.c_int),
    "rise_hip_server_release": ([_P], ctypes.c_int),
}


REQUEST_DONE, REQUEST_CANCELLED, REQUEST_TIMED_OUT, REQUEST_FAILED = 0, 1, 2, 3


class ServerResult(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("latency_ns", ctypes.c_int64), ("status", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


def _lib_render():
    L = _lib.lib()
    for name, (a, r) in _SIG.items():
        fn = getattr(L, name)
        fn.argtypes = a
        fn.restype = r
    return L


def render_desc(settings: RenderSettings, scene: Scene, spp: int | None = None, spp_offset: int = 0,
                spp_total: int | None = None) -> RenderDesc:
    d = RenderDesc()
    d.width = settings.width
    d.height = settings.height
    d.spp = settings.spp if spp is None else spp
    d.spp_offset = spp_offset
    d.spp_total = (settings.spp if spp_total is None else spp_total)
    d.lambda_begin = settings.lambda_begin
    d.lambda_end = settings.lambda_end
    d.max_diffuse_bounce = settings.max_diffuse_bounce
    d.max_glossy_bounce = settings.max_glossy_bounce
    d.max_transmission_bounce = settings.max_transmission_bounce
    d.rr_min_depth = settings.rr_min_depth
    d.rr_threshold = settings.rr_threshold
    d.seed = settings.seed
    c = scene.camera
    d.camera.position[:] = c.position
    d.camera.look_at[:] = c.look_at
    d.camera.up[:] = c.up
    d.camera.fov_deg = c.fov_deg
    d.sky_background = 1 if settings.sky_background else 0
    d.sampler = int(settings.sampler)
    return d


class PathTracingSpectralRasterizer:
    """A scene resident on the GPU plus a render entry point."""

    def __init__(self, scene: Scene):
        L = _lib_render()
        self.scene = scene
        self._keep = []
        spds = (SpdDesc * max(1, len(scene.spds)))()
        for k, s in enumerate(scene.spds):
            a = np.ascontiguousarray(s.amp, np.float32)
            self._keep.append(a)
            spds[k].lambda_begin = s.begin
            spds[k].lambda_end = s.end
            spds[k].n = -1 if s.sigmoid else a.size
            spds[k].amplitudes = a.ctypes.data
        mats = (MaterialDesc * max(1, len(scene.materials)))()
        for k, m in enumerate(scene.materials):
            mats[k].type = m.kind
            mats[k].reflectance_spd = m.reflectance
            mats[k].emission_spd = m.emission
            mats[k].emission_scale = m.emission_scale
            mats[k].specular_spd = m.specular
            mats[k].ior_spd = m.ior
            mats[k].extinction_spd = m.extinction
            mats[k].alpha = m.alpha
        objs = (ObjectDesc * max(1, len(scene.objects)))()
        for k, o in enumerate(scene.objects):
            t = np.ascontiguousarray(o.tris, np.float32).reshape(-1, 9)
            self._keep.append(t)
            objs[k].vertices = t.ctypes.data
            objs[k].n_tris = t.shape[0]
            objs[k].material = o.material
            objs[k].double_sided = 1 if o.double_sided else 0
        out = _P(None)
        if scene.media or not all(o.casts_shadows for o in scene.objects):
            med = (MediumDesc * max(1, len(scene.media)))()
            for k, m in enumerate(scene.media):
                med[k].sigma_a[:] = m.sigma_a
                med[k].sigma_s[:] = m.sigma_s
                med[k].hg_g = m.g
            om = (ObjectMedia * max(1, len(scene.objects)))()
            for k, o in enumerate(scene.objects):
                om[k].interior_medium = o.medium
                om[k].casts_shadows = 1 if o.casts_shadows else 0
            _lib.check(L.rise_hip_pt_scene_create_with_media(objs, len(scene.objects), mats, len(scene.materials),
                                                             spds, len(scene.spds), med, len(scene.media), om,
                                                             ctypes.byref(out)))
        else:
            _lib.check(L.rise_hip_pt_scene_create(objs, len(scene.objects), mats, len(scene.materials), spds,
                                                  len(scene.spds), ctypes.byref(out)))
        self._h = out
        if scene.lights:
            dl = (DirectionalLightDesc * len(scene.lights))()
            for k, l in enumerate(scene.lights):
                dl[k].direction[:] = l.direction
                dl[k].color[:] = l.color
                dl[k].power = l.power
            _lib.check(L.rise_hip_pt_scene_set_directional_lights(self._h, dl, len(scene.lights)))
        if scene.sky is not None:
            k = scene.sky
            sd = SkyDesc(k.elevation_deg, k.azimuth_deg, k.turbidity, k.scale, 1 if k.create_sun else 0, k.sun_power)
            _lib.check(L.rise_hip_pt_scene_set_sky(self._h, ctypes.byref(sd)))
        self._keep = []

    @classmethod
    def from_handle(cls, handle, scene: Scene) -> "PathTracingSpectralRasterizer":
        """Wraps a scene made resident elsewhere (rise_hip_scene_file_upload);
        `scene` supplies the camera."""
        self = cls.__new__(cls)
        self.scene = scene
        self._keep = []
        self._h = handle
        return self

    def rasterize_scene(self, settings: RenderSettings, spp: int | None = None, spp_offset: int = 0,
                        spp_total: int | None = None, want_sq: bool = False, stream=None, stats: bool = True,
                        film=None, film_sq=None):
        """Renders and returns (mean XYZ (H, W, 3) tensor, mean squared XYZ or
        None, stats dict or None).  Tensors stay on the GPU."""
        import torch

        d = render_desc(settings, self.scene, spp, spp_offset, spp_total)
        if film is None:
            film = torch.empty((settings.height, settings.width, 3), dtype=torch.float32, device="cuda")
        if want_sq and film_sq is None:
            film_sq = torch.empty_like(film)
        st = RenderStats() if stats else None
        s = _stream_handle(stream)
        _lib.check(_lib_render().rise_hip_render(self._h, ctypes.byref(d), film.data_ptr(),
                                                 film_sq.data_ptr() if want_sq else None,
                                                 ctypes.byref(st) if stats else None, s))
        return film, (film_sq if want_sq else None), (st.as_dict() if stats else None)

    def release(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib_render().rise_hip_pt_scene_release(self._h)
            self._h = _P(None)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def _stream_handle(stream) -> int:
    import torch

    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    return int(getattr(stream, "cuda_stream", stream))


class RenderBatch:
    """Several renders planned as one launch sequence (rise_hip_batch_*).

    ``items`` is a list of (PathTracingSpectralRasterizer, RenderSettings) or
    (rasterizer, settings, spp, spp_offset, spp_total).  Films are allocated
    here (``films[k]``: mean XYZ (H, W, 3); ``films_sq[k]`` when want_sq) and
    rewritten by every ``launch``; each is identical to what
    ``rasterize_scene`` gives for that item alone.  ``launch`` only enqueues
    kernels, so it may be captured in a CUDA/HIP graph.  ``tuning`` sets the
    launch shapes (rise_hip_batch_tuning fields; see rise_amd.autotune)."""

    def __init__(self, items, want_sq: bool = False, tuning: dict | None = None):
        import torch

        L = _lib_render()
        self.items = list(items)
        self.films, self.films_sq = [], []
        reqs = (RenderRequest * max(1, len(self.items)))()
        for k, it in enumerate(self.items):
            rast, settings = it[0], it[1]
            spp, spp_offset, spp_total = (list(it[2:]) + [None, 0, None][len(it) - 2:])[:3]
            film = torch.empty((settings.height, settings.width, 3), dtype=torch.float32, device="cuda")
            sq = torch.empty_like(film) if want_sq else None
            self.films.append(film)
            self.films_sq.append(sq)
            reqs[k].scene = rast._h.value
            reqs[k].desc = render_desc(settings, rast.scene, spp, spp_offset, spp_total)
            reqs[k].film_xyz = film.data_ptr()
            reqs[k].film_xyz_sq = sq.data_ptr() if want_sq else None
        out = _P(None)
        tn = BatchTuning(**(tuning or {}))
        _lib.check(L.rise_hip_batch_create_tuned(reqs, len(self.items), ctypes.byref(tn), ctypes.byref(out)))
        self._h = out

    def tuning(self) -> dict:
        """The launch shapes this batch resolved to."""
        t = BatchTuning()
        _lib.check(_lib_render().rise_hip_batch_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def launch(self, stream=None) -> None:
        _lib.check(_lib_render().rise_hip_batch_launch(self._h, _stream_handle(stream)))

    def stats(self, stream=None) -> dict:
        st = RenderStats()
        _lib.check(_lib_render().rise_hip_batch_get_stats(self._h, _stream_handle(stream), ctypes.byref(st)))
        return st.as_dict()

    def release(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib_render().rise_hip_batch_release(self._h)
            self._h = _P(None)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class RenderServer:
    """Continuous-batching render server (rise_hip_server_*): scenes stay
    resident, requests are submitted at any time and share passes with the
    requests already rendering.  ``submit`` returns a ticket and the film
    tensor the request will write; ``poll`` / ``drain`` report finished
    tickets with their submit-to-finish latency in seconds."""

    def __init__(self, hbm_budget_bytes: int = 0, max_requests: int = 0, tuning: dict | None = None, stream=None):
        import numpy as np  # noqa: F401

        self._stream = stream
        cfg = ServerConfig(hbm_budget_bytes, max_requests, 0, BatchTuning(**(tuning or {})))
        out = _P(None)
        _lib.check(_lib_render().rise_hip_server_create(ctypes.byref(cfg), _stream_handle(stream), ctypes.byref(out)))
        self._h = out
        self._keep = {}
        self._tk = (ctypes.c_uint64 * 512)()
        self._lat = (ctypes.c_int64 * 512)()
        self._res = (ServerResult * 512)()

    def submit(self, rast, settings, spp=None, spp_offset: int = 0, spp_total=None, want_sq: bool = False):
        import torch

        film = torch.empty((settings.height, settings.width, 3), dtype=torch.float32, device="cuda")
        sq = torch.empty_like(film) if want_sq else None
        q = RenderRequest()
        q.scene = rast._h.value
        q.desc = render_desc(settings, rast.scene, spp, spp_offset, spp_total)
        q.film_xyz = film.data_ptr()
        q.film_xyz_sq = sq.data_ptr() if want_sq else None
        t = ctypes.c_uint64()
        _lib.check(_lib_render().rise_hip_server_submit(self._h, ctypes.byref(q), ctypes.byref(t)))
        self._keep[t.value] = (rast, film, sq)
        return t.value, film, sq

    def _collect(self, fn) -> list:
        n = ctypes.c_size_t()
        _lib.check(fn(self._h, self._tk, self._lat, 512, ctypes.byref(n)))
        out = []
        for i in range(n.value):
            out.append((int(self._tk[i]), self._lat[i] * 1e-9))
            self._keep.pop(int(self._tk[i]), None)
        return out

    def poll(self) -> list:
        return self._collect(_lib_render().rise_hip_server_poll)

    def _collect_results(self, fn) -> list:
        n = ctypes.c_size_t()
        _lib.check(fn(self._h, self._res, 512, ctypes.byref(n)))
        out = []
        for i in range(n.value):
            r = self._res[i]
            out.append((int(r.ticket), r.latency_ns * 1e-9, int(r.status)))
            self._keep.pop(int(r.ticket), None)
        return out

    def poll_results(self) -> list:
        """Finished requests as (ticket, latency s, status: REQUEST_*)."""
        return self._collect_results(_lib_render().rise_hip_server_poll_results)

    def drain_results(self) -> list:
        done = []
        while True:
            got = self._collect_results(_lib_render().rise_hip_server_drain_results)
            done += got
            if len(got) < 512:
                return done

    def cancel(self, ticket: int) -> None:
        _lib.check(_lib_render().rise_hip_server_cancel(self._h, ticket))

    def set_timeout(self, ticket: int, seconds: float) -> None:
        _lib.check(_lib_render().rise_hip_server_set_timeout(self._h, ticket, int(seconds * 1e9)))

    def fail_next_allocations(self, n: int) -> None:
        """Test hook: the next n film allocations report out-of-memory."""
        _lib.check(_lib_render().rise_hip_server_fail_next_allocations(self._h, n))

    def drain(self) -> list:
        done = []
        while True:
            got = self._collect(_lib_render().rise_hip_server_drain)
            done += got
            if len(got) < 512:
                return done

    def stats(self) -> dict:
        st = ServerStats()
        _lib.check(_lib_render().rise_hip_server_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def release(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib_render().rise_hip_server_release(self._h)
            self._h = _P(None)
        self._keep = {}

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
