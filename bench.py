#!/usr/bin/env python3
"""Ray-cast benchmark on one MI355X.

Workload: closest hit for 6,073,600 rays (1920x1080 coherent pinhole rays +
4,000,000 incoherent rays) against a 1,000,000-triangle seeded soup
(centroids uniform in the unit cube, Gaussian edges sigma 0.003).  A step is
one closest-hit pass over the whole ray batch, rays and BVH already resident
in HBM.  ``value`` is Mrays/s over the timed steps.

The same JSON line carries:
* roofline: algorithmic bytes per ray B_ray = 32 + 80 * nodes/ray + 48 *
  triangles/ray (24 B ray in, 8 B result out; 80 B wide node; 48 B triangle),
  counts from the counted kernel variant on the same rays, times rays per
  launch, divided by the average launch time measured with HIP events on the
  launch stream; HBM traffic per launch from rocprofv3 counters when a
  measured summary is committed (profiles/closest_hit_traffic.json).
* cpu_baseline: the reference's own BVH4 mesh path (oracle/_ref, built with
  the reference's Linux flags) on this host's cores over an evenly spaced
  sample of the same rays, with its ids checked against the GPU's.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s + render requests/s with p50/p99 latency at ≤1% rel-RMSE, 1 GPU"
HBM_PEAK_GBS = 8000.0
TRI_BYTES = 48
RAY_IO_BYTES = 32


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def copy_calibration(steps: int) -> None:
    """A plain device copy of 1 GiB, for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE."""
    import torch

    n = 1 << 28  # float32 elements = 1 GiB
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(steps):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copied {steps} x {n * 4} bytes")


def load_kernel_counters(leg: str):
    """Per-kernel HBM traffic and VALU busy of a render leg from the committed
    rocprofv3 summary (profiles/path_tracing_kernel_counters.json), or None."""
    p = ROOT / "profiles" / "path_tracing_kernel_counters.json"
    if not p.exists():
        return None
    d = json.loads(p.read_text())
    k = d.get("legs", {}).get(leg)
    return {"from": "profiles/path_tracing_kernel_counters_batch_terrain.md", "kernels": k} if k else None


def camera_rays(cam, width: int, height: int):
    """Pinhole rays through the pixel centres of `cam` (pt_scene.Camera), SoA (6, n)."""
    import numpy as np

    pos, at, up = (np.asarray(v, np.float64) for v in (cam.position, cam.look_at, cam.up))
    w = at - pos
    w /= np.linalg.norm(w)
    u = np.cross(w, up)
    u /= np.linalg.norm(u)
    v = np.cross(u, w)
    half = np.tan(np.radians(cam.fov_deg) / 2.0)
    xs = ((np.arange(width) + 0.5) / width * 2.0 - 1.0) * half * width / height
    ys = (1.0 - (np.arange(height) + 0.5) / height * 2.0) * half
    X, Y = np.meshgrid(xs, ys)
    d = w[None, None] + X[..., None] * u + Y[..., None] * v
    d = d.reshape(-1, 3)
    o = np.broadcast_to(pos, d.shape)
    return np.ascontiguousarray(np.concatenate([o.T, d.T], 0), np.float32)


def terrain_cpu_baseline(scene, rs) -> dict | None:
    """The reference's own BVH4 mesh intersection (oracle/_ref, its Linux-flags
    build) on the terrain: closest hits of the 1920x1080 camera rays, then
    shadow rays from those hits toward the sun, on the host threads.  Path
    tracing adds sampling, shading and film work to these queries, so this
    rate bounds the reference's path-tracing Mrays/s from above."""
    import numpy as np

    import oracle

    if not oracle.ref_available("fast"):
        return None
    tris, _ = scene.flat()
    t0 = time.perf_counter()
    ref = oracle.RefMesh(tris, variant="fast")
    build_s = time.perf_counter() - t0
    th = cpu_threads()
    prim = camera_rays(scene.camera, rs.width, rs.height)
    t0 = time.perf_counter()
    pid, _, t64 = ref.closest(prim, threads=th)
    t_closest = time.perf_counter() - t0
    hit = pid >= 0
    dn = prim[3:][:, hit] / np.linalg.norm(prim[3:][:, hit], axis=0)
    p = prim[:3][:, hit] + dn * (t64[hit] * (1 - 1e-6))
    el, az = np.radians(scene.sky.elevation_deg), np.radians(scene.sky.azimuth_deg)
    sun = np.array([np.cos(el) * np.sin(az), np.sin(el), np.cos(el) * np.cos(az)])
    shadow = np.ascontiguousarray(np.concatenate([p, np.repeat(sun[:, None], p.shape[1], 1)], 0), np.float32)
    t0 = time.perf_counter()
    ref.any_hit(shadow, threads=th)
    t_any = time.perf_counter() - t0
    ref.close()
    n = prim.shape[1] + shadow.shape[1]
    return {"value": round(n / (t_closest + t_any) / 1e6, 3), "unit": "Mrays/s", "cores": th, "kind": "reference",
            "sample": f"{prim.shape[1]} camera rays (closest hit) + {shadow.shape[1]} sun shadow rays (any hit) "
                      f"on the terrain scene, RISE BVH4 mesh path built with its Linux flags; intersection only, "
                      f"an upper bound on its path-tracing rate; {t_closest + t_any:.1f} s of work; BVH build "
                      f"{build_s:.1f} s not counted"}


def load_traffic():
    p = ROOT / "profiles" / "closest_hit_traffic.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        return float(d["hbm_bytes_per_launch"])
    except Exception:
        return None


REL_RMSE_TARGET = 0.01
CORNELL_REFERENCE = ROOT / "tests" / "golden" / "cornell_oracle_1024spp.npz"


def _trials(pt, rs, spp: int, trials: int, seed0: int):
    """`trials` independent renders of `rs` at `spp` (render seeds seed0, seed0 + 1, ...):
    (films (K, H, W, 3) float64 on the GPU, ms per render, rays per render)."""
    import dataclasses

    import torch

    films, rays = [], 0
    for k in range(trials):
        r = dataclasses.replace(rs, spp=spp, seed=seed0 + k)
        film, _, st = pt.rasterize_scene(r, stats=(k == 0))
        if k == 0:
            rays = st["extension_rays"] + st["shadow_rays"]
        films.append(film.double())
    # Timed: the last request again (its plan is reused, as in a render loop).
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    pt.rasterize_scene(r, stats=False)
    b.record()
    torch.cuda.synchronize()
    return torch.stack(films), float(a.elapsed_time(b)), int(rays)


def _rel_noise(films) -> float:
    """Relative RMSE of one render from K independent ones: sqrt(mean over pixels
    and XYZ channels of the inter-render variance) / mean of the K-render mean.
    Unbiased renders (checked against the oracle) make this the distance to
    the converged image (RISE's K-trial measure, docs/UNIFIED_INTEGRATOR_BASELINES.md)."""
    var = films.var(dim=0, unbiased=True)
    return float(var.mean().sqrt() / films.mean(dim=0).mean())


def image_quality(pt, rs, spp: int, trials: int, seed0: int, ref=None):
    """Relative RMSE at `spp` from `trials` renders, the samples per pixel
    that reach REL_RMSE_TARGET under the 1/N law, and a measurement at that
    count (K renders, their own relative RMSE, time and Mrays/s).  With an
    oracle reference (dict: xyz, var_y, spp): each render's relative RMSE
    against it and the per-pixel z-test of the K-render mean (Y) against it
    with pooled variance (the inter-render variance and the reference's
    per-sample variance / its spp)."""
    import math

    import numpy as np
    import torch

    films, ms, rays = _trials(pt, rs, spp, trials, seed0)
    rel = _rel_noise(films)
    out = {"spp": spp, "trials": trials, "rel_rmse": round(rel, 5), "ms_per_render": round(ms, 3)}
    if ref is not None:
        rx = torch.from_numpy(ref["xyz"].astype(np.float64)).to(films.device)
        per = ((films - rx) ** 2).mean(dim=(1, 2, 3)).sqrt() / rx.mean()
        out["rel_rmse_vs_oracle"] = round(float(per.mean()), 5)
        out["oracle_spp"] = int(ref["spp"])
        gm = films.mean(dim=0)[..., 1]
        gv = films.var(dim=0, unbiased=True)[..., 1] / trials
        ov = torch.from_numpy(ref["var_y"].astype(np.float64)).to(films.device) / float(ref["spp"])
        z = (gm - rx[..., 1]) / torch.sqrt(gv + ov + 1e-30)
        out["z_test_pass_rate_Y"] = round(float((z.abs() < 3.29).double().mean()), 5)
        out["mean_Y_rel_diff_vs_oracle"] = float(gm.mean() / rx[..., 1].mean() - 1.0)
    need = int(math.ceil(spp * (rel / REL_RMSE_TARGET) ** 2 / 16.0) * 16)
    out["spp_for_1pct"] = need
    f1, ms1, rays1 = _trials(pt, rs, need, max(4, trials // 2), seed0 + 1000)
    out["at_spp_for_1pct"] = {
        "rel_rmse": round(_rel_noise(f1), 5),
        "trials": int(f1.shape[0]),
        "ms_per_render": round(ms1, 2),
        "mrays_per_s": round(rays1 / (ms1 / 1e3) / 1e6, 1),
        "renders_per_s": round(1e3 / ms1, 3),
    }
    del films, f1
    torch.cuda.empty_cache()
    return out


def render_leg(args, cpu_enabled: bool):
    """Cornell-box spectral path tracing: 256x256, 16 spp, 4 hero wavelengths,
    diffuse depth 5, Russian roulette from depth 3.  Rays = extension + shadow
    queries actually traced."""
    import numpy as np
    import torch

    from rise_amd import pt_scene
    from rise_amd.render import PathTracingSpectralRasterizer

    scene, rs = pt_scene.cornell_box()
    pt = PathTracingSpectralRasterizer(scene)
    film, _, stats = pt.rasterize_scene(rs)  # warmup + ray counts
    torch.cuda.synchronize()
    rays = stats["extension_rays"] + stats["shadow_rays"]
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    lat = []
    for _ in range(args.render_steps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        pt.rasterize_scene(rs, film=film, stats=False)
        b.record()
        torch.cuda.synchronize()
        lat.append(a.elapsed_time(b))
    start.record()
    for _ in range(args.render_steps):
        pt.rasterize_scene(rs, film=film, stats=False)
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / args.render_steps
    out = {
        "workload": "Cornell box (36 tris), 256x256, 16 spp, HWSS 4 wavelengths 380-720 nm, diffuse depth 5, RR from 3",
        "ms_per_render": round(ms, 3),
        "mrays_per_s": round(rays / (ms / 1e3) / 1e6, 2),
        "paths_per_s": round(stats["paths"] / (ms / 1e3), 1),
        "renders_per_s": round(1e3 / ms, 2),
        "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
        "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
        "rays_per_render": {"extension": stats["extension_rays"], "shadow": stats["shadow_rays"]},
        "cpu_baseline": None,
        "parity": None,
    }
    log(f"render: {ms:.2f} ms, {out['mrays_per_s']} Mrays/s, {out['paths_per_s']:.3g} paths/s")
    if not args.no_quality:
        ref = None
        if CORNELL_REFERENCE.exists():
            z = np.load(CORNELL_REFERENCE)
            ref = {"xyz": z["xyz"], "var_y": z["var_y"], "spp": int(z["spp"])}
        out["quality"] = image_quality(pt, rs, rs.spp, args.quality_trials, 5000, ref)
        out["rel_rmse"] = out["quality"]["rel_rmse"]
        log(f"render quality: {out['quality']}")
    if cpu_enabled:
        import oracle

        th = cpu_threads()
        t0 = time.perf_counter()
        sm, _, cnt = oracle.render(scene, rs, threads=th)
        sec = time.perf_counter() - t0
        cpu_rays = int(cnt[0] + cnt[1])
        out["cpu_baseline"] = {
            "value": round(cpu_rays / sec / 1e6, 4),
            "unit": "Mrays/s",
            "cores": th,
            "kind": "port",
            "sample": f"the whole render (65,536 pixels x 16 spp) with the FP64 C oracle, brute-force "
            f"intersection over the 36 triangles; {sec:.1f} s",
        }
        g = film.cpu().numpy().astype(np.float64)
        o = sm / rs.spp
        floor = 0.01 * np.abs(o[..., 1]).mean()
        rel = np.abs(g - o) / (np.abs(o) + floor)
        out["parity"] = {
            "against": "FP64 oracle on the same random streams",
            "pixels_within_1e-3": round(float(np.mean(np.all(rel <= 1e-3, axis=-1))), 5),
            "image_mean_Y_rel_diff": float(abs(g[..., 1].mean() / o[..., 1].mean() - 1.0)),
            "extension_rays_rel_diff": float(abs(stats["extension_rays"] / cnt[0] - 1.0)),
        }
        log(f"render CPU port: {out['cpu_baseline']['value']} Mrays/s on {th} threads; parity {out['parity']}")
    pt.release()
    return out


def terrain_leg(args):
    """Terrain-and-spheres spectral path tracing: a 2048x1024 value-noise
    heightfield (4,188,162 triangles) with 64 GGX-conductor and glass
    icospheres (327,680 triangles) under the spectral sky, 1920x1080, 64 spp,
    depth 8.  Rays = extension + shadow queries actually traced."""
    import numpy as np
    import torch

    from rise_amd import pt_scene
    from rise_amd.render import PathTracingSpectralRasterizer

    t0 = time.perf_counter()
    scene, rs = pt_scene.terrain_spheres()
    pt = PathTracingSpectralRasterizer(scene)
    setup = time.perf_counter() - t0
    film, _, stats = pt.rasterize_scene(rs)  # warm-up, plans the render, counts rays
    torch.cuda.synchronize()
    rays = stats["extension_rays"] + stats["shadow_rays"]
    lat = []
    for _ in range(max(1, args.terrain_steps)):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        pt.rasterize_scene(rs, film=film, stats=False)
        b.record()
        torch.cuda.synchronize()
        lat.append(a.elapsed_time(b))
    ms = float(np.mean(lat))
    y = film[..., 1].float().mean().item()
    out = {
        "workload": "terrain 2048x1024 value noise (4,188,162 tris) + 64 icospheres (GGX alpha U[0.05,0.5] / glass "
        "1.5), spectral sky, 1920x1080, 64 spp, depth 8",
        "n_tris": int(scene.n_tris),
        "ms_per_render": round(ms, 1),
        "mrays_per_s": round(rays / (ms / 1e3) / 1e6, 1),
        "paths_per_s": round(stats["paths"] / (ms / 1e3), 1),
        "rays_per_render": {"extension": stats["extension_rays"], "shadow": stats["shadow_rays"]},
        "mean_Y": round(y, 6),
        "setup_s": round(setup, 1),
        "kernel_counters": load_kernel_counters("terrain"),
    }
    log(f"terrain: {ms:.0f} ms, {out['mrays_per_s']} Mrays/s, {out['paths_per_s']:.3g} paths/s "
        f"({scene.n_tris} tris, setup {setup:.1f} s)")
    if not args.no_cpu_baseline:
        cb = terrain_cpu_baseline(scene, rs)
        if cb:
            out["cpu_baseline"] = cb
            log(f"terrain CPU reference (intersection only): {cb['value']} Mrays/s on {cb['cores']} threads")
    if not args.no_quality:
        import dataclasses

        small = dataclasses.replace(rs, width=480, height=270)
        q = image_quality(pt, small, rs.spp, max(4, args.quality_trials // 2), 9000)
        q["image"] = "480x270 (a quarter of 1920x1080 per axis), same camera"
        out["quality"] = q
        out["rel_rmse"] = q["rel_rmse"]
        log(f"terrain quality: {q}")
    pt.release()
    return out


def widened_leg(args, tris):
    """The stages either side of the hot path, each timed on its own:
    * .risemesh: writing the 1M-triangle soup with its BVH2 cache, reading it,
      and the 8-wide build from the file's cache vs from scratch (host); the
      scene reloaded from its saved 8-wide tree (no build);
    * RGB uplift: generating the 64^3 Jakob-Hanika table on the library's host pool (OMP_NUM_THREADS, at most 16 threads);
    * scene file: parsing the Cornell-box scene file and uploading its scene;
    * output: the per-pixel output kernel on a 1920x1080 film (HIP events on
      its stream; 12 B read + 4 B written per pixel against the HBM peak) and
      the whole PNG write (GPU resolve, copy, host framing, file)."""
    import tempfile
    from pathlib import Path

    import numpy as np
    import torch

    from rise_amd.output import output_desc, resolve, write_image
    from rise_amd.raycast import ObjectManager
    from rise_amd.rgb_uplift import RgbLut
    from rise_amd.risemesh import RiseMesh, write_risemesh
    from rise_amd.scene_file import SceneFile

    out = {}
    with tempfile.TemporaryDirectory() as td:
        # warm the upload path (first device allocations) on a small mesh
        write_risemesh(Path(td) / "small.risemesh", tris[:1000], with_bvh_cache=True)
        RiseMesh(Path(td) / "small.risemesh").object_manager().release()
        p = Path(td) / "soup.risemesh"
        t0 = time.perf_counter()
        write_risemesh(p, tris, with_bvh_cache=True)
        t1 = time.perf_counter()
        m = RiseMesh(p)
        t2 = time.perf_counter()
        om = m.object_manager()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        st = om.stats()
        om.release()
        write_risemesh(Path(td) / "plain.risemesh", tris, with_bvh_cache=False)
        m2 = RiseMesh(Path(td) / "plain.risemesh")
        t4 = time.perf_counter()
        om2 = m2.object_manager()
        torch.cuda.synchronize()
        t5 = time.perf_counter()
        st2 = om2.stats()
        om2.release()
        # the built 8-wide tree saved and read back: no build at all
        om3 = RiseMesh(p).object_manager()
        om3.save_cache(Path(td) / "soup.rhbvh")
        om3.release()
        t6 = time.perf_counter()
        om4 = ObjectManager.load_cache(Path(td) / "soup.rhbvh")
        torch.cuda.synchronize()
        t7 = time.perf_counter()
        st4 = om4.stats()
        om4.release()
        out["scene_cache"] = {"load_ms": round(1e3 * (t7 - t6), 1), "bvh_source": st4["bvh_source"]}
        out["risemesh"] = {"n_tris": int(tris.shape[0]), "write_with_cache_ms": round(1e3 * (t1 - t0), 1),
                           "read_ms": round(1e3 * (t2 - t1), 1),
                           "scene_from_cache_ms": round(1e3 * (t3 - t2), 1), "bvh_source_cached": st["bvh_source"],
                           "scene_from_scratch_ms": round(1e3 * (t5 - t4), 1),
                           "bvh_source_plain": st2["bvh_source"]}
        log(f"risemesh: write {out['risemesh']['write_with_cache_ms']} ms, read {out['risemesh']['read_ms']} ms, "
            f"scene from cache {out['risemesh']['scene_from_cache_ms']} ms vs from scratch "
            f"{out['risemesh']['scene_from_scratch_ms']} ms; scene from the saved 8-wide tree "
            f"{out['scene_cache']['load_ms']} ms")
    lut = RgbLut.generate(64)
    out["rgb_lut"] = {"resolution": 64, "generate_ms": round(lut.stats["build_ms"], 1),
                      "converged": lut.stats["converged"], "failed": lut.stats["failed"]}
    lut.release()
    log(f"rgb table: 64^3 generated in {out['rgb_lut']['generate_ms']} ms")
    t0 = time.perf_counter()
    sf = SceneFile(Path(__file__).resolve().parent / "tests" / "golden" / "scenes" / "cornell.RISEscene")
    t1 = time.perf_counter()
    pt = sf.rasterizer()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    pt.release()
    out["scene_file"] = {"parse_ms": round(1e3 * (t1 - t0), 2), "upload_ms": round(1e3 * (t2 - t1), 2)}
    g = torch.Generator(device="cuda").manual_seed(3)
    film = torch.rand((1080, 1920, 3), device="cuda", generator=g) * 2.0
    px = torch.empty((1080, 1920, 4), dtype=torch.uint8, device="cuda")
    d = output_desc("PNG", 8, "sRGB", "aces")
    s = torch.cuda.current_stream()
    for _ in range(3):
        resolve(film, px, d, s)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    a.record(s)
    for _ in range(reps):
        resolve(film, px, d, s)
    b.record(s)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    bytes_px = 16
    gbs = 1920 * 1080 * bytes_px / (ms / 1e3) / 1e9
    with tempfile.TemporaryDirectory() as td:
        t0 = time.perf_counter()
        write_image(film, Path(td) / "f.png", d)
        png_ms = 1e3 * (time.perf_counter() - t0)
    out["output"] = {"film": "1920x1080", "kernel_ms": round(ms, 4), "bytes_per_pixel": bytes_px,
                     "achieved_GBs": round(gbs, 1), "hbm_frac": round(gbs / 8000.0, 3),
                     "png_write_ms": round(png_ms, 1)}
    log(f"output: kernel {ms * 1e3:.1f} us per 1080p film ({gbs:.0f} GB/s, {gbs / 8000:.1%} of HBM peak), "
        f"PNG file {png_ms:.1f} ms")
    out["objects"] = objects_leg()
    return out


def objects_leg():
    """Objects with transforms: 16 objects on 4 meshes of 250,000 triangles
    (soups in [-1, 1]^3 with edges N(0, 0.005^2); 4 identity objects, 12
    rotated, scaled and translated ones: 4M world triangles), 2,000,000 rays aimed at them (three quarters) or random,
    closest hit with records off, then with full intersection records, then
    shadow queries on the same rays within half of each closest hit's range.  Timed with HIP events on the launch
    stream, against the same world triangles baked into one world-space
    scene (float vertices, so the hits agree only to rounding)."""
    import numpy as np
    import torch

    from rise_amd import raycast, scenes

    meshes, objs = scenes.instanced_scene(n_meshes=4, n_objects=16, tris_per_mesh=250000, seed=61, n_identity=4,
                                          sigma=0.0025)
    rays = scenes.object_rays(meshes, objs, 2_000_000, seed=62)
    log("objects: scene and rays made; building the object trees and the baked world tree")
    om = raycast.ObjectManager()
    for m in meshes:
        om.add_mesh(m)
    for ob in objs:
        om.add_object(int(ob["mesh"]), ob["to_world"], ob["to_object"])
    om.prepare_for_rendering()
    baked = []
    for ob in objs:
        m = ob["to_world"].reshape(4, 4)
        v = np.asarray(meshes[int(ob["mesh"])], np.float64).reshape(-1, 3)
        baked.append((v @ m[:3, :3] + m[3, :3]).astype(np.float32).reshape(-1, 9))
    flat = raycast.ObjectManager()
    flat.add_mesh(np.concatenate(baked))
    flat.prepare_for_rendering()
    r = torch.from_numpy(rays).to("cuda")
    n = rays.shape[1]
    s = torch.cuda.current_stream()
    res = {}
    for name, mgr in (("objects", om), ("baked_world", flat)):
        pid = torch.empty(n, dtype=torch.int32, device="cuda")
        t = torch.empty(n, dtype=torch.float32, device="cuda")
        for _ in range(2):
            mgr.intersect_ray(r, pid, t, stream=s)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        a.record(s)
        for _ in range(reps):
            mgr.intersect_ray(r, pid, t, stream=s)
        b.record(s)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        res[name] = {"ms": round(ms, 3), "mrays_s": round(n / ms / 1e3, 1),
                     "hit_fraction": round(float((pid >= 0).float().mean().item()), 4)}
        # Closest hit with full intersection records (rise_hip_intersect_closest_ex).
        hits = mgr.intersect_ray_ex(r, stream=s)
        a.record(s)
        for _ in range(reps):
            mgr.intersect_ray_ex(r, hits=hits, stream=s)
        b.record(s)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        res[name]["records"] = {"ms": round(ms, 3), "mrays_s": round(n / ms / 1e3, 1)}
        # Shadow queries: any hit within half the closest hit's world range
        # (misses: within 3e38).
        how_far = torch.where(pid >= 0, t * 0.5, torch.full_like(t, 3e38))
        occ = torch.empty(n, dtype=torch.uint8, device="cuda")
        mgr.intersect_shadow_ray(r, how_far, occ, stream=s)
        a.record(s)
        for _ in range(reps):
            mgr.intersect_shadow_ray(r, how_far, occ, stream=s)
        b.record(s)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        res[name]["shadow"] = {"ms": round(ms, 3), "mrays_s": round(n / ms / 1e3, 1),
                               "occluded_fraction": round(float(occ.float().mean().item()), 4)}
    om.release()
    flat.release()
    log(f"objects: 16 objects (4M world tris) {res['objects']['mrays_s']} Mrays/s vs the same triangles baked into "
        f"one world tree {res['baked_world']['mrays_s']} Mrays/s; with records {res['objects']['records']['mrays_s']} vs "
        f"{res['baked_world']['records']['mrays_s']}; shadow queries {res['objects']['shadow']['mrays_s']} "
        f"vs {res['baked_world']['shadow']['mrays_s']} Mrays/s")
    return {"workload": "16 objects on 4 meshes x 250,000 tris (4 identity, 12 transformed), 2,000,000 rays",
            **res}


SIZE_CLASSES = (("small", 1 << 18), ("medium", 1 << 22), ("large", 1 << 62))


def _size_class(r) -> str:
    """small: <= 2^18 paths (e.g. 128^2 x 16 spp), medium: <= 2^22, large: above."""
    p = r.width * r.height * r.spp
    return next(name for name, lim in SIZE_CLASSES if p <= lim)


def _serve(sv, rasts, timed):
    """Submits (arrival s, scene, settings) requests on time and polls until
    all are done.  Returns (makespan s, {ticket: latency s}, {ticket: index})."""
    lat, idx = {}, {}
    pending = list(enumerate(timed))
    t_start = time.perf_counter()
    while len(lat) < len(timed):
        now = time.perf_counter() - t_start
        while pending and pending[0][1][0] <= now:
            i, (_, k, r) = pending.pop(0)
            tk, _, _ = sv.submit(rasts[k], r)
            idx[tk] = i
        for tk, l in sv.poll():
            lat[tk] = l
        if pending and not sv.stats()["admitted"]:
            time.sleep(max(0.0, min(0.001, pending[0][1][0] - (time.perf_counter() - t_start))))
    return time.perf_counter() - t_start, lat, idx


def _class_latencies(reqs, lat, idx, alone_s=None):
    import numpy as np

    out = {}
    for name, _ in SIZE_CLASSES:
        ls = [lat[t] * 1e3 for t, i in idx.items() if _size_class(reqs[i][2]) == name]
        if not ls:
            continue
        e = {"requests": len(ls), "latency_ms_p50": round(float(np.percentile(ls, 50)), 1),
             "latency_ms_p99": round(float(np.percentile(ls, 99)), 1)}
        if alone_s is not None:
            a = [alone_s[i] * 1e3 for i in range(len(reqs)) if _size_class(reqs[i][2]) == name]
            e["alone_ms_mean"] = round(float(np.mean(a)), 2)
        out[name] = e
    return out


def serving_leg(args):
    """Serving: 16 resident scenes of 50k-2M triangles (Lambertian, GGX and
    glass, half with fog), three streams:
    * burst: 64 render requests at 128^2-1024^2 and 1-64 spp arriving as a
      Poisson stream at 200 requests/s (faster than the server can render),
    * sustained: the same 64 requests arriving at 60% of the burst's
      measured throughput, so latency is service time rather than queueing,
    * previews: 256 small requests (64^2-128^2, 1-4 spp) all at once.
    The server admits requests as they arrive and batches them continuously;
    latency is submit-to-finish per request, reported per size class
    (small <= 2^18 paths < medium <= 2^22 < large).  Compared with the same
    requests rendered one at a time with rise_hip_render, back to back (each
    timed alone for the per-class means)."""
    import numpy as np
    import torch

    from rise_amd import pt_scene
    from rise_amd.render import PathTracingSpectralRasterizer, RenderServer

    t0 = time.perf_counter()
    scenes, reqs = pt_scene.serving_workload(n_requests=args.serving_requests)
    rasts = [PathTracingSpectralRasterizer(s) for s in scenes]
    setup = time.perf_counter() - t0
    paths = sum(r.width * r.height * r.spp for _, _, r in reqs)
    log(f"serving: {len(reqs)} requests, {paths / 1e6:.1f} M paths over {len(scenes)} scenes "
        f"({sum(s.n_tris for s in scenes) / 1e6:.1f} M tris, setup {setup:.1f} s)")
    sv = RenderServer()
    for _, k, r in reqs[:4]:   # warm-up: plans kernels and the allocator
        sv.submit(rasts[k], pt_scene.RenderSettings(**{**r.__dict__, "width": 32, "height": 32, "spp": 1}))
    sv.drain()
    torch.cuda.synchronize()
    st0 = sv.stats()
    makespan, lat, idx = _serve(sv, rasts, reqs)
    st = sv.stats()
    rays = (st["extension_rays"] - st0["extension_rays"]) + (st["shadow_rays"] - st0["shadow_rays"])
    lats = np.array(sorted(lat.values())) * 1e3
    # The same requests one at a time through the single-request path.
    alone_s = []
    w0 = time.perf_counter()
    for _, k, r in reqs:
        a0 = time.perf_counter()
        rasts[k].rasterize_scene(r, stats=False)
        torch.cuda.synchronize()
        alone_s.append(time.perf_counter() - a0)
    seq_s = time.perf_counter() - w0
    burst_rps = len(reqs) / makespan
    # Sustained stream at 60% of the burst throughput.
    rate = 0.6 * burst_rps
    arr = pt_scene.poisson_arrivals(len(reqs), rate)
    sus = [(float(arr[i]), k, r) for i, (_, k, r) in enumerate(reqs)]
    sus_makespan, sus_lat, sus_idx = _serve(sv, rasts, sus)
    sus_l = np.array(sorted(sus_lat.values())) * 1e3
    # Previews: many small requests at once, against one at a time.
    prev = pt_scene.preview_requests(len(scenes))
    prev_t = [(0.0, k, r) for k, r in prev]
    _serve(sv, rasts, prev_t[:16])                      # warm-up of the small shapes
    pv_makespan, pv_lat, pv_idx = _serve(sv, rasts, prev_t)
    for k, r in prev[:16]:
        rasts[k].rasterize_scene(r, stats=False)
    torch.cuda.synchronize()
    p0 = time.perf_counter()
    for k, r in prev:
        rasts[k].rasterize_scene(r, stats=False)
        torch.cuda.synchronize()
    pv_seq = time.perf_counter() - p0
    sv.release()
    out = {
        "workload": "Poisson 200 req/s, 64 requests over 16 resident scenes (50k-2M tris; Lambertian/GGX/glass; "
        "half with a ground-fog box, homogeneous sigma_s 0.5 sigma_a 0.05 HG g 0.3), "
        "res {128,256,512,1024}^2, spp {1,4,16,64}, depth 6",
        "requests": len(reqs),
        "paths": int(paths),
        "makespan_s": round(makespan, 3),
        "requests_per_s": round(burst_rps, 2),
        "mrays_per_s": round(rays / makespan / 1e6, 1),
        "paths_per_s": round(paths / makespan, 1),
        "latency_ms_p50": round(float(np.percentile(lats, 50)), 1),
        "latency_ms_p99": round(float(np.percentile(lats, 99)), 1),
        "passes": st["passes"] - st0["passes"],
        "sequential_single_request": {"s": round(seq_s, 3), "requests_per_s": round(len(reqs) / seq_s, 2)},
        "arrival_span_s": round(reqs[-1][0], 3),
        "by_size": _class_latencies(reqs, lat, idx, alone_s),
        "sustained": {"rate_per_s": round(rate, 2), "arrival_span_s": round(float(arr[-1]), 2),
                      "makespan_s": round(sus_makespan, 3),
                      "latency_ms_p50": round(float(np.percentile(sus_l, 50)), 1),
                      "latency_ms_p99": round(float(np.percentile(sus_l, 99)), 1),
                      "by_size": _class_latencies(sus, sus_lat, sus_idx)},
        "previews": {"requests": len(prev), "sizes": "64^2-128^2, 1-4 spp",
                     "served_s": round(pv_makespan, 3), "served_requests_per_s": round(len(prev) / pv_makespan, 1),
                     "one_at_a_time_s": round(pv_seq, 3),
                     "one_at_a_time_requests_per_s": round(len(prev) / pv_seq, 1),
                     "speedup": round(pv_seq / pv_makespan, 2),
                     "latency_ms_p50": round(float(np.percentile([v * 1e3 for v in pv_lat.values()], 50)), 1),
                     "latency_ms_p99": round(float(np.percentile([v * 1e3 for v in pv_lat.values()], 99)), 1)},
    }
    log(f"serving: makespan {makespan:.2f} s, {out['requests_per_s']} req/s, {out['mrays_per_s']} Mrays/s, "
        f"p50 {out['latency_ms_p50']} ms, p99 {out['latency_ms_p99']} ms; sequential {seq_s:.2f} s; "
        f"by size {out['by_size']}")
    log(f"serving: sustained {out['sustained']}")
    log(f"serving: previews {out['previews']}")
    for r in rasts:
        r.release()
    return out


def batch_leg(args):
    """Batched rendering: 64 requests over 8 seeded Cornell-style scenes with
    10k-500k added triangles, resolutions {128..1024}^2, spp {1, 4, 16, 64},
    depth 6.  Timed three ways on one stream: every request rendered alone
    with rise_hip_render, back to back (the single-scene path); the whole
    set as one batch launch; and that launch replayed from a captured graph.
    Every batched film is checked bit-for-bit against the film of the same
    request rendered alone."""
    import numpy as np
    import torch

    from rise_amd import pt_scene
    from rise_amd.render import PathTracingSpectralRasterizer, RenderBatch

    t0 = time.perf_counter()
    scenes, reqs = pt_scene.batch_workload()
    rasts = [PathTracingSpectralRasterizer(s) for s in scenes]
    setup_s = time.perf_counter() - t0
    n_req = len(reqs)
    paths = sum(r.width * r.height * r.spp for _, r in reqs)
    log(f"batch: {n_req} requests, {paths / 1e6:.1f} M paths, scenes {[s.n_tris for s in scenes]} "
        f"(setup {setup_s:.1f} s)")

    items = [(rasts[k], r) for k, r in reqs]
    tuning = None
    tune_log = None
    if not args.no_autotune:
        from rise_amd import autotune

        t1 = time.perf_counter()
        tuning, results = autotune.tune(items, proxy_spp=1, budget_s=90.0)
        tune_log = {"proxy": "the same 64 requests at 1 spp", "seconds": round(time.perf_counter() - t1, 1),
                    "candidates": [{**tn, "ms": round(ms, 3)} for tn, ms in results]}
        log(f"autotune: best {tuning} ({tune_log['seconds']} s)")
    batch = RenderBatch(items, tuning=tuning)
    batch.launch()
    st = batch.stats()
    rays = st["extension_rays"] + st["shadow_rays"]
    steps = max(1, args.batch_steps)

    # Single-scene path: one rise_hip_render per request, back to back.
    alone = []
    for (k, r), f in zip(reqs, batch.films):
        rasts[k].rasterize_scene(r, film=torch.empty_like(f), stats=False)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for (k, r), f in zip(reqs, batch.films):
        film, _, _ = rasts[k].rasterize_scene(r, film=torch.empty_like(f), stats=False)
        alone.append(film)
    torch.cuda.synchronize()
    seq_s = time.perf_counter() - w0
    identical = all(torch.equal(a, f) for a, f in zip(alone, batch.films))
    del alone

    # One batch launch per step.
    lat = []
    for _ in range(steps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        batch.launch()
        b.record()
        torch.cuda.synchronize()
        lat.append(a.elapsed_time(b))
    batch_ms = float(np.mean(lat))

    # The same launch captured in a graph.
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        batch.launch(side)
    torch.cuda.synchronize()
    glat = []
    for _ in range(steps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        g.replay()
        b.record()
        torch.cuda.synchronize()
        glat.append(a.elapsed_time(b))
    graph_ms = float(np.mean(glat))
    del g
    best_ms = min(batch_ms, graph_ms)
    out = {
        "workload": "64 render requests over 8 seeded Cornell-style scenes + 10k-500k random tris, "
        "res {128,256,512,1024}^2, spp {1,4,16,64}, depth 6",
        "requests": n_req,
        "paths": int(paths),
        "rays": {"extension": st["extension_rays"], "shadow": st["shadow_rays"]},
        "sequential_single_scene": {"s": round(seq_s, 4), "requests_per_s": round(n_req / seq_s, 2),
                                    "mrays_per_s": round(rays / seq_s / 1e6, 1)},
        "batched": {"ms": round(batch_ms, 3), "requests_per_s": round(n_req / (batch_ms / 1e3), 2),
                    "mrays_per_s": round(rays / (batch_ms / 1e3) / 1e6, 1),
                    "paths_per_s": round(paths / (batch_ms / 1e3), 1)},
        "graph_replay": {"ms": round(graph_ms, 3), "requests_per_s": round(n_req / (graph_ms / 1e3), 2),
                         "latency_ms_p50": round(float(np.percentile(glat, 50)), 3),
                         "latency_ms_p99": round(float(np.percentile(glat, 99)), 3)},
        "speedup_vs_sequential": round(seq_s / (best_ms / 1e3), 3),
        "tuning": batch.tuning(),
        "autotune": tune_log,
        "films_identical_alone_vs_batched": bool(identical),
        "kernel_counters": load_kernel_counters("batch"),
    }
    log(f"batch: sequential {seq_s * 1e3:.1f} ms, batched {batch_ms:.1f} ms, graph {graph_ms:.1f} ms, "
        f"{out['batched']['mrays_per_s']} Mrays/s, identical films {identical}")
    batch.release()
    for r in rasts:
        r.release()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true", help="run only the 1 GiB copy kernel")
    ap.add_argument("--no-render", action="store_true", help="skip the Cornell-box render leg")
    ap.add_argument("--render-steps", type=int, default=5)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-rendering leg")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-autotune", action="store_true", help="batched leg with default launch shapes")
    ap.add_argument("--no-terrain", action="store_true", help="skip the terrain-and-spheres render leg")
    ap.add_argument("--terrain-steps", type=int, default=1)
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--serving-requests", type=int, default=64)
    ap.add_argument("--no-quality", action="store_true", help="skip the relative-RMSE measurements of the render legs")
    ap.add_argument("--quality-trials", type=int, default=16, help="independent renders per relative-RMSE estimate")
    ap.add_argument("--no-widened", action="store_true",
                    help="skip the mesh-file / RGB-table / scene-file / output-stage timings")
    args = ap.parse_args()

    if args.gpus != 1:
        print(
            "bench.py measures one GPU only: there is no multi-GPU path (a render would shard by image "
            "tiles or sample ranges, serving by whole requests).",
            file=sys.stderr,
        )
        return 2

    import numpy as np
    import torch

    if not torch.cuda.is_available():
        log("no GPU visible")
        return 1
    if args.calibrate_copy:
        copy_calibration(max(1, args.steps))
        return 0

    from rise_amd import raycast, scenes

    dev = torch.device("cuda:0")
    t0 = time.time()
    tris = scenes.soup_triangles()
    rays_np = scenes.cast_rays()
    n = rays_np.shape[1]
    log(f"workload generated: {tris.shape[0]} tris, {n} rays ({time.time() - t0:.1f}s)")

    om = raycast.ObjectManager()
    om.add_mesh(tris)
    om.prepare_for_rendering()
    st = om.stats()
    log(f"BVH: {st['n_nodes']} nodes, depth {st['depth']}, build {st['build_ms']:.0f} ms")

    rays = torch.from_numpy(rays_np).to(dev)
    pid = torch.empty(n, dtype=torch.int32, device=dev)
    tt = torch.empty(n, dtype=torch.float32, device=dev)
    counters = torch.zeros(2, dtype=torch.int64, device=dev)
    om.intersect_ray(rays, pid, tt, counters=counters)
    torch.cuda.synchronize()
    cnt = counters.cpu().numpy().astype(np.float64)
    nodes_per_ray = cnt[0] / n
    tris_per_ray = cnt[1] / n
    node_bytes = int(st["node_bytes"])  # node record stride fetched per visit
    b_ray = RAY_IO_BYTES + node_bytes * nodes_per_ray + TRI_BYTES * tris_per_ray
    log(f"per ray: {nodes_per_ray:.2f} nodes, {tris_per_ray:.2f} triangles -> {b_ray:.1f} algorithmic bytes")

    for _ in range(args.warmup):
        om.intersect_ray(rays, pid, tt)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        om.intersect_ray(rays, pid, tt)
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = start.elapsed_time(end)
    ms_per_step = ms / args.steps
    mrays = n * args.steps / (ms / 1e3) / 1e6
    achieved = b_ray * n / (ms_per_step / 1e3) / 1e9
    log(f"{args.steps} steps: {ms:.2f} ms (wall {wall * 1e3:.1f} ms), {mrays:.1f} Mrays/s, {achieved:.0f} GB/s algorithmic")
    gpu_ids = pid.cpu().numpy()
    hit_frac = float((gpu_ids >= 0).mean())

    # The same rays with the reference's Linux-build arithmetic (fused
    # multiply-adds), for the parity check against that build.
    om.set_arithmetic(raycast.ARITH_X86_FMA)
    om.intersect_ray(rays, pid, tt)
    torch.cuda.synchronize()
    s0 = torch.cuda.Event(enable_timing=True)
    s1 = torch.cuda.Event(enable_timing=True)
    s0.record()
    for _ in range(5):
        om.intersect_ray(rays, pid, tt)
    s1.record()
    torch.cuda.synchronize()
    fma_ms = s0.elapsed_time(s1) / 5
    gpu_ids_fma = pid.cpu().numpy()
    om.set_arithmetic(raycast.ARITH_IEEE)
    log(f"x86-FMA arithmetic: {fma_ms:.3f} ms per step; ids differ from the IEEE model on "
        f"{int((gpu_ids_fma != gpu_ids).sum())} rays")

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        import oracle

        if oracle.ref_available("fast"):
            th = cpu_threads()
            tb = time.time()
            ref = oracle.RefMesh(tris, "fast")
            build_s = time.time() - tb
            probe = np.arange(0, n, max(1, n // 20000))
            tp = time.perf_counter()
            ref.closest(rays_np[:, probe], threads=th)
            rate = len(probe) / max(1e-6, time.perf_counter() - tp)
            m = int(min(n, max(len(probe), rate * args.cpu_seconds)))
            idx = np.linspace(0, n - 1, m).astype(np.int64)
            sample = np.ascontiguousarray(rays_np[:, idx])
            tc = time.perf_counter()
            rp, _, _ = ref.closest(sample, threads=th)
            sec = time.perf_counter() - tc
            ref.close()
            cpu = {
                "value": round(m / sec / 1e6, 4),
                "unit": "Mrays/s",
                "cores": th,
                "kind": "reference",
                "sample": f"{m} of the {n} rays (evenly spaced), closest hit, RISE BVH4 mesh path "
                f"built with its Linux flags; {sec:.1f} s of work; BVH build {build_s:.1f} s not counted",
            }
            bad_fast = int((rp != gpu_ids_fma[idx]).sum())
            # The reference's source arithmetic (every operation rounded) on every ray.
            ref_i = oracle.RefMesh(tris, "ieee")
            rpi, _, _ = ref_i.closest(rays_np, threads=th)
            ref_i.close()
            bad_ieee = int((rpi != gpu_ids).sum())
            parity = {
                "against": "reference BVH4 mesh path (float pre-filter + double certifier) compiled from its sources",
                "rays_checked_ieee": int(n),
                "id_mismatches_ieee": bad_ieee,
                "rays_checked_x86_fma": m,
                "id_mismatches_x86_fma": bad_fast,
                "x86_fma_ms_per_step": round(fma_ms, 4),
            }
            log(f"CPU reference: {cpu['value']} Mrays/s on {th} threads; id mismatches {bad_ieee}/{n} "
                f"(IEEE build), {bad_fast}/{m} (Linux-flags build, FMA arithmetic)")
        else:
            log("oracle/_ref not built: no CPU baseline")

    render = None if args.no_render else render_leg(args, cpu_enabled=not args.no_cpu_baseline)
    batched = None if args.no_batch else batch_leg(args)
    terrain = None if args.no_terrain else terrain_leg(args)
    serving = None if args.no_serving else serving_leg(args)
    widened = None if args.no_widened else widened_leg(args, tris)

    traffic = load_traffic()
    line = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 traversal / f64 certify",
        "data": "synthetic",
        "config": {
            "workload": "ray-cast closest hit: 1,000,000-triangle seeded soup (centroids U[0,1]^3, edges N(0,0.003^2)), "
            "2,073,600 coherent pinhole rays (1920x1080, 40 deg) + 4,000,000 incoherent rays",
            "n_tris": int(tris.shape[0]),
            "rays_per_step": int(n),
            "hit_fraction": round(hit_frac, 4),
            "bvh": {"nodes": st["n_nodes"], "depth": st["depth"], "build_ms": round(st["build_ms"], 1)},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "closest_persistent",
            "bytes_per_ray": round(b_ray, 2),
            "node_record_bytes": node_bytes,
            "nodes_per_ray": round(nodes_per_ray, 3),
            "tris_per_ray": round(tris_per_ray, 3),
        },
        "cpu_baseline": cpu,
        "parity": parity,
        "render": render,
        "batch": batched,
        "terrain": terrain,
        "serving": serving,
        "widened": widened,
    }
    om.release()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
