"""The C ABI library loads on a machine without a GPU, exports every function
declared in include/*.h, and rejects bad arguments with status codes (no
compute calls here)."""
import ctypes

from rise_amd import _lib


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _lib.declared_symbols()
    assert "rise_hip_intersect_closest" in declared
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, f"declared but not exported: {missing}"


def test_version_and_error_strings():
    L = _lib.lib()
    assert L.rise_hip_version().decode().count(".") == 2
    assert isinstance(L.rise_hip_last_error(), bytes)


def test_null_arguments_rejected():
    L = _lib.lib()
    assert L.rise_hip_intersect_closest(None, None, 10, None, None, None) == 1
    assert b"scene" in L.rise_hip_last_error()
    assert L.rise_hip_intersect_any(None, None, None, 10, None, None) == 1
    assert L.rise_hip_scene_upload(None, 0, None) == 1
    assert L.rise_hip_scene_get_stats(None, None) == 1
    assert L.rise_hip_scene_release(None) == 0
    out = ctypes.c_void_p()
    desc = _lib.MeshDesc(None, 5, None)
    assert L.rise_hip_scene_upload(ctypes.byref(desc), 1, ctypes.byref(out)) == 1
    desc = _lib.MeshDesc(None, -1, None)
    assert L.rise_hip_scene_upload(ctypes.byref(desc), 1, ctypes.byref(out)) == 1


def test_render_and_batch_arguments_rejected():
    from rise_amd import render

    L = render._lib_render()
    out = ctypes.c_void_p()
    assert L.rise_hip_batch_create(None, 0, ctypes.byref(out)) == 1
    assert b"empty batch" in L.rise_hip_last_error()
    assert L.rise_hip_batch_create(None, 3, None) == 1
    # A request without a scene is rejected during planning, before any device call.
    reqs = (render.RenderRequest * 1)()
    reqs[0].desc.width = reqs[0].desc.height = reqs[0].desc.spp = reqs[0].desc.spp_total = 4
    assert L.rise_hip_batch_create(reqs, 1, ctypes.byref(out)) == 1
    assert b"request 0" in L.rise_hip_last_error() and not out.value
    assert L.rise_hip_batch_launch(None, None) == 1
    assert L.rise_hip_batch_get_stats(None, None, None) == 1
    assert L.rise_hip_batch_release(None) == 0
    assert L.rise_hip_render(None, None, None, None, None, None) == 1
    assert L.rise_hip_pt_scene_release(None) == 0
    # Server: bad configurations and NULL arguments are rejected before any device call.
    cfg = render.ServerConfig(-1, 0, 0, render.BatchTuning())
    assert L.rise_hip_server_create(ctypes.byref(cfg), None, ctypes.byref(out)) == 1
    cfg = render.ServerConfig(0, 0, 0, render.BatchTuning(0, 96, 0, 0))
    assert L.rise_hip_server_create(ctypes.byref(cfg), None, ctypes.byref(out)) == 1
    assert b"tuning" in L.rise_hip_last_error()
    assert L.rise_hip_server_create(None, None, None) == 1
    assert L.rise_hip_server_submit(None, None, None) == 1
    assert L.rise_hip_server_poll(None, None, None, 0, None) == 1
    assert L.rise_hip_server_release(None) == 0


def test_cpp_adapter_demo_links():
    # The C++ host adapters (rise_amd/host) build and load the library; with no
    # arguments the demo prints its usage without touching the GPU.
    import subprocess

    from rise_amd import build

    assert build.DEMO.exists()
    r = subprocess.run([str(build.DEMO)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 64 and "usage" in r.stderr
