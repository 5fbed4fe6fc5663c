"""The RISE-side binding (rise_amd/host/rise_binding.hpp) compiled against
RISE's own headers: an Implementation::ObjectManager that uploads its objects
(meshes in object space with their transforms) and an IRasterizer over the
spectral path tracer.  The translation unit below creates both, so any pure
virtual of IObjectManager / IRasterizer / IReference left unimplemented, or
any signature that does not match RISE's, fails the compile.  Syntax and types
only: linking needs RISE's library.  Skipped where the RISE tree is absent."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
RISE_LIB = Path("/root/reference/src/Library")

TU = r'''
#include "rise_binding.hpp"

int main()
{
    using namespace RISE;
    using namespace RISE::Implementation;
    HipAcceleratedObjectManager* om = new HipAcceleratedObjectManager(true, false, 4, 32);
    const IObjectManager* iface = om;
    (void)iface;
    HipPathTracingSpectralRasterizer* r = new HipPathTracingSpectralRasterizer("scene.RISEscene");
    IRasterizer* ir = r;
    ir->FreeRasterizerOutputs();
    r->release();
    om->release();
    return 0;
}
'''


@pytest.mark.skipif(not RISE_LIB.exists() or shutil.which("g++") is None, reason="RISE headers or g++ not present")
def test_binding_compiles_against_rise_headers():
    cmd = ["g++", "-std=gnu++17", "-fsyntax-only", "-x", "c++", "-Wall", "-DCOLORS_RGB", "-DMERSENNE53",
           "-DRISE_ENABLE_MAILBOXING", "-D__HIP_PLATFORM_AMD__", f"-I{RISE_LIB}", f"-I{ROOT / 'include'}",
           f"-I{ROOT / 'rise_amd' / 'host'}", "-I/opt/rocm/include", "-"]
    r = subprocess.run(cmd, input=TU, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
