"""The C++ host adapters (rise_amd/host/rise_hip_adapters.hpp) drive the GPU
path end to end from a C++ program: a ray cast with known answers, a tiny
spectral render, and a scene file loaded by LoadAsciiScene and rendered."""
import subprocess
from pathlib import Path

import pytest

from rise_amd import build
from tests.conftest import log_case

pytestmark = pytest.mark.gpu


SCENE = Path(__file__).resolve().parent / "golden" / "scenes" / "cornell.RISEscene"


@pytest.mark.parametrize("mode", ["raycast", "render", "scene"])
def test_cpp_adapter_demo(cuda, mode):
    log_case(f"adapter_demo {mode}")
    args = [str(build.DEMO), mode] + ([str(SCENE)] if mode == "scene" else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert r.stdout.startswith("ok ")
