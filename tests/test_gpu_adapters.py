"""The C++ host adapters (rise_amd/host/rise_hip_adapters.hpp) drive the GPU
path end to end from a C++ program: a ray cast with known answers and a tiny
spectral render."""
import subprocess

import pytest

from rise_amd import build
from tests.conftest import log_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["raycast", "render"])
def test_cpp_adapter_demo(cuda, mode):
    log_case(f"adapter_demo {mode}")
    r = subprocess.run([str(build.DEMO), mode], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert r.stdout.startswith("ok ")
