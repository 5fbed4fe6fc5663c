import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_built():
    from rise_amd import _lib

    if not _lib.LIB_PATH.exists():
        from rise_amd import build

        build.build()
    import oracle

    if not oracle.ORACLE_SO.exists():
        oracle.build(with_ref=oracle.REFERENCE_TREE.exists())


_ensure_built()


@pytest.fixture(scope="session")
def cuda():
    """The GPU device; GPU tests fail (not skip) without one."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return torch.device("cuda:0")


def log_case(msg: str) -> None:
    """Names the case on stdout before a GPU step (survives a crashed run)."""
    print(f"[case] {msg}", flush=True)
