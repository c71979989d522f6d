import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def gpu_case(msg: str) -> None:
    """Name the case on stdout before it touches the GPU (flushed)."""
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="session")
def ctx():
    from llama_amd.graph import GpuContext
    c = GpuContext(0)
    yield c
    c.close()
