import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "multigrid-fvm_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def pytest_sessionstart(session):
    """Initialise torch's GPU runtime before any test loads the solver
    library: torch carries its own HIP runtime, and it finds no device when
    the library's runtime has initialised the GPU first in the process."""
    try:
        import torch
        torch.cuda.is_available()
    except Exception:
        pass


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
