"""Shared pytest configuration: the ``gpu`` marker and the repository root on ``sys.path``."""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built HIP library")


def gpu_device():
    """The GPU device, or skip when there is none (GPU tests run only on the MI355X box)."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU device")
    return torch.device("cuda:0")
