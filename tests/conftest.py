"""Test configuration: the `gpu` marker and shared helpers.

CPU tier (`-m "not gpu"`): the oracle against closed forms and the compiled
reference host sources, the host surface, the C ABI exports.  GPU tier
(`-m gpu`): the HIP path through the C ABI against the oracle.
"""
from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and runs the HIP path")
    config.addinivalue_line("markers", "slow: long-running case")


def _gpu_count() -> int:
    try:
        import water_simulation_amd as ws
        return ws.device_count()
    except Exception:
        return 0


_HAS_GPU = None


def has_gpu() -> bool:
    global _HAS_GPU
    if _HAS_GPU is None:
        _HAS_GPU = _gpu_count() > 0
    return _HAS_GPU


def pytest_collection_modifyitems(config, items):
    for item in items:
        if "gpu" in item.keywords and not has_gpu():
            item.add_marker(pytest.mark.skip(reason="no HIP device"))
