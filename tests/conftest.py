"""Shared pytest setup: the `gpu` marker and import paths.

The host package lives in ``making-it-rain-webgpu_amd/mdcore`` and the CPU
oracle wrapper in ``oracle/oracle.py``; both directories go on ``sys.path``.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("making-it-rain-webgpu_amd", "oracle"):
    path = os.path.join(ROOT, sub)
    if path not in sys.path:
        sys.path.insert(0, path)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmdcore.so on the GPU)")
