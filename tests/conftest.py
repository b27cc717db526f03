"""pytest configuration: the `gpu` marker and import paths.

`-m "not gpu"` runs the oracle checks, host logic and the ABI-export check
on any CPU; `-m gpu` runs the parity tests through the C ABI on a GPU.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP library through the C ABI)")
