"""Test configuration: the `gpu` marker and import paths.

`-m "not gpu"` tests run on the CPU (oracle pins, host logic, ABI exports);
`-m gpu` tests need an MI355X and call the HIP library through its C ABI.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "oracle", "afid-murphfi_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
