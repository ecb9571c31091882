"""pytest configuration: the `gpu` marker and import paths.

Tests marked `gpu` need an MI355X (run with `-m gpu`); everything else runs on
the CPU.  The oracle (oracle/) and the product package (eto-heom_amd/) are put
on sys.path here so tests can import them by module name.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "oracle", "eto-heom_amd", "tests"):
    p = os.path.join(ROOT, sub) if sub else ROOT
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
