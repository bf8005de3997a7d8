"""Shared pytest setup.

* ``gpu`` marks tests that need an MI355X (run with ``-m gpu``).
* The drop-in package lives in ``famrdp-jax_amd/`` (importable as
  ``famrdp_jax``); the repository root holds the CPU oracle (``oracle``).
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "famrdp-jax_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X) and the built HIP library")
