"""Test configuration.

Markers:
  gpu  - needs a visible HIP device; run on the GPU box with ``-m gpu``.
Everything else runs on the CPU (``-m "not gpu"``) against the CPU oracle,
the reference CPU build in oracle/_ref (when present) and the native
library's host-side entry points.
"""
import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the GPU tests of this process."""
    from cudaedgeelimination_amd import backend

    context = backend.Context(0)
    yield context
    context.close()


def require_ref():
    import _oracle

    if not _oracle.ref_available():
        pytest.skip("reference CPU build oracle/_ref is not present")
