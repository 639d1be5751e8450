"""Shared pytest setup: the `gpu` marker and the library/oracle handles."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _ensure_built():
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    if not os.path.exists(os.path.join(ROOT, "albba_amd", "libalbba.so")):
        subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "albba_amd", "csrc")],
                       check=True)


_ensure_built()


@pytest.fixture(scope="session")
def ctx():
    import albba_amd as A

    c = A.Context(0)
    yield c
    c.close()
