import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cloth-cuda_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (ROOT, PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)
os.environ.setdefault("CLOTH_QUIET", "1")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); deselect with -m 'not gpu'")


def _built():
    libs = [os.path.join(PKG, "libclothhip.so"), os.path.join(PKG, "libcloth_core.so"),
            os.path.join(ORACLE, "_build", "liboracle.so")]
    return all(os.path.exists(p) for p in libs)


def pytest_sessionstart(session):
    # tests load what is built; build only when nothing is
    if not _built():
        import __graft_entry__
        __graft_entry__.build()
