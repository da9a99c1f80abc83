"""Test configuration: puts the repository root (for ``oracle``) and the
package directory ``tvb-root_amd`` (for ``tvb_hip``) on sys.path and
registers the ``gpu`` marker for tests that need an MI355X."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tvb-root_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running")
