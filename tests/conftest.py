"""Test configuration: import paths and the ``gpu`` marker.

The product package root (``radiative-transfer-amr_amd/``) plays the role of
the reference's ``src/`` directory (``tests/conftest.py:1-14`` of the
reference adds ``src`` to ``sys.path``); the repository root is added so the
oracle can be imported as the checker.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "radiative-transfer-amr_amd")
for p in (PKG, ROOT, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
