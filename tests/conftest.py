"""Shared fixtures.  GPU tests carry @pytest.mark.gpu; everything else runs on CPU."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def ref():
    """The reference LinA compiled from its own sources (oracle/_ref)."""
    from oracle.bind import Ref, REF_SO
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref/libreflina.so not built (needs /root/reference at build time)")
    return Ref()


@pytest.fixture(scope="session")
def gpu():
    from fem_amd import lina
    if lina.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need the MI355X")
    return lina
