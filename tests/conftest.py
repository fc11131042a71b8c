import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("ollinaxis-bib_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def oaxis_mod():
    import oaxis
    oaxis.lib()
    return oaxis


@pytest.fixture(scope="session")
def oracle_mod():
    import pyoracle
    pyoracle.lib()
    # the library's default stopping rule for the initial data's elliptic
    # solve; tests of the reference's rule alone turn it off
    pyoracle.ell_stationary(True)
    return pyoracle
