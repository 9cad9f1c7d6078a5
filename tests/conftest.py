import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (select with -m gpu)")


@pytest.fixture(scope="session")
def ref():
    """The reference's own hot-path sources compiled by oracle/Makefile."""
    from oracle import oracle as O
    if not O.reference_available():
        pytest.skip("oracle/_ref not built")
    return O.ReferenceLibrary()
