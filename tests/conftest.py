import importlib
import os
import shutil
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, GOLDEN):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU run, excluded from the default suite")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("AMCFD_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="long CPU run; set AMCFD_SLOW=1 to include")
    for item in items:
        if "slow" in item.keywords:
            item.add_marker(skip)


def amcfd_module():
    return importlib.import_module("am-cfd_amd")


@pytest.fixture(scope="session")
def amcfd():
    return amcfd_module()


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.load()
    return oracle


@pytest.fixture
def case_dir(tmp_path):
    """Write a named deck (tests/golden/decks.py) into a temp dir; returns a
    function case -> deck path."""
    import decks

    def make(case, **overrides):
        return decks.write_case(case, str(tmp_path / case), **overrides)
    return make


def golden_path(name):
    return os.path.join(GOLDEN, name)
