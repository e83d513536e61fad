import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def erm():
    return importlib.import_module("2d_and_3d_elastic_reverse_time_migration_amd")


@pytest.fixture(scope="session")
def oracle2d_mod():
    import oracle2d
    return oracle2d


@pytest.fixture(scope="session")
def oracle():
    import oracle3d
    oracle3d.set_threads(min(8, os.cpu_count() or 1))
    return oracle3d
