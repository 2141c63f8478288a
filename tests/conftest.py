import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle_lib
    oracle_lib.load()
    return oracle_lib


@pytest.fixture(scope="session")
def ylib():
    from yapt_amd import _lib
    return _lib


@pytest.fixture(scope="session")
def cornell():
    from yapt_amd import scenes
    return scenes.cornell_box()


@pytest.fixture(scope="session")
def shapes_small():
    """~25k-triangle random-shapes scene (60 shapes) for brute-force checks."""
    from yapt_amd import scenes
    return scenes.random_shapes(n_shapes=60, seed=7, cube=8.0, n_lights=2, light_emission=20.0)
