import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsong.so on the device)")


@pytest.fixture(scope="session")
def song():
    return importlib.import_module("2020_song_ex_amd")


@pytest.fixture(scope="session")
def datasets():
    return importlib.import_module("2020_song_ex_amd.datasets")


@pytest.fixture(scope="session")
def oracle():
    return importlib.import_module("oracle")
