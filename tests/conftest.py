import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope="session")
def orc():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def gpu():
    """A library context on cuda:0.  No fallback: the HIP library must load."""
    from inmyroom_vulkan_amd import Collider
    c = Collider(leaf_max=4)
    yield c
    c.close()


@pytest.fixture(scope="session")
def gpu_l1():
    from inmyroom_vulkan_amd import Collider
    c = Collider(leaf_max=1)
    yield c
    c.close()


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def case(msg):
    """Name the case on stdout before a GPU step (flushed)."""
    print(f"[case] {msg}", flush=True)
