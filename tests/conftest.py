import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP modules")


@pytest.fixture(scope="session")
def oracle():
    from scene_runner import ORACLE_SO, SceneModule
    return SceneModule(ORACLE_SO)


@pytest.fixture(scope="session")
def gpu():
    from scene_runner import GPU_SO, SceneModule
    if not os.path.exists(GPU_SO):
        pytest.fail(f"GPU scene module not built: {GPU_SO} (run __graft_entry__.build())")
    return SceneModule(GPU_SO)
