import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    # The checker and the library are built in-tree; build them here only when
    # absent (a fresh checkout on a CPU machine).
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        import oracle

        oracle.build()
    if not os.path.exists(os.path.join(ROOT, "rainmd_amd", "librainmd_amd.so")):
        from rainmd_amd import build

        build.build()


@pytest.fixture(scope="session")
def golden():
    import json

    def load(name):
        with open(os.path.join(ROOT, "tests", "golden", "openmm", name)) as f:
            return json.load(f)

    return load
