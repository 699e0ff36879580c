import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine through the C ABI)")
    config.addinivalue_line("markers", "slow: larger CPU oracle cases")


@pytest.fixture(scope="session")
def engine():
    # torch's HIP runtime must come up before the library's: a process whose first
    # HIP initialisation was the library's sees no GPU through torch afterwards
    import torch  # noqa: F401
    from soap_gpu_amd.engine import Engine
    eng = Engine(0)
    yield eng
    eng.close()


@pytest.fixture(scope="session")
def tables():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "input_spec_tables.json")) as f:
        return json.load(f)
