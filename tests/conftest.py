import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not oracle.LIB_PATH.exists():
        oracle.build()
    return oracle


def load_golden(name):
    d = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    return {k: d[k] for k in d.files}


def golden_names():
    import json
    return sorted(json.loads((GOLDEN / "manifest.json").read_text()))
