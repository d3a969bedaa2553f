import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); the parity tests proper")


@pytest.fixture(scope="session")
def golden():
    import json
    data = json.loads((GOLDEN / "golden.json").read_text())
    return {c["name"]: c for c in data["cases"]}
