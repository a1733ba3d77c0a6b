import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


def load_golden(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        g = json.load(f)
    g["deck_path"] = os.path.join(GOLDEN, g["deck"])
    return g


# fixture files holding many generated decks' tallies (not one deck each)
WORKLOAD_SETS = ("mixed_batch", "serve_stream")


def golden_names():
    """Single-deck fixtures (tests/golden/<name>.json with decks/<name>.inp)."""
    return sorted(f[:-5] for f in os.listdir(GOLDEN) if f.endswith(".json") and f[:-5] not in WORKLOAD_SETS)


def load_workload_set(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def gpu_available():
    import quicksilver_amd as q
    q.lib()  # raises if the library is missing
    return True
