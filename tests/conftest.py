import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

DECKS = os.path.join(ROOT, "cudaprotoplanetarydisks_amd", "decks")
REFERENCE_IN = "/root/reference/in"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfargo_hip.so on cuda:0)")
    config.addinivalue_line("markers", "slow: long-running case")


from cudaprotoplanetarydisks_amd.workloads import BASE_DECK, write_deck, write_planets  # noqa: E402,F401


@pytest.fixture
def deck(tmp_path):
    def make(name="d.par", **kw):
        kw.setdefault("OutputDir", str(tmp_path) + "/")
        return write_deck(tmp_path / name, **kw)
    return make
