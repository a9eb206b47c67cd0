import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU oracle case")


@pytest.fixture(scope="session")
def transfers():
    """Device context on GPU 0 (fails loudly when the HIP library or GPU is missing)."""
    from tessera_amd.mpm import HipMpmTransfers

    return HipMpmTransfers(0)


@pytest.fixture(autouse=True)
def _fresh_tune_cache(request):
    """GPU tests start with an empty autotuner cache: a shape tuned by an
    earlier test must not replace another test's default launch shape."""
    if request.node.get_closest_marker("gpu") is not None:
        from tessera_amd.mpm import tune_cache_clear

        tune_cache_clear()
    yield
