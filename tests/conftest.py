import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (built from oracle/ with its Makefile if needed)."""
    from oracle import oracle as o

    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        o.build()
    return o


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a real gfx950 device; fails loudly (never skips) when
    the library or the device is missing."""
    from paintop_amd import hip

    info = hip.probe()
    assert info["arch"].startswith("gfx950"), info
    print(f"[gpu] {info['name']} {info['arch']} CUs={info['compute_units']}", flush=True)
    return hip
