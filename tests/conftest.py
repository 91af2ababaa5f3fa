import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES_DIR = os.path.join(GOLDEN, "cases")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than ~20 s")


def golden_cases(with_states=None):
    names = sorted(os.listdir(CASES_DIR))
    if with_states is None:
        return names
    out = []
    for n in names:
        has = n != "config1_honeycomb_20k"
        if has == with_states:
            out.append(n)
    return out


@pytest.fixture(scope="session")
def gpu_ctx():
    from gpuqt_amd import native
    print(f"[gpu] opening device 0", flush=True)
    ctx = native.Context(0)
    yield ctx
    ctx.close()
