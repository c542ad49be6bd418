import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
if GOLDEN not in sys.path:
    sys.path.insert(0, GOLDEN)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def ctx():
    from klampt_amd import engine
    c = engine.Context(0)
    yield c


class engine_params:
    """Context manager: set some engine parameters of a context for a block
    and restore the previous values afterwards."""

    def __init__(self, ctx, **changes):
        self.ctx, self.changes = ctx, changes

    def __enter__(self):
        self.old = self.ctx.set_params(**self.changes)
        return self

    def __exit__(self, *exc):
        self.ctx.set_params(**self.old)
        return False
